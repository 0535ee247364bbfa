/*
 * rivet_gen.h — seeded synthetic work lists for the reconstruction path.
 *
 * The reference has no stored vectors for its kernels: its tests drive
 * every kernel with a seeded LCG (src/dsp/hevc.rs:1444-1449,
 * src/dsp/u16_sweep.rs:39-48). This generator uses the same LCG to build
 * whole pictures of the shapes the benchmark configurations name: a CU
 * quadtree with symmetric / AMP partitions, quarter-sample vectors, a TU
 * quadtree with sparse Laplacian coefficients (scaled values, or levels
 * with qP and optional scaling lists), per-CU QP, per-CTB SAO
 * (HEVC); macroblocks with 16x16..4x4 partitions, 4x4 / 8x8 transforms and
 * intra macroblocks (H.264). The generated descriptions are what a parser
 * would hand the device: they feed the product path and the oracle alike.
 * No libm: Laplacian magnitudes use fixed-point thresholds.
 */
#ifndef RIVET_GEN_H
#define RIVET_GEN_H

#include "rivet_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rv_gen_hevc_cfg {
    int32_t width, height, bit_depth;
    int32_t b_picture;      /* 0: P (list 0 only); 1: B                    */
    int32_t bi_percent;     /* B: % of PUs bi-predicted (never 8x4 / 4x8)  */
    int32_t explicit_wp;    /* 1: explicit WP, denominators 6, weights
                               32..96, offsets +-32 (syntax units)          */
    int32_t intra_percent;  /* % of 8x8 CUs with 4x4 luma TUs made intra   */
    int32_t nz_permille;    /* per-coefficient nonzero probability x 1000  */
    int32_t lap_b;          /* Laplacian scale (8, 12, 16)                 */
    int32_t qp_min, qp_max;
    int32_t mv_range;       /* |mv| <= mv_range quarter samples            */
    int32_t deblock, sao;
    int32_t poc;
    int32_t ref_poc[2];     /* POC of the list-0 / list-1 reference        */
    int32_t ref_slot[2];    /* DPB slot of the list-0 / list-1 reference   */
    int32_t levels;         /* 1: coded TUs carry levels (RV_TU_LEVELS) at
                               the CU's qP (chroma through Table 8-10)     */
    int32_t scaling;        /* 1: seeded scaling lists (factors 4..64)     */
    int32_t intra_device;   /* 1: intra_percent of CUs of every size intra,
                               predicted on the device (RV_INTRA_DEVICE)
                               with seeded modes; 8x8 CUs split into 4x4
                               luma TUs are PART_NxN                       */
    int32_t strong_intra_smoothing, constrained_intra_pred;  /* copied */
    /* Slices and tiles: tile_cols x tile_rows uniformly spaced tiles (0 or
     * 1: one), n_slices slices of about equal CTB counts in tile-scan
     * order (0 or 1: one slice, no per-CTB tables). lf_across_slices /
     * lf_across_tiles: the flags (-1: each slice's flag seeded). slice_vary:
     * 1 seeds each slice's deblocking_disabled (1 in 5) and offsets
     * (-6..6, even). exempt_percent: % of CUs with loop-filter exemptions
     * (RV_CU_LF_EXEMPT, with RV_CU_BYPASS half the time, or
     * RV_CU_DBK_EXEMPT). None of them changes the other seeded streams. */
    int32_t tile_cols, tile_rows, n_slices;
    int32_t lf_across_slices, lf_across_tiles, slice_vary;
    int32_t exempt_percent;
} rv_gen_hevc_cfg;

typedef struct rv_gen_h264_cfg {
    int32_t width, height;  /* coded, multiples of 16                     */
    int32_t b_picture;
    int32_t bi_percent;     /* B: % of inter partitions bi-predicted       */
    int32_t weighting;      /* RV_H264_DEFAULT / RV_H264_IMPLICIT          */
    int32_t intra_percent;  /* % of macroblocks intra                      */
    int32_t nz_permille;
    int32_t lap_b;
    int32_t qp_min, qp_max;
    int32_t mv_range;
    int32_t deblock;
    int32_t poc;
    int32_t ref_poc[2];
    int32_t ref_slot[2];
    int32_t levels;         /* 1: blocks carry levels (RV_BLK_LEVELS): chroma
                               DC groups, and intra macroblocks without 8x8
                               transforms Intra_16x16 (luma DC group) half
                               the time                                     */
    int32_t scaling;        /* 1: seeded scaling lists (factors 4..64)     */
    int32_t intra_device;   /* 1: the intra macroblocks are RV_INTRA_DEVICE
                               with a seeded kind (I4X4 / I8X8 / I16X16)
                               and seeded modes among those their
                               neighbours allow                            */
    int32_t constrained_intra_pred;  /* copied */
} rv_gen_h264_cfg;

typedef struct rv_gen rv_gen;

int rv_gen_hevc(const rv_gen_hevc_cfg* cfg, uint64_t seed, rv_gen** out);
int rv_gen_h264(const rv_gen_h264_cfg* cfg, uint64_t seed, rv_gen** out);
/* The generated description (owned by the generator object). */
const rv_hevc_pic* rv_gen_hevc_pic(const rv_gen* g);
const rv_h264_pic* rv_gen_h264_pic(const rv_gen* g);
void rv_gen_free(rv_gen* g);
/* Uniform random planes, tightly packed, 4:2:0 (1 byte per sample at 8
 * bits, uint16 above). */
int rv_gen_planes(int width, int height, int bit_depth, uint64_t seed, void* y, void* cb, void* cr);
/* Derived sub-seed for (seed, a, b, c): stream / picture / purpose. */
uint64_t rv_gen_subseed(uint64_t seed, uint64_t a, uint64_t b, uint64_t c);

#ifdef __cplusplus
}
#endif
#endif
