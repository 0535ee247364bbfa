/*
 * rivet_hip.h — C ABI of the MI355X (gfx950) picture-reconstruction tier.
 *
 * Two levels, both plain C (no C++ or torch types cross this boundary):
 *
 *  (1) Work-list formats and the batched picture API. A host decoder (the
 *      parser, which stays on the CPU) describes one inter picture as
 *      host-packed lists — coding units, prediction units, transform units
 *      with their coefficients (scaled, or levels that the device scales),
 *      SAO parameters — and the
 *      library reconstructs it on the device: motion compensation, inverse
 *      transform + residual add, deblocking, SAO. Reference pictures stay
 *      resident in device memory (a per-session decoded-picture buffer).
 *
 *  (2) The per-entry parity API (rv_hevc_u8_*, rv_hevc_u16_*, rv_h264_u8_*):
 *      one function per entry of the reference's kernel tables
 *      (HevcDsp<S>, src/dsp/hevc.rs:165-229; H264Dsp<S>, src/dsp/h264.rs:
 *      123-170), same argument meaning, host buffers copied to the device
 *      and back per call. It exists so a table-for-table drop-in and the
 *      kernel-level parity tests can call single kernels.
 *
 * Every entry returns int: 0 (RV_OK) or a negative RV_E_* code. Nothing
 * aborts or throws across the boundary; rv_last_error() gives a
 * thread-local message for the last failure on the calling thread. The
 * reference's kernels are infallible Rust functions that panic on
 * out-of-range slices; here the same preconditions are checked and refused
 * with RV_E_RANGE.
 *
 * Threading: every entry is re-entrant. Sessions are independent and may
 * be driven from different host threads; launches serialise on an internal
 * lock. Completion is observed with rv_sync or markers (HIP events).
 * Kernels run on one library stream; plane uploads and downloads use a
 * second stream ordered after the kernels they depend on.
 */
#ifndef RIVET_HIP_H
#define RIVET_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Errors                                                              */
/* ------------------------------------------------------------------ */
#define RV_OK 0
#define RV_E_INVALID (-1)  /* bad argument (null pointer, bad enum)     */
#define RV_E_NODEVICE (-2) /* no HIP device / rv_init not called        */
#define RV_E_HIP (-3)      /* a HIP runtime call failed (message kept)  */
#define RV_E_NOMEM (-4)    /* device or pinned allocation failed        */
#define RV_E_RANGE (-5)    /* size / depth / fraction outside the range */
#define RV_E_STATE (-6)    /* call out of order (e.g. job still running) */
#define RV_E_DEVICE (-7)   /* a kernel's bounded wait gave up since the last
                              check: outputs of the launches completed since
                              then are not valid (reported once, by the
                              synchronising call that sees it) */

/* Limits of the work-list formats. */
#define RV_MAX_REFS 4 /* reference indices per list (DPB of up to 4 refs) */

/* ------------------------------------------------------------------ */
/* Shared types                                                        */
/* ------------------------------------------------------------------ */

/* A motion vector in quarter luma samples (src/hevc/frame.rs:20-26,
 * src/h264/frame.rs:18-24). */
typedef struct rv_mv {
    int16_t x, y;
} rv_mv;

/* SAO parameters of one CTB for one component (SaoParams,
 * src/hevc/pic.rs:10-17). offsets are SaoOffsetVal[1..=4], already scaled
 * by log2OffsetScale. */
typedef struct rv_sao {
    uint8_t type_idx;      /* 0 off, 1 band, 2 edge                     */
    uint8_t band_or_class; /* sao_band_position (band) / SaoEoClass     */
    int16_t offsets[4];
} rv_sao;

/* ------------------------------------------------------------------ */
/* HEVC work lists (Main / Main10, 4:2:0)                               */
/* ------------------------------------------------------------------ */

/* One coding unit: what the deblocking side data needs (QpY, prediction
 * mode, coding-block edges; src/hevc/ctu.rs:751-800) plus, for an intra
 * CU, where its host-supplied prediction block lives or the modes the
 * device predicts it with. */
typedef struct rv_hevc_cu {
    uint16_t x, y;      /* luma top-left                               */
    uint8_t log2_size;  /* 3..6                                        */
    int8_t qp_y;        /* QpY of the CU                               */
    uint8_t intra;      /* 0 inter, RV_INTRA_HOST or RV_INTRA_DEVICE   */
    uint8_t part_mode;  /* PartMode: 0 PART_2Nx2N, 3 PART_NxN (device
                           intra reads it; otherwise informational)   */
    uint32_t pred_off;  /* RV_INTRA_HOST: sample offset into intra_pred:
                           Y (s*s) then Cb and Cr ((s/2)^2 each), raster */
    uint8_t luma_mode[4]; /* RV_INTRA_DEVICE: IntraPredModeY (0..34) of
                           the PU; PART_NxN: of the four PUs, z order  */
    uint8_t chroma_mode;  /* RV_INTRA_DEVICE: IntraPredModeC (0..34), the
                           derived mode (intra_chroma_pred_mode 4 and the
                           mode-34 substitution resolved)              */
    uint8_t exempt;     /* loop-filter exemptions of the CU's samples, as
                           PicInfo::filter_exempt (src/hevc/pic.rs:77-80,
                           set at src/hevc/ctu.rs:763-766): RV_CU_LF_EXEMPT
                           | RV_CU_BYPASS | RV_CU_DBK_EXEMPT, or 0      */
    uint8_t pad[2];
} rv_hevc_cu;

/* rv_hevc_cu.exempt: deblocking and SAO leave the samples alone (PCM with
 * pcm_loop_filter_disabled_flag, or cu_transquant_bypass_flag). */
#define RV_CU_LF_EXEMPT 1u
/* cu_transquant_bypass_flag (recorded with RV_CU_LF_EXEMPT). */
#define RV_CU_BYPASS 2u
/* Deblocking only leaves the samples alone (palette_mode_flag). */
#define RV_CU_DBK_EXEMPT 4u

/* Intra CU whose prediction the host supplies in intra_pred (copied in
 * stage 0). */
#define RV_INTRA_HOST 1
/* Intra CU predicted on the device (src/hevc/intra.rs, called per
 * transform block from src/hevc/ctu.rs:1727-1833): every TU of the CU is
 * listed, luma and chroma, coded or not, and the TUs of such CUs appear in
 * decoding order (CTBs in raster order, z-scan within). Each TU is
 * predicted from the reconstructed samples around it, then its residual
 * added, in dependency order after the transform stage. */
#define RV_INTRA_DEVICE 2

/* One prediction unit (the arguments of predict_block,
 * src/hevc/inter.rs:555-567). */
typedef struct rv_hevc_pu {
    uint16_t x, y;      /* luma top-left                               */
    uint8_t w, h;       /* luma size, 4..64                            */
    uint8_t dir;        /* bit 0: list 0 used, bit 1: list 1 used      */
    uint8_t pad0;
    int8_t ref_idx[2];  /* per list, -1 = unused                       */
    uint16_t pad1;
    rv_mv mv[2];        /* quarter-sample vectors per list             */
} rv_hevc_pu;

/* One nonzero transform coefficient: pos = y * n + x in its n x n TU. A
 * coded TU's coefficients are the run nz[coef_off .. coef_off + n_nz)
 * (positions distinct); every other coefficient of the TU is zero — the
 * dense n x n raster of IdctFn (src/dsp/hevc.rs:18-21) is these values
 * scattered into zeros. val is the scaled coefficient d[x][y], or with
 * RV_TU_LEVELS the level TransCoeffLevel that stage 1 scales. */
typedef struct rv_nz {
    uint16_t pos;
    int16_t val;
} rv_nz;

/* One transform unit. Luma TUs are listed whether coded or not (their
 * edges feed the deblocking filter, src/hevc/ctu.rs:1493-1510); chroma TUs
 * only when coded. */
typedef struct rv_hevc_tu {
    uint16_t x, y;      /* top-left in component samples              */
    uint8_t comp;       /* 0 Y, 1 Cb, 2 Cr                             */
    uint8_t log2;       /* 2..5                                        */
    uint8_t flags;      /* RV_TU_DST | RV_TU_CODED | RV_TU_LEVELS |
                           RV_TU_INTRA                                 */
    uint8_t max_x;      /* nonzero bounds, inclusive (IdctFn,          */
    uint8_t max_y;      /*   src/dsp/hevc.rs:18-21)                    */
    uint8_t qp;         /* RV_TU_LEVELS: qP of the scaling (Qp'Y, Qp'Cb
                           or Qp'Cr, QpBdOffset included: 0 ..
                           51 + 6 * (bit_depth - 8)); else ignored     */
    uint16_t n_nz;      /* coded: number of nonzero coefficients       */
    uint32_t coef_off;  /* coded: first entry in nz                    */
} rv_hevc_tu;

#define RV_TU_DST 1u   /* 4x4 DST (intra luma 4x4, src/hevc/ctu.rs:2041) */
#define RV_TU_CODED 2u /* cbf: coefficients present                      */
/* The TU's nz values are levels: stage 1 scales them (scale_coefficients,
 * src/hevc/residual.rs:1223-1283, as called at src/hevc/ctu.rs:2016-2031)
 * with qp and the picture's scaling lists before the inverse transform.
 * Transform-skip, RDPCM and transquant-bypass TUs have no place in these
 * lists. */
#define RV_TU_LEVELS 4u
/* RV_TU_LEVELS of an intra CU: scaling matrices 0..2 (matrixId = cIdx),
 * else 3..5 (3 + cIdx), src/hevc/ctu.rs:2021. */
#define RV_TU_INTRA 8u

/* Resolved scaling factors of a picture (the lists the reference's
 * ScalingSource::List carries, src/hevc/residual.rs:1209-1215, chosen at
 * src/hevc/ctu.rs:2016-2029): default, predicted and explicit lists already
 * resolved by the host. Raster order (y * size + x); the 16x16 and 32x32
 * factors are the 8x8 grid, each entry covering a 2x2 / 4x4 square,
 * except (0, 0), which is dc. */
typedef struct rv_hevc_scaling {
    uint8_t list4[6][16];     /* sizeId 0, matrixId 0..5               */
    uint8_t list8[3][6][64];  /* sizeId 1..3 (8x8, 16x16, 32x32)       */
    uint8_t dc[2][6];         /* sizeId 2..3: scaling_list_dc_coef     */
} rv_hevc_scaling;

/* Loop-filter parameters of one slice (SliceFilterParams,
 * src/hevc/pic.rs:19-35; the chroma QP offsets are the PPS's, in
 * rv_hevc_pic). */
typedef struct rv_hevc_slice {
    int32_t deblocking_disabled;        /* slice_deblocking_filter_disabled_flag */
    int32_t beta_offset, tc_offset;     /* slice_*_offset_div2 * 2                */
    int32_t loop_filter_across_slices;  /* slice_loop_filter_across_slices_enabled_flag */
} rv_hevc_slice;

/* One HEVC P/B picture to reconstruct. Planes are the session's DPB
 * slots; ref_slot names them per list and reference index. */
typedef struct rv_hevc_pic {
    int32_t width, height;  /* coded luma size, multiples of 8         */
    int32_t bit_depth;      /* 8..12                                   */
    int32_t log2_ctb;       /* 4..6                                    */
    int32_t poc;
    int32_t num_ref[2];
    int32_t ref_slot[2][RV_MAX_REFS];
    int32_t ref_poc[2][RV_MAX_REFS];
    int32_t ref_long_term[2][RV_MAX_REFS];
    /* Explicit weighted prediction (8.5.3.3.4.3; Weighting::Explicit,
     * src/hevc/inter.rs:71-84, derived as in src/hevc/ctu.rs:1254-1292). */
    int32_t weighted;                        /* 0 default, 1 explicit  */
    int32_t log2_wd[3];                      /* shift1 folded in        */
    int32_t wp_w[2][RV_MAX_REFS][3];
    int32_t wp_o[2][RV_MAX_REFS][3];         /* in sample units         */
    /* Loop filters (SliceFilterParams, src/hevc/pic.rs:19-35). */
    int32_t deblock;                          /* 1 = on                  */
    int32_t beta_offset, tc_offset;           /* *_div2 * 2              */
    int32_t cb_qp_offset, cr_qp_offset;       /* PPS offsets             */
    int32_t sao_luma, sao_chroma;             /* slice_sao_*_flag        */
    /* Work lists. */
    int32_t n_cus, n_pus, n_tus;
    const rv_hevc_cu* cus;
    const rv_hevc_pu* pus;
    const rv_hevc_tu* tus;
    const rv_nz* nz;             /* nonzero coefficients of coded TUs    */
    uint32_t n_nz;
    uint32_t n_intra_pred;
    const uint16_t* intra_pred;  /* samples, any depth, as uint16       */
    const rv_sao* sao;           /* 3 per CTB (Y, Cb, Cr), raster CTBs   */
    /* Scaling lists of RV_TU_LEVELS TUs; NULL = flat 16 (scaling lists
     * off, ScalingSource::Flat). */
    const rv_hevc_scaling* scaling;
    /* Intra prediction on the device (RV_INTRA_DEVICE CUs). */
    int32_t strong_intra_smoothing;  /* strong_intra_smoothing_enabled_flag */
    int32_t constrained_intra_pred;  /* constrained_intra_pred_flag         */
    /* Slices and tiles (PicInfo::ctb_slice / ctb_slice_addr / slices and
     * Geometry::ctb_tile / ctb_rs_to_ts, src/hevc/pic.rs:41-96). NULL
     * ctb_slice: the picture is one slice and one tile, filtered with
     * deblock, beta_offset and tc_offset above. Otherwise every CTB
     * (raster order) names its slice in slices[], its SliceAddrRs, its
     * tile and its tile-scan address; a slice's deblocking_disabled and
     * offsets apply to the edges whose q block lies in its CTBs, and
     * deblocking and SAO cross slice and tile boundaries as
     * src/hevc/deblock.rs:120-165 and src/hevc/sao.rs:193-238 allow
     * (deblock = 0 still turns deblocking off for the whole picture; a
     * slice's SAO flags are its CTBs' rv_sao.type_idx). */
    int32_t n_slices;
    const rv_hevc_slice* slices;
    const uint16_t* ctb_slice;
    const uint32_t* ctb_slice_addr;
    const uint16_t* ctb_tile;
    const uint32_t* ctb_ts;
    int32_t loop_filter_across_tiles;  /* loop_filter_across_tiles_enabled_flag */
} rv_hevc_pic;

/* ------------------------------------------------------------------ */
/* H.264 work lists (High, 8-bit, 4:2:0, frame pictures)                */
/* ------------------------------------------------------------------ */

/* One macroblock (raster order, every MB of the picture): the facts the
 * deblocking filter reads (MbInfo, src/h264/mb.rs:322-369) and, for
 * macroblocks predicted on the device, the prediction modes the parser
 * read (MbLayer's intra16_mode / intra_modes / chroma_mode,
 * src/h264/recon.rs:611-828). */
typedef struct rv_h264_mb {
    uint8_t intra;          /* 0 inter, RV_INTRA_HOST (prediction
                               supplied in intra_pred) or RV_INTRA_DEVICE */
    uint8_t transform_8x8;
    int8_t qp;              /* QP_Y                                   */
    int8_t qpc[2];          /* QP_C for Cb, Cr                        */
    uint8_t pad;
    uint16_t nz_mask;       /* 4x4 luma blocks with coefficients (raster;
                               all four of an 8x8 when any)             */
    uint32_t pred_off;      /* RV_INTRA_HOST: offset into intra_pred
                               (Y 256, Cb 64, Cr 64)                    */
    uint8_t intra_kind;     /* RV_INTRA_DEVICE: RV_H264_I4X4, RV_H264_I8X8
                               or RV_H264_I16X16                        */
    uint8_t chroma_mode;    /* RV_INTRA_DEVICE: intra_chroma_pred_mode
                               (0 DC, 1 horizontal, 2 vertical, 3 plane) */
    uint8_t pad2[2];
    uint8_t modes[16];      /* RV_INTRA_DEVICE: Intra4x4PredMode per 4x4
                               block in raster order (I4X4); the 8x8
                               blocks' modes at raster 0, 2, 8, 10 (I8X8);
                               Intra16x16PredMode in modes[0] (I16X16) */
} rv_h264_mb;

/* Kinds of RV_INTRA_DEVICE H.264 macroblocks. Their residual blocks are
 * listed like any other (4x4 for I4X4 and I16X16, the latter with has_dc;
 * 8x8 for I8X8) and added after the prediction of their block, in decoding
 * order within the macroblock; macroblocks are predicted in dependency
 * order after the transform stage. */
#define RV_H264_I4X4 0
#define RV_H264_I8X8 1
#define RV_H264_I16X16 2

/* One inter partition (the arguments of predict_partition,
 * src/h264/inter.rs:172-185). */
typedef struct rv_h264_part {
    uint16_t x, y;       /* luma top-left                              */
    uint8_t w, h;        /* 4, 8 or 16                                 */
    uint8_t dir;         /* bit 0: list 0, bit 1: list 1               */
    uint8_t pad0;
    int8_t ref_idx[2];
    uint16_t pad1;
    rv_mv mv[2];
} rv_h264_part;

/* One coded residual block (the arguments of Residual4Fn / Residual8Fn,
 * src/dsp/h264.rs:109-116). */
typedef struct rv_h264_blk {
    uint16_t x, y;      /* top-left in component samples              */
    uint8_t comp;       /* 0 Y, 1 Cb, 2 Cr                            */
    uint8_t size;       /* 4 or 8                                     */
    uint8_t has_dc;     /* 1: dc replaces position 0 (chroma DC, or
                           the luma DC of an Intra_16x16 macroblock)   */
    uint8_t flags;      /* RV_BLK_LEVELS                              */
    int32_t dc;         /* has_dc: the scaled DC; with RV_BLK_LEVELS
                           the offset in coefs of the DC levels of the
                           block's group (below)                      */
    uint32_t coef_off;  /* offset of size*size int32 in coefs         */
} rv_h264_blk;

/* The block's coefs are levels, raster order: stage 1 scales them as the
 * reference's parser does (dequant_level with the MbDequant tables,
 * src/h264/mb.rs:1363-1428; LevelScale from Dequant::new,
 * src/h264/transform.rs:24-56), with the macroblock's QP_Y (luma) or
 * QP_C (chroma), its intra / inter lists and the picture's scaling lists.
 * With has_dc, position 0 is not read: dc is the offset of the group's DC
 * levels, raster — 4 for a chroma component of the macroblock
 * (chroma_dc_transform_420, src/h264/transform.rs:216-228), 16 for the
 * luma of an Intra_16x16 macroblock (luma_dc_transform, :179-213), as
 * called at src/h264/recon.rs:700-716 and 1200-1226 — and the block takes
 * the transformed, scaled DC of its own position in the group. */
#define RV_BLK_LEVELS 1u

/* Scaling lists of a picture (ScalingLists, src/h264/sps.rs), resolved by
 * the host (fall-back rules applied), raster order. */
typedef struct rv_h264_scaling {
    uint8_t list4[6][16];   /* Y, Cb, Cr intra; Y, Cb, Cr inter         */
    uint8_t list8[2][64];   /* Y intra, Y inter                         */
} rv_h264_scaling;

#define RV_H264_DEFAULT 0
#define RV_H264_EXPLICIT 1
#define RV_H264_IMPLICIT 2

typedef struct rv_h264_pic {
    int32_t width, height;   /* coded luma size, multiples of 16        */
    int32_t bit_depth;       /* 8                                       */
    int32_t poc;
    int32_t num_ref[2];
    int32_t ref_slot[2][RV_MAX_REFS];
    int32_t ref_poc[2][RV_MAX_REFS];
    int32_t ref_long_term[2][RV_MAX_REFS];
    int32_t weighting;       /* RV_H264_DEFAULT / EXPLICIT / IMPLICIT     */
    int32_t log_wd[3];       /* explicit only                            */
    int32_t wp_w[2][RV_MAX_REFS][3];
    int32_t wp_o[2][RV_MAX_REFS][3];
    int32_t deblock;         /* 1 = on (disable_deblocking_filter_idc 0) */
    int32_t offset_a, offset_b;
    int32_t n_mbs, n_parts, n_blks;
    const rv_h264_mb* mbs;
    const rv_h264_part* parts;
    const rv_h264_blk* blks;
    const int32_t* coefs;
    uint32_t n_coefs;
    uint32_t n_intra_pred;
    const uint16_t* intra_pred;
    /* Scaling lists of RV_BLK_LEVELS blocks; NULL = flat 16
     * (Flat_4x4_16 / Flat_8x8_16). */
    const rv_h264_scaling* scaling;
    /* constrained_intra_pred_flag: RV_INTRA_DEVICE macroblocks predict
     * only from intra macroblocks. */
    int32_t constrained_intra_pred;
} rv_h264_pic;

/* ------------------------------------------------------------------ */
/* Device, sessions, jobs                                              */
/* ------------------------------------------------------------------ */

/* Select the device (mirrors Cpu::detect_honouring_env, src/dsp/mod.rs:
 * 200,304: chosen once). Must precede every device call. Returns
 * RV_E_NODEVICE without a gfx950 device, or when the environment sets
 * RIVET_HIP_DISABLE=1 (the analogue of H26X_NO_SIMD, src/dsp/mod.rs:
 * 295-350): the caller then keeps its own CPU tier. There is no CPU
 * fallback inside this library. */
int rv_init(int device_ordinal);
/* "hip-gfx950" once initialised (mirrors Cpu::rung, src/dsp/mod.rs:271). */
const char* rv_tier_name(void);
/* Thread-local message of the last failure on this thread. */
const char* rv_last_error(void);
/* Number of visible HIP devices (0 without a GPU; never fails). */
int rv_device_count(void);

#define RV_CODEC_HEVC 0
#define RV_CODEC_H264 1

typedef struct rv_session rv_session;

/* A decode session: one stream's DPB planes resident on the device plus
 * one reconstruction scratch picture. dpb_slots reference/output slots of
 * width x height (luma, 4:2:0) at bit_depth. */
int rv_session_create(int codec, int width, int height, int bit_depth,
                      int dpb_slots, rv_session** out);
/* Returns at once; the DPB's memory is reused once earlier work is done. */
int rv_session_destroy(rv_session* s);
/* Device bytes a session of this shape holds (DPB + scratch), for
 * admission by footprint. */
int64_t rv_session_footprint(int codec, int width, int height, int bit_depth,
                             int dpb_slots);
/* Host planes are tightly packed (row stride = plane width), 1 byte per
 * sample at 8 bits, 2 (little-endian uint16) above. Both calls return when
 * the copy is complete. Each first waits for the latest launch that
 * included a job of this session (an upload must not overwrite a slot
 * that is still being read; a download must see the picture written
 * there), and never for other sessions' work. Calls from several threads
 * may overlap. */
int rv_dpb_upload(rv_session* s, int slot, const void* y, const void* cb,
                  const void* cr);
int rv_plane_download(rv_session* s, int slot, void* y, void* cb, void* cr);

/* Decoded picture hash (H.265 D.3.20) of DPB slot `slot`, computed on the
 * device without a plane download. Replaces the hashing half of
 * hash::verify (src/hevc/hash.rs:109-158): per component, the plane's
 * samples in raster order as bytes, low byte first above 8 bits, then
 * `type` (the SEI's hash_type):
 *   RV_HASH_MD5       md5 (:191-250): out = 3 x 16 bytes (Y, Cb, Cr);
 *   RV_HASH_CRC       crc16 (:160-174): out = 3 x uint16;
 *   RV_HASH_CHECKSUM  checksum (:176-189): out = 3 x uint32.
 * Comparing with the SEI's values is the caller's (PartialEq on
 * PictureHash, :12-21). Waits like rv_plane_download for the slot's
 * writer. MD5 is serial within a component: about 106 ms for a 4K 8-bit
 * picture, for verification rather than the decode loop; CRC and checksum
 * take about 0.15 and 0.1 ms. RV_E_INVALID for an unknown type. */
#define RV_HASH_MD5 0
#define RV_HASH_CRC 1
#define RV_HASH_CHECKSUM 2
int rv_picture_hash(rv_session* s, int slot, int type, void* out);

/* rv_picture_hash that allocates nothing. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture elsewhere in the process, tight serving loops, harnesses
 * that forbid hipMalloc). Same arguments, results and ordering after the
 * slot's writer; the hash kernels' working memory is a fixed device array
 * of the library (about 51 KiB per device, held until the process ends)
 * instead of a buffer from the library's pool, and the result returns on
 * the kernel stream instead of through an event. Pictures of more than
 * RV_HASH_NOALLOC_MAX_ROWS rows (8K UHD has 4320) are refused with
 * RV_E_RANGE. Calls serialise on a lock of their own, since they share
 * the array; the array keeps the last call's contents. */
#define RV_HASH_NOALLOC_MAX_ROWS 4352
int rv_picture_hash_noalloc(rv_session* s, int slot, int type, void* out);

/* The MD5 of each of n pictures packed planar: its Y rows, then Cb, then
 * Cr, cropped to the output size widths[i] x heights[i] (chroma
 * ceil(w / 2) x ceil(h / 2)), samples above 8 bits as two bytes low byte
 * first — the digest of the per-frame line examples/h26xdec.rs:120-131
 * prints for Picture::into_packed (src/picture.rs:155-163). Picture i is
 * DPB slot slots[i] of sessions[i]; the hashes run after the latest launch
 * that covered each session, one workgroup per picture in one launch.
 * out receives 16 bytes per picture. MD5 is serial within a picture
 * (about 0.1 GB/s per picture), so many pictures per call is the way to
 * throughput. RV_E_RANGE for a slot outside its DPB or a size above the
 * session's. */
int rv_frame_md5(int n, rv_session* const* sessions, const int* slots, const int* widths, const int* heights,
                 uint8_t* out);

/* A prepared picture: the work lists copied into device memory (they are
 * not referenced after the call returns). The job writes its picture into
 * DPB slot out_slot of its session when launched; it may be launched many
 * times (it reads only its references). */
typedef struct rv_job rv_job;
int rv_hevc_prepare(rv_session* s, const rv_hevc_pic* pic, int out_slot,
                    rv_job** out);
int rv_h264_prepare(rv_session* s, const rv_h264_pic* pic, int out_slot,
                    rv_job** out);
/* rv_job_free, rv_batch_free and rv_session_destroy never wait for the
 * GPU: device memory goes back to the library's pool once the work launched
 * before the free has completed (a completion event), and later prepares
 * and sessions reuse it instead of allocating. A batch launched from a
 * captured graph is the exception: freeing it waits for the graph. */
int rv_job_free(rv_job* j);

/* A batch: a set of prepared jobs launched together. Creating it uploads
 * the small per-batch tables (picture descriptors, per-stage item
 * offsets); it can then be launched many times. */
typedef struct rv_batch rv_batch;
int rv_batch_create(rv_job* const* jobs, int n, rv_batch** out);
int rv_batch_free(rv_batch* b);

#define RV_LAUNCH_PLAIN 0   /* one kernel sequence per picture            */
#define RV_LAUNCH_BATCHED 1 /* one launch per stage over all pictures     */
#define RV_LAUNCH_GRAPH 2   /* BATCHED, replayed from a captured hipGraph */

/* Launch a batch asynchronously on the library's stream. */
int rv_batch_launch(rv_batch* b, int mode);
/* Block until everything launched so far has completed. */
int rv_sync(void);
/* Launch the batch's BATCHED stages `reps` times with a HIP event between
 * consecutive stages on the library's stream, wait, and write the average
 * duration of each stage in milliseconds (out_ms[0..4]; for a batch with
 * both codecs, HEVC stages then H.264 stages are summed per index). */
int rv_batch_profile(rv_batch* b, int reps, float out_ms[5]);
/* Record a completion marker after everything launched so far; poll or
 * wait for it (0 = complete, 1 = pending). Markers are recycled by id. */
int rv_mark(int* id_out);
int rv_mark_query(int id);
int rv_mark_wait(int id);
/* Milliseconds between two markers (both complete). */
int rv_mark_elapsed(int id_start, int id_end, float* ms);


/* Kernel variants and the autotuner. No counterpart in the reference,
 * whose tables have one kernel per SIMD tier; the reference's C3
 * configuration asks for tile shapes from an autotuner. Each HEVC stage has
 * variants: variant 0 serves any picture; the others only 8-bit or only
 * deeper pictures, compiled for one sample path with their own register and
 * occupancy budgets. Every variant gives identical output.
 * rv_batch_tune times each variant the batch's pictures allow inside full
 * reconstructions of the batch (one untimed run, then `reps` timed runs; the
 * minimum of the stage's time counts), keeps the fastest per stage for the
 * batch, and records it for its depth class (all 8-bit, all deeper, mixed).
 * A batch without its own choice launches with its class's recorded
 * variants; when its class has no record yet, its first RV_LAUNCH_BATCHED
 * or RV_LAUNCH_GRAPH launch tunes it first (a few full reconstructions of
 * the batch), unless RIVET_HIP_TUNE=0 is set. H.264 stages have one variant
 * each. */
int rv_batch_tune(rv_batch* b, int reps);
int rv_batch_variants(const rv_batch* b, int out[5]);
/* RV_E_RANGE when the batch's sample depths rule the variant out. */
int rv_batch_set_variant(rv_batch* b, int stage, int variant);
int rv_variant_count(int codec, int stage);
const char* rv_variant_name(int codec, int stage, int variant);

/* Stage switches (mirror H26X_NO_DEBLOCK / H26X_NO_SAO,
 * src/hevc/decoder.rs:1247-1248): bit 0 skip deblocking, bit 1 skip SAO.
 * Also read once from RIVET_HIP_NO_DEBLOCK / RIVET_HIP_NO_SAO. */
int rv_set_stage_mask(int mask);

/* Algorithmic bytes of a prepared job per kernel class (for roofline
 * reporting): mc, tx, dbk, sao. */
int rv_job_bytes(const rv_job* j, int64_t out[4]);

/* Achievable device-memory bandwidth: a plain 16-byte-vector copy kernel
 * (rv_copy_probe) over `bytes` bytes, `reps` times; writes (read + write
 * bytes) / average kernel time in GB/s. Also the known-size kernel that
 * calibrates hardware traffic counters. */
int rv_probe_copy(int64_t bytes, int reps, double* gbs);

/* ------------------------------------------------------------------ */
/* Per-entry parity API                                                */
/* ------------------------------------------------------------------ */
/*
 * One function per entry of HevcDsp<S> (src/dsp/hevc.rs:165-229) and
 * H264Dsp<u8> (src/dsp/h264.rs:123-170), arguments in the entry's order
 * and meaning. Each Rust slice `&[T]` / `&mut [T]` becomes a pointer and
 * its length in elements (`*_len`). Before anything runs, the call checks
 * that every element the kernel reads or writes lies inside its slice, and
 * returns RV_E_RANGE where the reference would panic on an out-of-range
 * index. Host buffers: the touched span is copied to the device, one
 * kernel runs on the library stream, written spans are copied back, and
 * the call returns when they are. Calls serialise on the library lock.
 * Samples are uint8_t (u8 tables) or uint16_t (u16 tables, 9..16-bit).
 * Entries whose arguments carry no sample type have one function.
 *
 * Not covered: the intra entries (intra_planar / intra_dc /
 * intra_angular: intra prediction stays with the host decoder) and the
 * MBAFF-only deblock_luma8_v / deblock_luma8_v_intra.
 */

/* IdctFn (src/dsp/hevc.rs:18-21): idct[log2 - 2], log2 = 2..5, coeffs
 * n x n in place; entries outside max_x / max_y are taken as zero. */
int rv_hevc_idct(int log2, int16_t* coeffs, size_t coeffs_len, int bd_shift, size_t max_x,
                 size_t max_y);
/* idst4 (trType 1), 4 x 4 in place. */
int rv_hevc_idst4(int16_t* coeffs, size_t coeffs_len, int bd_shift, size_t max_x, size_t max_y);
/* qpel_v2 / epel_v2: vertical 8 / 4 taps over 14-bit rows, >> 6. */
int rv_hevc_qpel_v2(int16_t* dst, size_t dst_len, const int16_t* src, size_t src_len,
                    size_t src_stride, size_t w, size_t h, size_t frac);
int rv_hevc_epel_v2(int16_t* dst, size_t dst_len, const int16_t* src, size_t src_len,
                    size_t src_stride, size_t w, size_t h, size_t frac);

/* The entries over samples, for S = uint8_t (_u8_) and uint16_t (_u16_).
 * AddResidualFn, CopyFn, InterpFn, UniFn, BiFn, WeightedUniFn,
 * WeightedBiFn, UniInterpFn / BiInterpFn (the fused entries; their `tmp`
 * scratch argument is not needed on the device and is omitted), SaoBandFn,
 * SaoEdgeFn, LumaDeblockFn (_v: lines `stride` apart, p/q 1 apart; _h the
 * other way round) and ChromaDeblockFn. Flags are bytes (0 / 1). */
#define RV_HEVC_SAMPLE_ENTRIES(S, sfx)                                                           \
    int rv_hevc_##sfx##_add_residual(S* dst, size_t dst_len, size_t stride, const int16_t* res,  \
                                     size_t res_len, size_t n, int max);                          \
    int rv_hevc_##sfx##_qpel_copy(int16_t* dst, size_t dst_len, const S* src, size_t src_len,     \
                                  size_t src_stride, size_t w, size_t h, int shift);              \
    int rv_hevc_##sfx##_qpel_h(int16_t* dst, size_t dst_len, const S* src, size_t src_len,        \
                               size_t src_stride, size_t w, size_t h, size_t frac, int shift);    \
    int rv_hevc_##sfx##_qpel_v(int16_t* dst, size_t dst_len, const S* src, size_t src_len,        \
                               size_t src_stride, size_t w, size_t h, size_t frac, int shift);    \
    int rv_hevc_##sfx##_epel_copy(int16_t* dst, size_t dst_len, const S* src, size_t src_len,     \
                                  size_t src_stride, size_t w, size_t h, int shift);              \
    int rv_hevc_##sfx##_epel_h(int16_t* dst, size_t dst_len, const S* src, size_t src_len,        \
                               size_t src_stride, size_t w, size_t h, size_t frac, int shift);    \
    int rv_hevc_##sfx##_epel_v(int16_t* dst, size_t dst_len, const S* src, size_t src_len,        \
                               size_t src_stride, size_t w, size_t h, size_t frac, int shift);    \
    int rv_hevc_##sfx##_uni(S* dst, size_t dst_len, size_t stride, const int16_t* src,            \
                            size_t src_len, size_t w, size_t h, int shift, int max);              \
    int rv_hevc_##sfx##_bi(S* dst, size_t dst_len, size_t stride, const int16_t* a, size_t a_len, \
                           const int16_t* b, size_t b_len, size_t w, size_t h, int shift,         \
                           int max);                                                              \
    int rv_hevc_##sfx##_weighted_uni(S* dst, size_t dst_len, size_t stride, const int16_t* src,   \
                                     size_t src_len, size_t w, size_t h, int log2_wd, int wt,     \
                                     int o, int max);                                             \
    int rv_hevc_##sfx##_weighted_bi(S* dst, size_t dst_len, size_t stride, const int16_t* a,      \
                                    size_t a_len, const int16_t* b, size_t b_len, size_t w,       \
                                    size_t h, int log2_wd, int w0, int w1, int o0, int o1,        \
                                    int max);                                                     \
    int rv_hevc_##sfx##_qpel_uni(S* dst, size_t dst_len, size_t dst_stride, const S* src,         \
                                 size_t src_len, size_t src_stride, size_t w, size_t h,           \
                                 size_t fx, size_t fy, unsigned bit_depth);                       \
    int rv_hevc_##sfx##_epel_uni(S* dst, size_t dst_len, size_t dst_stride, const S* src,         \
                                 size_t src_len, size_t src_stride, size_t w, size_t h,           \
                                 size_t fx, size_t fy, unsigned bit_depth);                       \
    int rv_hevc_##sfx##_qpel_bi(S* dst, size_t dst_len, size_t dst_stride, const S* src,          \
                                size_t src_len, size_t src_stride, size_t w, size_t h, size_t fx, \
                                size_t fy, const int16_t* other, size_t other_len,                \
                                unsigned bit_depth);                                              \
    int rv_hevc_##sfx##_epel_bi(S* dst, size_t dst_len, size_t dst_stride, const S* src,          \
                                size_t src_len, size_t src_stride, size_t w, size_t h, size_t fx, \
                                size_t fy, const int16_t* other, size_t other_len,                \
                                unsigned bit_depth);                                              \
    int rv_hevc_##sfx##_sao_band(S* dst, size_t dst_len, size_t dst_stride, const S* src,         \
                                 size_t src_len, size_t src_stride, size_t w, size_t h,           \
                                 const int16_t table[32], int shift, int max);                    \
    int rv_hevc_##sfx##_sao_edge(S* dst, size_t dst_len, const S* src, size_t src_len,            \
                                 size_t origin, size_t stride, size_t w, size_t h, ptrdiff_t na,  \
                                 ptrdiff_t nb, const int16_t off[5], int max);                    \
    int rv_hevc_##sfx##_deblock_luma_v(S* data, size_t data_len, size_t off, size_t stride,       \
                                       const int32_t beta[2], const int32_t tc[2],                \
                                       const uint8_t no_p[2], const uint8_t no_q[2], int max);    \
    int rv_hevc_##sfx##_deblock_luma_h(S* data, size_t data_len, size_t off, size_t stride,       \
                                       const int32_t beta[2], const int32_t tc[2],                \
                                       const uint8_t no_p[2], const uint8_t no_q[2], int max);    \
    int rv_hevc_##sfx##_deblock_chroma_v(S* data, size_t data_len, size_t off, size_t stride,     \
                                         const int32_t tc[4], const uint8_t no_p[4],              \
                                         const uint8_t no_q[4], int max);                         \
    int rv_hevc_##sfx##_deblock_chroma_h(S* data, size_t data_len, size_t off, size_t stride,     \
                                         const int32_t tc[4], const uint8_t no_p[4],              \
                                         const uint8_t no_q[4], int max);
RV_HEVC_SAMPLE_ENTRIES(uint8_t, u8)
RV_HEVC_SAMPLE_ENTRIES(uint16_t, u16)

/* H264Dsp<u8> (8-bit). Prediction scratch blocks are PRED_STRIDE = 16
 * samples per row (src/dsp/h264.rs:22-24). */
#define RV_H264_PRED_STRIDE 16
#define RV_H264_NO_DC INT32_MIN /* NO_DC (src/dsp/h264.rs:118) */
/* qpel[pos], pos = yf * 4 + xf: src at the six-tap window's top-left (two
 * samples left of and two rows above the block). */
int rv_h264_u8_qpel(int pos, uint8_t* dst, size_t dst_len, const uint8_t* src, size_t src_len,
                    size_t src_stride, size_t w, size_t h, int max);
int rv_h264_u8_chroma(uint8_t* dst, size_t dst_len, const uint8_t* src, size_t src_len,
                      size_t stride, size_t w, size_t h, int xf, int yf);
int rv_h264_u8_copy(uint8_t* dst, size_t dst_len, size_t stride, const uint8_t* src,
                    size_t src_len, size_t w, size_t h);
int rv_h264_u8_avg(uint8_t* dst, size_t dst_len, size_t stride, const uint8_t* a, size_t a_len,
                   const uint8_t* b, size_t b_len, size_t w, size_t h);
int rv_h264_u8_weighted_uni(uint8_t* dst, size_t dst_len, size_t stride, const uint8_t* src,
                            size_t src_len, size_t w, size_t h, int log_wd, int wt, int o,
                            int max);
int rv_h264_u8_weighted_bi(uint8_t* dst, size_t dst_len, size_t stride, const uint8_t* a,
                           size_t a_len, const uint8_t* b, size_t b_len, size_t w, size_t h,
                           int log_wd, int w0, int w1, int o0, int o1, int max);
/* Deblocking: 16 luma / 8 chroma lines; tc0[i] < 0 leaves its lines
 * alone; the _intra entries are bS 4. */
int rv_h264_u8_deblock_luma_v(uint8_t* data, size_t data_len, size_t off, size_t stride,
                              int alpha, int beta, const int16_t tc0[4], int max);
int rv_h264_u8_deblock_luma_h(uint8_t* data, size_t data_len, size_t off, size_t stride,
                              int alpha, int beta, const int16_t tc0[4], int max);
int rv_h264_u8_deblock_luma_v_intra(uint8_t* data, size_t data_len, size_t off, size_t stride,
                                    int alpha, int beta, int max);
int rv_h264_u8_deblock_luma_h_intra(uint8_t* data, size_t data_len, size_t off, size_t stride,
                                    int alpha, int beta, int max);
int rv_h264_u8_deblock_chroma_v(uint8_t* data, size_t data_len, size_t off, size_t stride,
                                int alpha, int beta, const int16_t tc0[4], int max);
int rv_h264_u8_deblock_chroma_h(uint8_t* data, size_t data_len, size_t off, size_t stride,
                                int alpha, int beta, const int16_t tc0[4], int max);
int rv_h264_u8_deblock_chroma_v_intra(uint8_t* data, size_t data_len, size_t off,
                                      size_t stride, int alpha, int beta, int max);
int rv_h264_u8_deblock_chroma_h_intra(uint8_t* data, size_t data_len, size_t off,
                                      size_t stride, int alpha, int beta, int max);
/* Transforms: 8-bit arithmetic in wrapping i16 (src/dsp/h264.rs:766-1066). */
int rv_h264_u8_idct4_add(uint8_t* dst, size_t dst_len, size_t stride, const int16_t coeffs[16],
                         int max);
int rv_h264_u8_idct8_add(uint8_t* dst, size_t dst_len, size_t stride, const int16_t coeffs[64],
                         int max);
int rv_h264_u8_idct4_dc_add(uint8_t* dst, size_t dst_len, size_t stride, int dc, int max);
int rv_h264_u8_idct8_dc_add(uint8_t* dst, size_t dst_len, size_t stride, int dc, int max);
/* residual4: dc == RV_H264_NO_DC keeps coefs[0]; otherwise dc replaces it. */
int rv_h264_u8_residual4(uint8_t* dst, size_t dst_len, size_t stride, const int32_t coefs[16],
                         int dc, int max);
int rv_h264_u8_residual8(uint8_t* dst, size_t dst_len, size_t stride, const int32_t coefs[64],
                         int max);

/* sizeof of the work-list structs, for bindings to check their layout:
 * 0 rv_hevc_cu, 1 rv_hevc_pu, 2 rv_hevc_tu, 3 rv_sao, 4 rv_hevc_pic,
 * 5 rv_h264_mb, 6 rv_h264_part, 7 rv_h264_blk, 8 rv_h264_pic, 9 rv_nz,
 * 10 rv_hevc_scaling, 11 rv_h264_scaling; -1 for an unknown index. */
int rv_abi_size(int which);

#ifdef __cplusplus
}
#endif
#endif /* RIVET_HIP_H */
