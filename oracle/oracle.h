/*
 * oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * The CPU oracle: a C++ restatement of the reference's scalar dsp kernels
 * (src/dsp/hevc.rs, src/dsp/h264.rs) and of the drivers that decide their
 * call geometry (src/hevc/inter.rs, deblock.rs, sao.rs; src/h264/inter.rs,
 * recon.rs, deblock.rs). Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker: the product path
 * (librivet_hip.so) never links, loads or calls it.
 *
 * Sample buffers are passed as void* with a sample size: 1 byte (8-bit
 * pictures, the reference's u8 tables) or 2 bytes (uint16, deeper).
 */
#ifndef RIVET_ORACLE_H
#define RIVET_ORACLE_H

#include <stdint.h>

#include "../include/rivet_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- HEVC dsp entries (HevcDsp<S>, src/dsp/hevc.rs:165-229) ---------- */
void orc_hevc_idct(int16_t* coeffs, int log2, int bd_shift, int max_x, int max_y);
void orc_hevc_idst4(int16_t* coeffs, int bd_shift);
/* scale_coefficients (src/hevc/residual.rs:1223-1283) of one n x n raster
 * (n = 1 << log2) in place; list = NULL for flat scaling, else the TU's
 * list (16 or 64 raster entries) and, for 16x16 / 32x32, its dc. */
void orc_hevc_scale_coefficients(int16_t* coeffs, int log2, int qp, int bit_depth, const uint8_t* list, int dc,
                                 int max_x, int max_y);
/* Intra prediction (src/hevc/intra.rs). A block's reference samples in
 * substitution order: scan[0 .. 2n) = p[-1][2n-1] .. p[-1][0], scan[2n] =
 * p[-1][-1], scan[2n+1 .. 4n] = p[0][-1] .. p[2n-1][-1]. */
void orc_hevc_intra_tables(int32_t* angle33, int32_t* inv15);
void orc_hevc_intra_substitute(uint16_t* scan, const uint8_t* avail, int n, int bit_depth);
/* predict_prepared with the filter decision; luma = 1 for the luma
 * component of 4:2:0 (smoothing and boundary filters); dst[y * n + x]. */
void orc_hevc_intra_pred(int32_t* dst, const uint16_t* scan, int n, int mode, int luma, int strong, int bit_depth);
/* The availability of TU tu_index's reference samples in scan order. */
int orc_hevc_intra_avail(const rv_hevc_pic* pic, int tu_index, uint8_t* avail);
/* The plain matrix product with the two clips (src/dsp/hevc.rs:1417-1442),
 * an independent statement the butterfly must equal. */
void orc_hevc_idct_matrix(const int16_t* coeffs, int n, int bd_shift, int16_t* out);
void orc_hevc_add_residual(int ss, void* dst, int stride, const int16_t* res, int n, int max);
void orc_hevc_copy(int ss, int16_t* dst, const void* src, int src_stride, int w, int h, int shift);
void orc_hevc_qpel_h(int ss, int16_t* dst, const void* src, int src_stride, int w, int h, int frac, int shift);
void orc_hevc_qpel_v(int ss, int16_t* dst, const void* src, int src_stride, int w, int h, int frac, int shift);
void orc_hevc_qpel_v2(int16_t* dst, const int16_t* src, int src_stride, int w, int h, int frac);
void orc_hevc_epel_h(int ss, int16_t* dst, const void* src, int src_stride, int w, int h, int frac, int shift);
void orc_hevc_epel_v(int ss, int16_t* dst, const void* src, int src_stride, int w, int h, int frac, int shift);
void orc_hevc_epel_v2(int16_t* dst, const int16_t* src, int src_stride, int w, int h, int frac);
void orc_hevc_uni(int ss, void* dst, int stride, const int16_t* src, int w, int h, int shift, int max);
void orc_hevc_bi(int ss, void* dst, int stride, const int16_t* a, const int16_t* b, int w, int h,
                 int shift, int max);
void orc_hevc_weighted_uni(int ss, void* dst, int stride, const int16_t* src, int w, int h,
                           int log2_wd, int wt, int o, int max);
void orc_hevc_weighted_bi(int ss, void* dst, int stride, const int16_t* a, const int16_t* b, int w,
                          int h, int log2_wd, int w0, int w1, int o0, int o1, int max);
/* Fused interpolation + default combine (src/dsp/hevc.rs:887-991); luma
 * selects the 8-tap (qpel) or 4-tap (epel) kernel. other == NULL: uni. */
void orc_hevc_mc_fused(int ss, int luma, void* dst, int dst_stride, const void* src,
                       int src_stride, int w, int h, int fx, int fy, const int16_t* other,
                       int bit_depth);
void orc_hevc_sao_band(int ss, void* dst, int dst_stride, const void* src, int src_stride, int w,
                       int h, const int16_t* table32, int shift, int max);
void orc_hevc_sao_edge(int ss, void* dst, const void* src, int origin, int stride, int w, int h,
                       int na, int nb, const int16_t* off5, int max);
/* vertical = 1: _v (lines stride apart, p/q 1 apart); 0: _h. */
void orc_hevc_deblock_luma(int ss, int vertical, void* data, int off, int stride,
                           const int32_t beta[2], const int32_t tc[2], const uint8_t no_p[2],
                           const uint8_t no_q[2], int max);
void orc_hevc_deblock_chroma(int ss, int vertical, void* data, int off, int stride,
                             const int32_t tc[4], const uint8_t no_p[4], const uint8_t no_q[4],
                             int max);

/* ---- HEVC whole picture ------------------------------------------------
 * slot_planes[3 * slot + c] = tightly packed plane c of DPB slot `slot`.
 * out[3] receives the reconstructed picture (tightly packed). stop_after:
 * 1 = prediction only, 2 = + residual, 3 = + deblocking, 0/4 = + SAO. */
int orc_hevc_reconstruct(const rv_hevc_pic* pic, const void* const* slot_planes, int n_slots,
                         void* const out[3], int stop_after);
/* The same on `threads` threads: PUs and TUs in ranges, deblocking in
 * bands of 16 luma rows per pass, SAO by CTB rows (the CPU baseline's
 * parallelism within a picture). Output identical to the above. */
int orc_hevc_reconstruct_mt(const rv_hevc_pic* pic, const void* const* slot_planes, int n_slots,
                            void* const out[3], int stop_after, int threads);

/* Boundary strengths of the deblocking filter (boundary_strengths,
 * src/hevc/deblock.rs:62-180) from a picture's work lists: ver[by*w4+bx]
 * for the vertical edge at the left of 4x4 block (bx, by), hor[...] for
 * the horizontal edge above it; w4 = width/4, h4 = height/4 entries each,
 * 0 where no edge is filtered. */
int orc_hevc_bs_map(const rv_hevc_pic* pic, uint8_t* ver, uint8_t* hor);

/* ---- H.264 dsp entries (H264Dsp<S>, src/dsp/h264.rs:123-170), 8-bit -- */
void orc_h264_qpel(int pos, uint8_t* dst, const uint8_t* src, int stride, int w, int h, int max);
void orc_h264_chroma(uint8_t* dst, const uint8_t* src, int stride, int w, int h, int xf, int yf);
void orc_h264_copy(uint8_t* dst, int stride, const uint8_t* src, int w, int h);
void orc_h264_avg(uint8_t* dst, int stride, const uint8_t* a, const uint8_t* b, int w, int h);
void orc_h264_idct4_add(uint8_t* dst, int stride, const int16_t* coeffs, int max);
void orc_h264_idct8_add(uint8_t* dst, int stride, const int16_t* coeffs, int max);
/* dc_add_scalar, n = 4 or 8. */
void orc_h264_dc_add(uint8_t* dst, int stride, int dc, int n, int max);
void orc_h264_weighted_uni(uint8_t* dst, int stride, const uint8_t* src, int w, int h, int log_wd,
                           int wt, int o, int max);
void orc_h264_weighted_bi(uint8_t* dst, int stride, const uint8_t* a, const uint8_t* b, int w,
                          int h, int log_wd, int w0, int w1, int o0, int o1, int max);
/* tc0 == NULL: the bS 4 (intra) kernel. chroma selects the 8-line 4:2:0
 * chroma kernel, else the 16-line luma one. */
void orc_h264_deblock(int chroma, int vertical, uint8_t* data, int off, int stride, int alpha,
                      int beta, const int16_t* tc0, int max);
void orc_h264_residual4(uint8_t* dst, int stride, const int32_t* coefs, int dc, int has_dc,
                        int max);
void orc_h264_residual8(uint8_t* dst, int stride, const int32_t* coefs, int max);
/* LevelScale4x4 / LevelScale8x8 (Dequant::new, src/h264/transform.rs:24-56)
 * of list `list` at qP % 6 = m, raster pos; sl = NULL: flat 16. */
int orc_h264_level_scale4(const rv_h264_scaling* sl, int list, int m, int pos);
int orc_h264_level_scale8(const rv_h264_scaling* sl, int list, int m, int pos);
/* dequant_level (src/h264/mb.rs:1420-1428). */
int32_t orc_h264_dequant_level(int32_t level, int32_t table, int shift);
/* luma_dc_transform / chroma_dc_transform_420 (src/h264/transform.rs:
 * 179-228), in place on 16 / 4 raster DC levels. */
void orc_h264_luma_dc(int32_t* dc, int scale00, int qp);
void orc_h264_chroma_dc(int32_t* dc, int scale00, int qp);
void orc_h264_implicit_pair(int cur_poc, int poc0, int poc1, int long0, int long1, int32_t out[2]);

int orc_h264_reconstruct(const rv_h264_pic* pic, const void* const* slot_planes, int n_slots,
                         void* const out[3], int stop_after);
/* Boundary strengths of the deblocking filter (src/h264/deblock.rs:
 * 97-292, 617-756) from a picture's work lists: per macroblock (raster),
 * 8 words: the vertical edges 0..3 (0 = the macroblock's left edge), then
 * the horizontal edges 0..3 (0 = its top edge). Byte k of a word is the bS
 * of the edge's 4-sample segment k (top to bottom / left to right). */
int orc_h264_bs_map(const rv_h264_pic* pic, uint32_t* out);
/* H.264 intra prediction (8.3, src/h264/intra.rs) on a caller's plane. */
int orc_h264_intra_pred(int kind, int n, int mode, int av, uint8_t* d, int stride);
void orc_h264_filter8(const int32_t in[25], int av, int32_t out[25]);
int orc_h264_block_avail(const rv_h264_pic* pic, int addr, int bx, int by, int size);

/* ---- Decoded picture hash (src/hevc/hash.rs) -------------------------
 * Over one component's message (n bytes): crc16 (:160-174), checksum
 * (:176-189; width in samples, wide = two bytes per sample) and md5
 * (:191-250). */
uint16_t orc_crc16(const uint8_t* data, size_t n);
uint32_t orc_checksum(const uint8_t* data, size_t n, int width, int wide);
void orc_md5(const uint8_t* data, size_t n, uint8_t out[16]);

#ifdef __cplusplus
}
#endif
#endif
