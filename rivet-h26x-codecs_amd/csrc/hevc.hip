// HEVC picture reconstruction on gfx950: five stages, each one launch over
// every picture of a batch (work items located by a prefix search).
//
//   stage 0  side data (per-4x4 QP / prediction mode / edges / PU map),
//            intra prediction blocks, motion-compensated tiles
//   stage 1  inverse DCT / DST + residual add; boundary strengths
//   stage 2  deblocking, all vertical edges
//   stage 3  deblocking, all horizontal edges
//   stage 4  SAO from the deblocked picture into the DPB slot
//
// Arithmetic restates the reference's scalar kernels (src/dsp/hevc.rs) and
// drivers (src/hevc/inter.rs, deblock.rs, sao.rs) bit for bit; each device
// function cites what it follows.
#include <hip/hip_runtime.h>

#include "rv_device.h"
#include "rv_kernels.h"

namespace rv {
namespace {

// Table 8-12 / 8-13 (src/hevc/tables_gen.rs:173-191).
__constant__ int8_t kQpel[4][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0},
    {-1, 4, -10, 58, 17, -5, 1, 0},
    {-1, 4, -11, 40, 40, -11, 4, -1},
    {0, 1, -5, 17, 58, -10, 4, -1},
};
__constant__ int8_t kEpel[8][4] = {
    {0, 0, 0, 0},    {-2, 58, 10, -2}, {-4, 54, 16, -2}, {-6, 46, 28, -4},
    {-4, 36, 36, -4}, {-4, 28, 46, -6}, {-2, 16, 54, -4}, {-2, 10, 58, -2},
};
// tC' / beta' (src/hevc/tables_gen.rs:195-199).
__constant__ uint8_t kTc[54] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4,
                                4, 4, 5, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 22, 24};
__constant__ uint8_t kBeta[52] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  6,  7,
                                  8,  9,  10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 22, 24, 26, 28, 30, 32,
                                  34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64};
// The 32 distinct magnitudes of the 8.6.4.2 matrix (its first column).
__constant__ int16_t kDctC[33] = {64, 90, 90, 90, 89, 88, 87, 85, 83, 82, 80, 78, 75, 73, 70, 67, 64,
                                  61, 57, 54, 50, 46, 43, 38, 36, 31, 25, 22, 18, 13, 9,  4,  0};
__constant__ int8_t kDst[4][4] = {{29, 55, 74, 84}, {74, 74, 0, -74}, {84, -29, -74, 55}, {55, -84, 74, -29}};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }
__device__ __forceinline__ int sgn(int v) { return (v > 0) - (v < 0); }

// Entry (k, n) of the 32-point matrix: +-c[m], m = (2n + 1) k mod 128
// folded by the cosine symmetries.
__device__ __forceinline__ int dct_entry(int k, int n) {
    int a = ((2 * n + 1) * k) & 127;
    if (a <= 32) return kDctC[a];
    if (a <= 64) return -kDctC[64 - a];
    if (a <= 96) return -kDctC[a - 64];
    return kDctC[128 - a];
}

__device__ __forceinline__ int find_pic(const int32_t* pre, int n, int it) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= it) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------
// Stage 0: side data, intra blocks, motion compensation
// ---------------------------------------------------------------------

template <typename S>
__device__ void intra_copy(const HevcPicDev& P, const rv_hevc_cu& cu, int tid) {
    int s = 1 << cu.log2_size;
    const uint16_t* pr = P.intra_pred + cu.pred_off;
    for (int c = 0; c < 3; ++c) {
        int n = c ? s / 2 : s, px = c ? cu.x / 2 : cu.x, py = c ? cu.y / 2 : cu.y;
        S* dst = (S*)P.cur[c];
        for (int e = tid; e < n * n; e += 64) {
            int yy = e / n, xx = e - yy * n;
            dst[(size_t)(py + yy) * P.pitch[c] + px + xx] = (S)pr[e];
        }
        pr += n * n;
    }
}

// One component of one list of an MC tile: the 14-bit prediction
// (interp_scalar, src/dsp/hevc.rs:823-882) of up to NOUT outputs per lane.
// The (w + taps - 1) x (h + taps - 1) window is staged in LDS from
// coalesced row loads, coordinates clamped to the picture (the replicated
// border of src/hevc/frame.rs:169-214 and the gather of
// src/hevc/inter.rs:116-123 read the same samples).
template <typename S, int TAPS, int WIN_W, int OUT_W, int NOUT>
__device__ void mc_component(const S* __restrict__ ref, int pitch, int pw, int ph, int xi, int yi, int fx, int fy,
                             int bw, int bh, int bd, int16_t* win, int16_t* tmp, int tid, int out[NOUT]) {
    constexpr int REACH = TAPS / 2 - 1;
    const int ww = bw + TAPS - 1, hh = bh + TAPS - 1;
    const int x0 = xi - REACH, y0 = yi - REACH;
    // Stage the window: lanes cover WIN_W-wide row slots.
    for (int e = tid; e < hh * WIN_W; e += 64) {
        int r = e / WIN_W, c = e - r * WIN_W;
        if (c < ww) {
            int gx = clampi(x0 + c, 0, pw - 1), gy = clampi(y0 + r, 0, ph - 1);
            win[r * WIN_W + c] = (int16_t)ref[(size_t)gy * pitch + gx];
        }
    }
    __syncthreads();
    const int shift1 = min(bd, 12) - 8, shift3 = 14 - bd;
    const int8_t* fxt = TAPS == 8 ? kQpel[fx] : kEpel[fx];
    const int8_t* fyt = TAPS == 8 ? kQpel[fy] : kEpel[fy];
    if (fx != 0 && fy != 0) {
        // First stage over every window row, truncated to 14 bits (i16).
        for (int e = tid; e < hh * OUT_W; e += 64) {
            int r = e / OUT_W, c = e - r * OUT_W;
            if (c < bw) {
                int acc = 0;
#pragma unroll
                for (int k = 0; k < TAPS; ++k) acc += fxt[k] * (int)win[r * WIN_W + c + k];
                tmp[r * OUT_W + c] = (int16_t)(acc >> shift1);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k2 = 0; k2 < NOUT; ++k2) {
        int o = tid + 64 * k2;
        int oy = o / OUT_W, ox = o - oy * OUT_W;
        int v = 0;
        if (ox < bw && oy < bh) {
            if (fx == 0 && fy == 0) {
                v = (int16_t)((int)win[(oy + REACH) * WIN_W + ox + REACH] << shift3);
            } else if (fy == 0) {
                int acc = 0;
#pragma unroll
                for (int k = 0; k < TAPS; ++k) acc += fxt[k] * (int)win[(oy + REACH) * WIN_W + ox + k];
                v = (int16_t)(acc >> shift1);
            } else if (fx == 0) {
                int acc = 0;
#pragma unroll
                for (int k = 0; k < TAPS; ++k) acc += fyt[k] * (int)win[(oy + k) * WIN_W + ox + REACH];
                v = (int16_t)(acc >> shift1);
            } else {
                int acc = 0;
#pragma unroll
                for (int k = 0; k < TAPS; ++k) acc += fyt[k] * (int)tmp[(oy + k) * OUT_W + ox];
                v = (int16_t)(acc >> 6);
            }
        }
        out[k2] = v;
    }
    __syncthreads();
}

// The combine of predict_block (src/hevc/inter.rs:712-741) with the
// scalar combiners (src/dsp/hevc.rs:997-1087).
__device__ __forceinline__ int combine(const HevcPicDev& P, int c, int dir, const int8_t* ref_idx, int a, int b,
                                       int bd) {
    const int max = (1 << bd) - 1;
    const bool both = dir == 3;
    if (!P.weighted) {
        if (!both) {
            int shift = 14 - bd;
            int v = dir == 1 ? a : b;
            return clampi((v + (1 << (shift - 1))) >> shift, 0, max);
        }
        int shift = 15 - bd;
        return clampi((a + b + (1 << (shift - 1))) >> shift, 0, max);
    }
    const int lwd = P.log2_wd[c];
    if (!both) {
        int l = dir == 1 ? 0 : 1;
        int v = l == 0 ? a : b;
        int wt = P.wp_w[l][ref_idx[l]][c], o = P.wp_o[l][ref_idx[l]][c];
        if (lwd >= 1) return clampi(((v * wt + (1 << (lwd - 1))) >> lwd) + o, 0, max);
        return clampi(v * wt + o, 0, max);
    }
    int w0 = P.wp_w[0][ref_idx[0]][c], w1 = P.wp_w[1][ref_idx[1]][c];
    int o0 = P.wp_o[0][ref_idx[0]][c], o1 = P.wp_o[1][ref_idx[1]][c];
    return clampi((a * w0 + b * w1 + ((o0 + o1 + 1) << lwd)) >> (lwd + 1), 0, max);
}

template <typename S>
__device__ void mc_tile(const HevcPicDev& P, const McTile& t, int tid, int16_t* win, int16_t* tmp) {
    const rv_hevc_pu pu = P.pus[t.pu];
    const int x = pu.x + t.dx, y = pu.y + t.dy, tw = t.tw, th = t.th, bd = P.bd;
    const int dir = pu.dir;
    // Luma: outputs in 16-wide slots, four per lane.
    {
        int pred[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
        for (int l = 0; l < 2; ++l) {
            if (!((dir >> l) & 1)) continue;
            const S* ref = (const S*)P.ref[l][pu.ref_idx[l]][0];
            int mvx = pu.mv[l].x, mvy = pu.mv[l].y;
            mc_component<S, 8, 32, 16, 4>(ref, P.pitch[0], P.pw[0], P.ph[0], x + (mvx >> 2), y + (mvy >> 2), mvx & 3,
                                          mvy & 3, tw, th, bd, win, tmp, tid, pred[l]);
        }
        S* dst = (S*)P.cur[0];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int o = tid + 64 * k, oy = o >> 4, ox = o & 15;
            if (ox < tw && oy < th)
                dst[(size_t)(y + oy) * P.pitch[0] + x + ox] = (S)combine(P, 0, dir, pu.ref_idx, pred[0][k], pred[1][k], bd);
        }
    }
    // Chroma (4:2:0): eighth-sample vectors in chroma units
    // (src/hevc/inter.rs:581-594), outputs in 8-wide slots, one per lane.
    const int cw = tw >> 1, ch = th >> 1, cx = x >> 1, cy = y >> 1;
    for (int c = 1; c < 3; ++c) {
        int pred[2][1] = {{0}, {0}};
        for (int l = 0; l < 2; ++l) {
            if (!((dir >> l) & 1)) continue;
            const S* ref = (const S*)P.ref[l][pu.ref_idx[l]][c];
            int mvx = pu.mv[l].x, mvy = pu.mv[l].y;
            mc_component<S, 4, 16, 8, 1>(ref, P.pitch[c], P.pw[c], P.ph[c], cx + (mvx >> 3), cy + (mvy >> 3), mvx & 7,
                                         mvy & 7, cw, ch, bd, win, tmp, tid, pred[l]);
        }
        int oy = tid >> 3, ox = tid & 7;
        if (ox < cw && oy < ch) {
            S* dst = (S*)P.cur[c];
            dst[(size_t)(cy + oy) * P.pitch[c] + cx + ox] = (S)combine(P, c, dir, pu.ref_idx, pred[0][0], pred[1][0], bd);
        }
    }
}

__global__ void __launch_bounds__(64) hevc_stage0(BatchDev b) {
    __shared__ int16_t win[23 * 32];
    __shared__ int16_t tmp[23 * 16];
    const int32_t* pre = b.prefix + 0 * (b.n + 1);
    const int it = blockIdx.x;
    const int p = find_pic(pre, b.n, it);
    const HevcPicDev& P = *b.hevc[p];
    int i = it - pre[p];
    const int tid = threadIdx.x;
    if (i < P.n_cus) {
        // CU: QP, prediction mode, coding-block edges (src/hevc/ctu.rs:751-800).
        const rv_hevc_cu cu = P.cus[i];
        int n4 = 1 << (cu.log2_size - 2);
        int bx0 = cu.x >> 2, by0 = cu.y >> 2;
        for (int e = tid; e < n4 * n4; e += 64) {
            int by = by0 + e / n4, bx = bx0 + e % n4;
            size_t q = (size_t)by * P.w4 + bx;
            bool left = bx == bx0, top = by == by0;
            P.qp[q] = cu.qp_y;
            P.cuinfo[q] = (uint8_t)(cu.intra | (left << 1) | (top << 2));
            if (cu.intra) P.pumap[q] = 0x3FFFFFFFu | ((uint32_t)left << 30) | ((uint32_t)top << 31);
        }
        if (cu.intra) {
            if (P.bd == 8) intra_copy<uint8_t>(P, cu, tid);
            else intra_copy<uint16_t>(P, cu, tid);
        }
        return;
    }
    i -= P.n_cus;
    if (i < P.n_pus) {
        const rv_hevc_pu pu = P.pus[i];
        int nw = pu.w >> 2, nh = pu.h >> 2, bx0 = pu.x >> 2, by0 = pu.y >> 2;
        for (int e = tid; e < nw * nh; e += 64) {
            int by = by0 + e / nw, bx = bx0 + e % nw;
            P.pumap[(size_t)by * P.w4 + bx] = (uint32_t)i | ((uint32_t)(bx == bx0) << 30) | ((uint32_t)(by == by0) << 31);
        }
        return;
    }
    i -= P.n_pus;
    if (i < P.n_ltus) {
        // Luma TU edges and cbf (src/hevc/ctu.rs:1493-1510).
        const rv_hevc_tu tu = P.ltus[i];
        int n4 = 1 << (tu.log2 - 2), bx0 = tu.x >> 2, by0 = tu.y >> 2;
        uint8_t cbf = (tu.flags & RV_TU_CODED) ? 1 : 0;
        for (int e = tid; e < n4 * n4; e += 64) {
            int by = by0 + e / n4, bx = bx0 + e % n4;
            P.tuinfo[(size_t)by * P.w4 + bx] = (uint8_t)(cbf | ((bx == bx0) << 1) | ((by == by0) << 2));
        }
        return;
    }
    i -= P.n_ltus;
    const McTile t = P.tiles[i];
    if (P.bd == 8) mc_tile<uint8_t>(P, t, tid, win, tmp);
    else mc_tile<uint16_t>(P, t, tid, win, tmp);
}

// ---------------------------------------------------------------------
// Stage 1: inverse transform + residual add; boundary strengths
// ---------------------------------------------------------------------

// idct_scalar / idst4_scalar (src/dsp/hevc.rs:499-556) in matrix form: the
// partial butterfly is the matrix product regrouped exactly (the
// reference's butterfly_equals_matrix, :1511-1536), and coefficients
// outside the nonzero bounds contribute nothing. Stage 1 rounds by 7 and
// clips to i16; stage 2 rounds by bdShift = 20 - BitDepth and clips; then
// add_residual_scalar (:558-566). `lane` indexes the TU's outputs.
template <typename S>
__device__ void tu_residual(const HevcPicDev& P, const rv_hevc_tu& tu, int lane, int nlanes, int16_t* c,
                            int16_t* t1, const int8_t* mat, bool active) {
    const int log2 = tu.log2, n = 1 << log2, nn = n * n;
    const bool dst = (tu.flags & RV_TU_DST) != 0;
    const int step = 32 >> log2;
    // The DST ignores the bounds (idst4_scalar takes all sixteen).
    const int mx = dst ? 3 : tu.max_x, my = dst ? 3 : tu.max_y;
    if (active)
        for (int e = lane; e < nn; e += nlanes) c[e] = P.coefs[tu.coef_off + e];
    __syncthreads();
    // Stage 1 (columns): t1[i][x] = clip((sum_j M[j][i] c[j][x] + 64) >> 7).
    if (active)
        for (int e = lane; e < nn; e += nlanes) {
            int i = e >> log2, x = e & (n - 1);
            int s = 0;
            if (x <= mx)
                for (int j = 0; j <= my; ++j) {
                    int m = dst ? (int)kDst[j][i] : (int)mat[(j * step) * 32 + i];
                    s += m * (int)c[j * n + x];
                }
            t1[e] = (int16_t)clampi((s + 64) >> 7, -32768, 32767);
        }
    __syncthreads();
    const int bd_shift = 20 - P.bd, round2 = 1 << (bd_shift - 1), max = (1 << P.bd) - 1;
    if (active) {
        S* plane = (S*)P.cur[tu.comp];
        int pitch = P.pitch[tu.comp];
        for (int e = lane; e < nn; e += nlanes) {
            int y = e >> log2, i = e & (n - 1);
            int s = 0;
            for (int j = 0; j <= mx; ++j) {
                int m = dst ? (int)kDst[j][i] : (int)mat[(j * step) * 32 + i];
                s += m * (int)t1[y * n + j];
            }
            int r = clampi((s + round2) >> bd_shift, -32768, 32767);
            S* d = plane + (size_t)(tu.y + y) * pitch + tu.x + i;
            *d = (S)clampi((int)*d + r, 0, max);
        }
    }
}

struct MotionRef {
    int uses[2];
    int delta[2], lt[2];
    int mv[2][2];
};

__device__ __forceinline__ MotionRef motion_of(const HevcPicDev& P, uint32_t pm) {
    MotionRef m;
    const uint32_t idx = pm & 0x3FFFFFFFu;
    if (idx >= (uint32_t)P.n_pus) {
        m.uses[0] = m.uses[1] = 0;
        m.delta[0] = m.delta[1] = m.lt[0] = m.lt[1] = 0;
        m.mv[0][0] = m.mv[0][1] = m.mv[1][0] = m.mv[1][1] = 0;
        return m;
    }
    const rv_hevc_pu pu = P.pus[idx];
    for (int l = 0; l < 2; ++l) {
        m.uses[l] = (pu.dir >> l) & 1;
        int ri = m.uses[l] ? pu.ref_idx[l] : 0;
        m.delta[l] = m.uses[l] ? P.ref_delta[l][ri] : 0;
        m.lt[l] = m.uses[l] ? P.ref_lt[l][ri] : 0;
        m.mv[l][0] = pu.mv[l].x;
        m.mv[l][1] = pu.mv[l].y;
    }
    return m;
}

__device__ __forceinline__ bool mv_far(const int* p, const int* q) {
    return abs(p[0] - q[0]) >= 4 || abs(p[1] - q[1]) >= 4;
}

// motion_bs (src/hevc/deblock.rs:15-60).
__device__ int motion_bs(const MotionRef& a, const MotionRef& b) {
    int na = a.uses[0] + a.uses[1], nb = b.uses[0] + b.uses[1];
    if (na != nb) return 1;
    if (na == 1) {
        int la = a.uses[0] ? 0 : 1, lb = b.uses[0] ? 0 : 1;
        if (a.delta[la] != b.delta[lb] || a.lt[la] != b.lt[lb]) return 1;
        return mv_far(a.mv[la], b.mv[lb]);
    }
    bool a0b0 = a.delta[0] == b.delta[0] && a.lt[0] == b.lt[0];
    bool a1b1 = a.delta[1] == b.delta[1] && a.lt[1] == b.lt[1];
    bool a0b1 = a.delta[0] == b.delta[1] && a.lt[0] == b.lt[1];
    bool a1b0 = a.delta[1] == b.delta[0] && a.lt[1] == b.lt[0];
    if (!((a0b0 && a1b1) || (a0b1 && a1b0))) return 1;
    bool a0a1 = a.delta[0] == a.delta[1] && a.lt[0] == a.lt[1];
    if (!a0a1) {
        if (a0b0) return mv_far(a.mv[0], b.mv[0]) || mv_far(a.mv[1], b.mv[1]);
        return mv_far(a.mv[0], b.mv[1]) || mv_far(a.mv[1], b.mv[0]);
    }
    bool c1 = mv_far(a.mv[0], b.mv[0]) || mv_far(a.mv[1], b.mv[1]);
    bool c2 = mv_far(a.mv[0], b.mv[1]) || mv_far(a.mv[1], b.mv[0]);
    return c1 && c2;
}

// boundary_strengths (src/hevc/deblock.rs:62-180), one 4x4 block per
// thread: bS of its left and top edges on the 8x8 grid; one slice, one tile.
__device__ void block_bs(const HevcPicDev& P, int q) {
    const int w4 = P.w4;
    const int bx = q % w4, by = q / w4, x = bx * 4, y = by * 4;
    int out = 0;
    if (P.deblock) {
        const uint8_t ci = P.cuinfo[q], ti = P.tuinfo[q];
        const uint32_t pm = P.pumap[q];
        const bool intra_q = ci & 1;
        const bool tu_l = ((ci >> 1) | (ti >> 1)) & 1, pu_l = (pm >> 30) & 1;
        const bool tu_t = ((ci >> 2) | (ti >> 2)) & 1, pu_t = (pm >> 31) & 1;
        if (x > 0 && (x & 7) == 0 && (tu_l || pu_l)) {
            int p = q - 1;
            int b;
            if (intra_q || (P.cuinfo[p] & 1)) b = 2;
            else if (tu_l && ((P.tuinfo[p] | ti) & 1)) b = 1;
            else b = motion_bs(motion_of(P, P.pumap[p]), motion_of(P, pm));
            out |= b;
        }
        if (y > 0 && (y & 7) == 0 && (tu_t || pu_t)) {
            int p = q - w4;
            int b;
            if (intra_q || (P.cuinfo[p] & 1)) b = 2;
            else if (tu_t && ((P.tuinfo[p] | ti) & 1)) b = 1;
            else b = motion_bs(motion_of(P, P.pumap[p]), motion_of(P, pm));
            out |= b << 2;
        }
    }
    P.bs[q] = (uint8_t)out;
}

__global__ void __launch_bounds__(64) hevc_stage1(BatchDev b) {
    __shared__ int16_t cbuf[32 * 32];
    __shared__ int16_t t1[32 * 32];
    __shared__ int8_t mat[32 * 32];
    const int32_t* pre = b.prefix + 1 * (b.n + 1);
    const int it = blockIdx.x;
    const int p = find_pic(pre, b.n, it);
    const HevcPicDev& P = *b.hevc[p];
    int i = it - pre[p];
    const int tid = threadIdx.x;
    const int n4items = (P.n_tx4 + 3) >> 2;
    const int ntx = n4items + P.n_txbig;
    if (i < ntx) {
        for (int e = tid; e < 32 * 32; e += 64) mat[e] = (int8_t)dct_entry(e >> 5, e & 31);
        if (i < n4items) {
            // Four 4x4 TUs, sixteen lanes each.
            int k = i * 4 + (tid >> 4);
            bool active = k < P.n_tx4;
            rv_hevc_tu tu = P.ctus[active ? k : P.n_tx4 - 1];
            int16_t* c = cbuf + (tid >> 4) * 16;
            int16_t* tt = t1 + (tid >> 4) * 16;
            if (P.bd == 8) tu_residual<uint8_t>(P, tu, tid & 15, 16, c, tt, mat, active);
            else tu_residual<uint16_t>(P, tu, tid & 15, 16, c, tt, mat, active);
        } else {
            const rv_hevc_tu tu = P.ctus[P.n_tx4 + (i - n4items)];
            if (P.bd == 8) tu_residual<uint8_t>(P, tu, tid, 64, cbuf, t1, mat, true);
            else tu_residual<uint16_t>(P, tu, tid, 64, cbuf, t1, mat, true);
        }
        return;
    }
    i -= ntx;
    int q = i * 64 + tid;
    if (q < P.w4 * P.h4) block_bs(P, q);
}

// ---------------------------------------------------------------------
// Stages 2, 3: deblocking
// ---------------------------------------------------------------------

// luma_edge_scalar (src/dsp/hevc.rs:1234-1324) on one 4-line segment held
// in registers: s[line][k], k = 0..7 for p3 p2 p1 p0 q0 q1 q2 q3.
__device__ void luma_segment(int s[4][8], int beta, int tc, int max, bool& changed) {
    auto P_ = [&](int line, int k) { return s[line][3 - k]; };  // p_k
    auto Q_ = [&](int line, int k) { return s[line][4 + k]; };  // q_k
    int dp0 = abs(P_(0, 2) - 2 * P_(0, 1) + P_(0, 0));
    int dp3 = abs(P_(3, 2) - 2 * P_(3, 1) + P_(3, 0));
    int dq0 = abs(Q_(0, 2) - 2 * Q_(0, 1) + Q_(0, 0));
    int dq3 = abs(Q_(3, 2) - 2 * Q_(3, 1) + Q_(3, 0));
    int dpq0 = dp0 + dq0, dpq3 = dp3 + dq3, dp = dp0 + dp3, dq = dq0 + dq3;
    if (dpq0 + dpq3 >= beta) return;
    auto dsam = [&](int line, int dpq) {
        return dpq < (beta >> 2) &&
               abs(P_(line, 3) - P_(line, 0)) + abs(Q_(line, 0) - Q_(line, 3)) < (beta >> 3) &&
               abs(P_(line, 0) - Q_(line, 0)) < ((5 * tc + 1) >> 1);
    };
    bool strong = dsam(0, 2 * dpq0) && dsam(3, 2 * dpq3);
    bool dep = dp < ((beta + (beta >> 1)) >> 3);
    bool deq = dq < ((beta + (beta >> 1)) >> 3);
    changed = true;
#pragma unroll
    for (int line = 0; line < 4; ++line) {
        int p0 = s[line][3], p1 = s[line][2], p2 = s[line][1], p3 = s[line][0];
        int q0 = s[line][4], q1 = s[line][5], q2 = s[line][6], q3 = s[line][7];
        if (strong) {
            s[line][3] = clampi((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3, p0 - 2 * tc, p0 + 2 * tc);
            s[line][2] = clampi((p2 + p1 + p0 + q0 + 2) >> 2, p1 - 2 * tc, p1 + 2 * tc);
            s[line][1] = clampi((2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3, p2 - 2 * tc, p2 + 2 * tc);
            s[line][4] = clampi((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3, q0 - 2 * tc, q0 + 2 * tc);
            s[line][5] = clampi((p0 + q0 + q1 + q2 + 2) >> 2, q1 - 2 * tc, q1 + 2 * tc);
            s[line][6] = clampi((p0 + q0 + q1 + 3 * q2 + 2 * q3 + 4) >> 3, q2 - 2 * tc, q2 + 2 * tc);
        } else {
            int delta = (9 * (q0 - p0) - 3 * (q1 - p1) + 8) >> 4;
            if (abs(delta) < tc * 10) {
                delta = clampi(delta, -tc, tc);
                s[line][3] = clampi(p0 + delta, 0, max);
                s[line][4] = clampi(q0 - delta, 0, max);
                if (dep) {
                    int d = clampi((((p2 + p0 + 1) >> 1) - p1 + delta) >> 1, -(tc >> 1), tc >> 1);
                    s[line][2] = clampi(p1 + d, 0, max);
                }
                if (deq) {
                    int d = clampi((((q2 + q0 + 1) >> 1) - q1 - delta) >> 1, -(tc >> 1), tc >> 1);
                    s[line][5] = clampi(q1 + d, 0, max);
                }
            }
        }
    }
}

// Luma beta / tc of one segment (src/hevc/deblock.rs:264-279).
__device__ __forceinline__ void luma_thresholds(const HevcPicDev& P, int b, int qp_p, int qp_q, int* beta, int* tc) {
    int qp = (qp_p + qp_q + 1) >> 1, sh = P.bd - 8;
    *beta = (int)kBeta[clampi(qp + P.beta_offset, 0, 51)] << sh;
    *tc = (int)kTc[clampi(qp + 2 * (b - 1) + P.tc_offset, 0, 53)] << sh;
}

// chroma_qp (src/hevc/ctu.rs:3072-3084), 4:2:0.
__device__ __forceinline__ int chroma_qp420(int qpi) {
    if (qpi < 30) return qpi;
    if (qpi >= 43) return qpi - 6;
    const int t[14] = {29, 30, 31, 32, 33, 33, 34, 34, 35, 35, 36, 36, 37, 37};
    return t[qpi - 30];
}

// Chroma tc for Cb (c = 0) / Cr (c = 1) (src/hevc/deblock.rs:281-294).
__device__ __forceinline__ int chroma_tc(const HevcPicDev& P, int qp_p, int qp_q, int c) {
    int qpi = ((qp_p + qp_q + 1) >> 1) + (c == 0 ? P.cb_qp_offset : P.cr_qp_offset);
    int qpc = qpi < 0 ? qpi : chroma_qp420(qpi);
    return (int)kTc[clampi(qpc + 2 + P.tc_offset, 0, 53)] << (P.bd - 8);
}

// chroma_edge_scalar (src/dsp/hevc.rs:1326-1353), two lines. d points at
// q0 of line 0; `step` crosses the edge, `along` moves between lines.
template <typename S>
__device__ void chroma_lines(S* d, long step, long along, int tc, int max) {
    if (tc == 0) return;
#pragma unroll
    for (int line = 0; line < 2; ++line) {
        S* b = d + line * along;
        int p0 = b[-step], p1 = b[-2 * step], q0 = b[0], q1 = b[step];
        int delta = clampi((((q0 - p0) << 2) + p1 - q1 + 4) >> 3, -tc, tc);
        b[-step] = (S)clampi(p0 + delta, 0, max);
        b[0] = (S)clampi(q0 - delta, 0, max);
    }
}

template <typename S>
__device__ void deblock_v_seg(const HevcPicDev& P, int t) {
    const int nx = P.w >> 3;
    const int kx = t % nx, j = t / nx;
    if (kx == 0 || j >= P.h4) return;
    const int q = j * P.w4 + 2 * kx;
    const int b = P.bs[q] & 3;
    if (!b) return;
    const int x = kx * 8, y = j * 4, max = (1 << P.bd) - 1;
    const int qpp = P.qp[q - 1], qpq = P.qp[q];
    int beta, tc;
    luma_thresholds(P, b, qpp, qpq, &beta, &tc);
    if (beta != 0 || tc != 0) {
        S* Y = (S*)P.cur[0];
        const int pitch = P.pitch[0];
        int s[4][8];
#pragma unroll
        for (int line = 0; line < 4; ++line) {
            const S* r = Y + (size_t)(y + line) * pitch + x - 4;
#pragma unroll
            for (int k = 0; k < 8; ++k) s[line][k] = r[k];
        }
        bool changed = false;
        luma_segment(s, beta, tc, max, changed);
        if (changed) {
#pragma unroll
            for (int line = 0; line < 4; ++line) {
                S* r = Y + (size_t)(y + line) * pitch + x - 4;
#pragma unroll
                for (int k = 1; k < 7; ++k) r[k] = (S)s[line][k];
            }
        }
    }
    // Chroma: bS 2 on the 16-sample luma grid, two chroma lines.
    if (b == 2 && (x & 15) == 0) {
        for (int c = 1; c < 3; ++c) {
            int tcc = chroma_tc(P, qpp, qpq, c - 1);
            S* pl = (S*)P.cur[c];
            chroma_lines(pl + (size_t)(y >> 1) * P.pitch[c] + (x >> 1), 1, P.pitch[c], tcc, max);
        }
    }
}

template <typename S>
__device__ void deblock_h_seg(const HevcPicDev& P, int t) {
    const int w4 = P.w4;
    const int i = t % w4, e = 1 + t / w4;
    const int y = e * 8;
    if (y >= P.h) return;
    const int q = (2 * e) * w4 + i;
    const int b = (P.bs[q] >> 2) & 3;
    if (!b) return;
    const int x = i * 4, max = (1 << P.bd) - 1;
    const int qpp = P.qp[q - w4], qpq = P.qp[q];
    int beta, tc;
    luma_thresholds(P, b, qpp, qpq, &beta, &tc);
    if (beta != 0 || tc != 0) {
        S* Y = (S*)P.cur[0];
        const int pitch = P.pitch[0];
        int s[4][8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const S* r = Y + (size_t)(y - 4 + k) * pitch + x;
#pragma unroll
            for (int line = 0; line < 4; ++line) s[line][k] = r[line];
        }
        bool changed = false;
        luma_segment(s, beta, tc, max, changed);
        if (changed) {
#pragma unroll
            for (int k = 1; k < 7; ++k) {
                S* r = Y + (size_t)(y - 4 + k) * pitch + x;
#pragma unroll
                for (int line = 0; line < 4; ++line) r[line] = (S)s[line][k];
            }
        }
    }
    if (b == 2 && (y & 15) == 0) {
        for (int c = 1; c < 3; ++c) {
            int tcc = chroma_tc(P, qpp, qpq, c - 1);
            S* pl = (S*)P.cur[c];
            chroma_lines(pl + (size_t)(y >> 1) * P.pitch[c] + (x >> 1), P.pitch[c], 1, tcc, max);
        }
    }
}

__global__ void __launch_bounds__(256) hevc_stage2(BatchDev b) {
    const int32_t* pre = b.prefix + 2 * (b.n + 1);
    const int it = blockIdx.x;
    const int p = find_pic(pre, b.n, it);
    const HevcPicDev& P = *b.hevc[p];
    if (!P.deblock) return;
    int t = (it - pre[p]) * 256 + threadIdx.x;
    if (t >= (P.w >> 3) * P.h4) return;
    if (P.bd == 8) deblock_v_seg<uint8_t>(P, t);
    else deblock_v_seg<uint16_t>(P, t);
}

__global__ void __launch_bounds__(256) hevc_stage3(BatchDev b) {
    const int32_t* pre = b.prefix + 3 * (b.n + 1);
    const int it = blockIdx.x;
    const int p = find_pic(pre, b.n, it);
    const HevcPicDev& P = *b.hevc[p];
    if (!P.deblock) return;
    int t = (it - pre[p]) * 256 + threadIdx.x;
    if (t >= ((P.h >> 3) - 1) * P.w4 || P.h < 16) return;
    if (P.bd == 8) deblock_h_seg<uint8_t>(P, t);
    else deblock_h_seg<uint16_t>(P, t);
}

// ---------------------------------------------------------------------
// Stage 4: SAO (src/hevc/sao.rs:110-431; kernels src/dsp/hevc.rs:1181-1224)
// ---------------------------------------------------------------------
// Every output sample is written: CTBs without SAO are copied. With one
// slice and one tile, a neighbour is usable exactly when it lies inside
// the picture, so a sample is edge-filtered iff both of its class's
// neighbours are inside the picture (the interior + ring of sao_ctb).
template <typename S>
__device__ void sao_group(const HevcPicDev& P, int c, int gx, int y) {
    const int pw = P.pw[c], ph = P.ph[c], pitch = P.pitch[c];
    const S* src = (const S*)P.cur[c];
    S* dst = (S*)P.out[c];
    const int x0 = gx * 4;
    const int sh = c ? 1 : 0;
    const int rx = (x0 << sh) >> P.log2_ctb, ry = (y << sh) >> P.log2_ctb;
    rv_sao sp = P.sao[3 * (ry * P.wc + rx) + c];
    const bool on = c == 0 ? P.sao_luma : P.sao_chroma;
    const size_t row = (size_t)y * pitch;
    int v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = src[row + x0 + k];
    const int max = (1 << P.bd) - 1;
    int o[4] = {v[0], v[1], v[2], v[3]};
    if (on && sp.type_idx == 1) {
        const int shift = P.bd - 5;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int band = ((v[k] >> shift) - sp.band_or_class) & 31;
            if (band < 4) o[k] = clampi(v[k] + sp.offsets[band], 0, max);
        }
    } else if (on && sp.type_idx == 2) {
        const int cls = sp.band_or_class & 3;
        const int hx = cls == 1 ? 0 : (cls == 3 ? 1 : -1);  // first neighbour dx
        const int vy = cls == 0 ? 0 : -1;                   // first neighbour dy
        const int off5[5] = {sp.offsets[0], sp.offsets[1], 0, sp.offsets[2], sp.offsets[3]};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int x = x0 + k;
            int xa = x + hx, ya = y + vy, xb = x - hx, yb = y - vy;
            if (xa < 0 || xa >= pw || xb < 0 || xb >= pw || ya < 0 || ya >= ph || yb < 0 || yb >= ph) continue;
            int a = src[(size_t)ya * pitch + xa], bb = src[(size_t)yb * pitch + xb];
            int e = 2 + sgn(v[k] - a) + sgn(v[k] - bb);
            o[k] = clampi(v[k] + off5[e], 0, max);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) dst[row + x0 + k] = (S)o[k];
}

__global__ void __launch_bounds__(256) hevc_stage4(BatchDev b) {
    const int32_t* pre = b.prefix + 4 * (b.n + 1);
    const int it = blockIdx.x;
    const int p = find_pic(pre, b.n, it);
    const HevcPicDev& P = *b.hevc[p];
    int t = (it - pre[p]) * 256 + threadIdx.x;
    const int gy = P.pw[0] >> 2, gc = P.pw[1] >> 2;
    const int ny = gy * P.ph[0], nc = gc * P.ph[1];
    int c, gx, y;
    if (t < ny) {
        c = 0;
        y = t / gy;
        gx = t - y * gy;
    } else if (t < ny + 2 * nc) {
        t -= ny;
        c = 1 + (t >= nc);
        if (t >= nc) t -= nc;
        y = t / gc;
        gx = t - y * gc;
    } else {
        return;
    }
    if (P.bd == 8) sao_group<uint8_t>(P, c, gx, y);
    else sao_group<uint16_t>(P, c, gx, y);
}

}  // namespace

// Work items per stage for one picture (the grid unit of each stage).
void hevc_stage_items(HevcPicDev& P) {
    P.items[0] = P.n_cus + P.n_pus + P.n_ltus + P.n_tiles;
    P.items[1] = ((P.n_tx4 + 3) >> 2) + P.n_txbig + (P.w4 * P.h4 + 63) / 64;
    P.items[2] = ((P.w >> 3) * P.h4 + 255) / 256;
    P.items[3] = (P.h >= 16) ? ((((P.h >> 3) - 1) * P.w4) + 255) / 256 : 0;
    int ny = (P.pw[0] >> 2) * P.ph[0], nc = (P.pw[1] >> 2) * P.ph[1];
    P.items[4] = (ny + 2 * nc + 255) / 256;
}

hipError_t hevc_launch_stage(int stage, const BatchDev& b, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    switch (stage) {
        case 0: hipLaunchKernelGGL(hevc_stage0, dim3(total), dim3(64), 0, st, b); break;
        case 1: hipLaunchKernelGGL(hevc_stage1, dim3(total), dim3(64), 0, st, b); break;
        case 2: hipLaunchKernelGGL(hevc_stage2, dim3(total), dim3(256), 0, st, b); break;
        case 3: hipLaunchKernelGGL(hevc_stage3, dim3(total), dim3(256), 0, st, b); break;
        case 4: hipLaunchKernelGGL(hevc_stage4, dim3(total), dim3(256), 0, st, b); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace rv
