// H.264 picture reconstruction on gfx950 (8-bit, 4:2:0, frame pictures).
//
//   stage 0  intra prediction blocks, motion-compensated partitions,
//            per-4x4 partition map
//   stage 1  4x4 / 8x8 inverse transforms + residual add (one block per lane)
//   stage 2  boundary strengths and filter thresholds per macroblock edge
//   stage 3  deblocking as a macroblock wavefront inside one workgroup per
//            picture: MB (x, y) after MB (x - 1, y) and MB (x + 1, y - 1),
//            which gives the raster order's result (src/h264/deblock.rs:
//            617-813); carried samples stay in LDS
//   stage 4  (none: the picture is reconstructed in its DPB slot)
//
// Arithmetic restates src/dsp/h264.rs bit for bit (wrapping i16 transforms
// at 8 bits, :766-1066).
#include <hip/hip_runtime.h>

#include "rv_device.h"
#include "rv_kernels.h"

namespace rv {
namespace {

// Tables 8-16 / 8-17 of H.264 (src/h264/tables_gen.rs:331-392).
__constant__ uint8_t kAlpha[52] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,   0,   0,   0,   0,   0,   4,   4,
                                   5,  6,  7,  8,  9,  10, 12, 13, 15, 17, 20,  22,  25,  28,  32,  36,  40,  45,
                                   50, 56, 63, 71, 80, 90, 101, 113, 127, 144, 162, 182, 203, 226, 255, 255};
__constant__ uint8_t kBeta[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  0,  0,  0,  0,  2,  2,  2,  3,  3,  3,  3,  4,  4,  4,
                                  6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 18, 18};
__constant__ uint8_t kTc0[52][3] = {
    {0, 0, 0},  {0, 0, 0},  {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},    {0, 0, 0},
    {0, 0, 0},  {0, 0, 0},  {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},    {0, 0, 0},
    {0, 0, 0},  {0, 0, 1},  {0, 0, 1},   {0, 0, 1},   {0, 0, 1},   {0, 1, 1},   {0, 1, 1},    {1, 1, 1},
    {1, 1, 1},  {1, 1, 1},  {1, 1, 1},   {1, 1, 2},   {1, 1, 2},   {1, 1, 2},   {1, 1, 2},    {1, 2, 3},
    {1, 2, 3},  {2, 2, 3},  {2, 2, 4},   {2, 3, 4},   {2, 3, 4},   {3, 3, 5},   {3, 4, 6},    {3, 4, 6},
    {4, 5, 7},  {4, 5, 8},  {4, 6, 9},   {5, 7, 10},  {6, 8, 11},  {6, 8, 13},  {7, 10, 14},  {8, 11, 16},
    {9, 12, 18}, {10, 13, 20}, {11, 15, 23}, {13, 17, 25}};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

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
// Stage 0: intra blocks, motion compensation
// ---------------------------------------------------------------------

__device__ __forceinline__ int tap6(int a, int b, int c, int d, int e, int f) {
    return a - 5 * b + 20 * c + 20 * d - 5 * e + f;
}

// qpel_scalar<XF, YF> (src/dsp/h264.rs:427-493) over a window staged in
// LDS: w is the window (stride LS), origin two left / two above the block.
template <int LS> struct WinS {
    const int16_t* w;
    int mx = 255;  // sample maximum
    __device__ int G(int x, int y) const { return w[(y + 2) * LS + x + 2]; }
    __device__ int b1(int x, int yy) const {
        const int16_t* r = w + yy * LS + x;
        return tap6(r[0], r[1], r[2], r[3], r[4], r[5]);
    }
    __device__ int h1(int xx, int y) const {
        const int16_t* c = w + y * LS + xx;
        return tap6(c[0], c[LS], c[2 * LS], c[3 * LS], c[4 * LS], c[5 * LS]);
    }
    __device__ int b(int x, int y) const { return clampi((b1(x, y + 2) + 16) >> 5, 0, mx); }
    __device__ int h(int x, int y) const { return clampi((h1(x + 2, y) + 16) >> 5, 0, mx); }
    __device__ int j(int x, int y) const {
        int j1 = tap6(b1(x, y), b1(x, y + 1), b1(x, y + 2), b1(x, y + 3), b1(x, y + 4), b1(x, y + 5));
        return clampi((j1 + 512) >> 10, 0, mx);
    }
};
using Win = WinS<32>;

template <class Q> __device__ int qpel_at(const Q& q, int pos, int x, int y) {
    switch (pos) {
        case 0: return q.G(x, y);
        case 1: return (q.G(x, y) + q.b(x, y) + 1) >> 1;
        case 2: return q.b(x, y);
        case 3: return (q.G(x + 1, y) + q.b(x, y) + 1) >> 1;
        case 4: return (q.G(x, y) + q.h(x, y) + 1) >> 1;
        case 8: return q.h(x, y);
        case 12: return (q.G(x, y + 1) + q.h(x, y) + 1) >> 1;
        case 10: return q.j(x, y);
        case 5: return (q.b(x, y) + q.h(x, y) + 1) >> 1;
        case 7: return (q.b(x, y) + q.h(x + 1, y) + 1) >> 1;
        case 13: return (q.h(x, y) + q.b(x, y + 1) + 1) >> 1;
        case 15: return (q.h(x + 1, y) + q.b(x, y + 1) + 1) >> 1;
        case 6: return (q.b(x, y) + q.j(x, y) + 1) >> 1;
        case 14: return (q.j(x, y) + q.b(x, y + 1) + 1) >> 1;
        case 9: return (q.h(x, y) + q.j(x, y) + 1) >> 1;
        default: return (q.j(x, y) + q.h(x + 1, y) + 1) >> 1;  // 11: (3, 2)
    }
}

// The combine of predict_partition (src/h264/inter.rs:292-331) with
// copy / avg / weighted_uni / weighted_bi (src/dsp/h264.rs:522-597).
__device__ __forceinline__ int combine(const H264PicDev& P, int c, int dir, const int8_t* ri, int a, int b) {
    const bool both = dir == 3;
    const int single = dir == 1 ? 0 : 1;
    int log_wd, w0 = 1, w1 = 1, o0 = 0, o1 = 0;
    if (P.weighting == RV_H264_EXPLICIT) {
        log_wd = P.log_wd[c];
        if (dir & 1) {
            w0 = P.wp_w[0][ri[0]][c];
            o0 = P.wp_o[0][ri[0]][c];
        }
        if (dir & 2) {
            w1 = P.wp_w[1][ri[1]][c];
            o1 = P.wp_o[1][ri[1]][c];
        }
    } else if (P.weighting == RV_H264_IMPLICIT && both) {
        log_wd = 5;
        w0 = P.imp_w[ri[0]][ri[1]][0];
        w1 = P.imp_w[ri[0]][ri[1]][1];
    } else {
        return both ? (a + b + 1) >> 1 : (single == 0 ? a : b);
    }
    if (!both) {
        int v = single == 0 ? a : b, wt = single == 0 ? w0 : w1, o = single == 0 ? o0 : o1;
        if (log_wd >= 1) return clampi(((v * wt + (1 << (log_wd - 1))) >> log_wd) + o, 0, 255);
        return clampi(v * wt + o, 0, 255);
    }
    return clampi(((a * w0 + b * w1 + (1 << log_wd)) >> (log_wd + 1)) + ((o0 + o1 + 1) >> 1), 0, 255);
}

// qpel_at as terms: every position is one term, or the rounded mean of two,
// among G (full sample at column offset bit 4, row offset bit 5), b
// (horizontal half sample, row offset bit 6), h (vertical half sample,
// column offset bit 7) and j (centre half sample). One byte per position.
constexpr uint32_t kQG = 1, kQB = 2, kQH = 4, kQJ = 8;
__device__ __forceinline__ uint32_t qpel_terms(int pos) {
    // Positions 0..7: G, G+b, b, G(x+1)+b, G+h, b+h, b+j, b+h(x+1);
    // 8..15: h, h+j, j, j+h(x+1), G(y+1)+h, h+b(y+1), j+b(y+1), h(x+1)+b(y+1).
    constexpr uint64_t lo = 0x860A060513020301ull, hi = 0xC64A46258C080C04ull;
    return (uint32_t)(((pos & 8) ? hi : lo) >> (8 * (pos & 7))) & 0xFFu;
}

// The window of qpel_scalar for stage 0: the luma window as bytes (stride
// LS; w points at column 0, which sits `sh` bytes into its row) and the
// six-tap sums b1 of its rows at columns 0..w-1 (stride BS), computed once
// per partition.
template <int LS, int BS> struct WinB {
    const uint8_t* w;
    const int16_t* hb;
    int mx = 255;  // sample maximum
    __device__ int G(int x, int y) const { return w[(y + 2) * LS + x + 2]; }
    __device__ int b1(int x, int yy) const { return hb[yy * BS + x]; }
    __device__ int h1(int xx, int y) const {
        const uint8_t* c = w + y * LS + xx;
        return tap6(c[0], c[LS], c[2 * LS], c[3 * LS], c[4 * LS], c[5 * LS]);
    }
    __device__ int b(int x, int y) const { return clampi((b1(x, y + 2) + 16) >> 5, 0, mx); }
    __device__ int h(int x, int y) const { return clampi((h1(x + 2, y) + 16) >> 5, 0, mx); }
    __device__ int j(int x, int y) const {
        int j1 = tap6(b1(x, y), b1(x, y + 1), b1(x, y + 2), b1(x, y + 3), b1(x, y + 4), b1(x, y + 5));
        return clampi((j1 + 512) >> 10, 0, mx);
    }
};

// One partition on NT lanes: 64 for partitions up to 16 x 16, 16 for
// those up to 8 x 8 (four per wave). Its LDS (bytes) holds the luma windows
// of both lists (up to (w + 5) x (h + 5), stride LS), the chroma windows of
// both lists and planes (up to (w/2 + 1) x (h/2 + 1), stride CS), then the
// six-tap row sums of both luma windows (int16, (h + 5) x w, stride MW).
template <int NT> struct PartShape {
    static constexpr int MW = NT == 64 ? 16 : 8;     // largest width / height
    static constexpr int LS = NT == 64 ? 32 : 16;    // luma window stride
    static constexpr int LR = MW + 5;                // luma window rows
    static constexpr int CS = NT == 64 ? 16 : 8;     // chroma window stride
    static constexpr int CR = MW / 2 + 1;            // chroma window rows
    static constexpr int LWIN = LR * LS, CWIN = CR * CS;
    static constexpr int BWIN = LR * MW;             // int16 per list
    static constexpr int BOFF = 2 * LWIN + 4 * CWIN;  // bytes before the row sums
    static constexpr int LDS = BOFF + 4 * BWIN;       // bytes per partition
    static constexpr int LPASS = (LR * LS + NT - 1) / NT, CPASS = (CR * CS + NT - 1) / NT;
    static constexpr int ND = MW / 4 + 2;  // aligned dwords covering a luma window row
    static constexpr int DPASS = (LR * ND + NT - 1) / NT;
    static constexpr int QN = MW / 4;      // four-column groups of a row-sum row
    static constexpr int BPASS = (LR * QN + NT - 1) / NT;
    static_assert(BOFF % 8 == 0 && LDS % 8 == 0 && 4 * ND <= LS, "LDS layout");
};
constexpr int kStage0Lds = PartShape<64>::LDS > 4 * PartShape<16>::LDS ? PartShape<64>::LDS : 4 * PartShape<16>::LDS;

// Six-tap sums 1, -5, 20, 20, -5, 1 of four consecutive window columns
// c..c+3 with v_dot4_i32_i8: the samples x are biased to x - 128 (top bit
// flipped) to make them signed, and the taps sum to 32, so each sum starts
// at 128 * 32. d holds the window bytes from column c - sh on (sh 0..3) in
// byte order; the taps of bytes 6 and 7 of each 8-byte group are 0.
__device__ __forceinline__ void tap6x4(const uint32_t (&d)[4], uint32_t sh, int (&o)[4]) {
    constexpr uint32_t t0 = 0x1414FB01u, t1 = 0x000001FBu;
    const uint32_t x0 = d[0] ^ 0x80808080u, x1 = d[1] ^ 0x80808080u, x2 = d[2] ^ 0x80808080u,
                   x3 = d[3] ^ 0x80808080u;
    const uint32_t e0 = __builtin_amdgcn_alignbyte(x1, x0, sh), e1 = __builtin_amdgcn_alignbyte(x2, x1, sh),
                   e2 = __builtin_amdgcn_alignbyte(x3, x2, sh);
    o[0] = __builtin_amdgcn_sdot4((int)e1, (int)t1, __builtin_amdgcn_sdot4((int)e0, (int)t0, 4096, false), false);
#pragma unroll
    for (int j = 1; j < 4; ++j)
        o[j] = __builtin_amdgcn_sdot4((int)__builtin_amdgcn_alignbyte(e2, e1, j), (int)t1,
                                      __builtin_amdgcn_sdot4((int)__builtin_amdgcn_alignbyte(e1, e0, j), (int)t0, 4096, false),
                                      false);
}

// `valid` false: a lane group without a partition, which still meets the
// workgroup barriers.
template <int NT>
__device__ __forceinline__ void mc_part(const H264PicDev& P, int idx, int tid, uint8_t* lds, bool valid) {
    using S = PartShape<NT>;
    const rv_h264_part pt = P.parts[valid ? idx : 0];
    const int x = pt.x, y = pt.y, w = pt.w, h = pt.h, dir = valid ? pt.dir : 0;
    // Partition map for the deblocking filter.
    for (int e = tid; valid && e < (w >> 2) * (h >> 2); e += NT) {
        int bx = (x >> 2) + e % (w >> 2), by = (y >> 2) + e / (w >> 2);
        P.partmap[(size_t)by * P.w4 + bx] = (uint32_t)idx;
    }
    // Every window's samples are requested before any is stored, so the
    // workgroup waits for memory once. Windows clamp to the picture (the
    // replicated border of src/h264/frame.rs and the gather of
    // src/h264/inter.rs:127-139 read the same samples). A luma window whose
    // columns lie inside the plane is read and stored as aligned dwords of
    // clamped rows, column 0 at byte sh = x0 & 3 of its row; one crossing
    // the left or right edge sample by sample, column 0 at byte 0.
    const int ww = w + 5, hh = h + 5, cw = w >> 1, ch = h >> 1;
    int lv[2][S::LPASS], cv[2][2][S::CPASS];
    int shl[2] = {0, 0}, ndl[2] = {0, 0};
    bool ins[2] = {false, false};
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        if (!((dir >> l) & 1)) continue;
        const int mvx = pt.mv[l].x, mvy = pt.mv[l].y;
        const uint8_t* ref = (const uint8_t*)P.ref[l][pt.ref_idx[l]][0];
        const int x0 = x + (mvx >> 2) - 2, y0 = y + (mvy >> 2) - 2;
        const int sh = x0 & 3, nd = (sh + ww + 3) >> 2;
        ins[l] = x0 >= 0 && x0 - sh + 4 * nd <= P.pw[0];
        shl[l] = ins[l] ? sh : 0, ndl[l] = nd;
        if (ins[l]) {
#pragma unroll
            for (int k = 0; k < S::DPASS; ++k) {
                const int e = tid + NT * k, r = e / S::ND, d = e - r * S::ND;
                if (d < nd && r < hh)
                    lv[l][k] = (int)*(const uint32_t*)(ref + (size_t)clampi(y0 + r, 0, P.ph[0] - 1) * P.pitch[0] +
                                                       (x0 - sh) + 4 * d);
            }
        } else {
#pragma unroll
            for (int k = 0; k < S::LPASS; ++k) {
                const int e = tid + NT * k, r = e / S::LS, c = e % S::LS;
                if (c < ww && r < hh)
                    lv[l][k] = ref[(size_t)clampi(y0 + r, 0, P.ph[0] - 1) * P.pitch[0] + clampi(x0 + c, 0, P.pw[0] - 1)];
            }
        }
        const int xc0 = (x >> 1) + (mvx >> 3), yc0 = (y >> 1) + (mvy >> 3);
#pragma unroll
        for (int c2 = 0; c2 < 2; ++c2) {
            const uint8_t* rc = (const uint8_t*)P.ref[l][pt.ref_idx[l]][1 + c2];
#pragma unroll
            for (int k = 0; k < S::CPASS; ++k) {
                const int e = tid + NT * k, r = e / S::CS, c = e % S::CS;
                if (c <= cw && r <= ch)
                    cv[l][c2][k] = rc[(size_t)clampi(yc0 + r, 0, P.ph[1 + c2] - 1) * P.pitch[1 + c2] +
                                      clampi(xc0 + c, 0, P.pw[1 + c2] - 1)];
            }
        }
    }
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        if (!((dir >> l) & 1)) continue;
        uint8_t* lw = lds + l * S::LWIN;
        if (ins[l]) {
#pragma unroll
            for (int k = 0; k < S::DPASS; ++k) {
                const int e = tid + NT * k, r = e / S::ND, d = e - r * S::ND;
                if (d < ndl[l] && r < hh) *(uint32_t*)(lw + r * S::LS + 4 * d) = (uint32_t)lv[l][k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < S::LPASS; ++k) {
                const int e = tid + NT * k, r = e / S::LS, c = e % S::LS;
                if (c < ww && r < hh) lw[r * S::LS + c] = (uint8_t)lv[l][k];
            }
        }
#pragma unroll
        for (int c2 = 0; c2 < 2; ++c2)
#pragma unroll
            for (int k = 0; k < S::CPASS; ++k) {
                const int e = tid + NT * k, r = e / S::CS, c = e % S::CS;
                if (c <= cw && r <= ch) lds[2 * S::LWIN + (l * 2 + c2) * S::CWIN + r * S::CS + c] = (uint8_t)cv[l][c2][k];
            }
    }
    __syncthreads();
    // Row sums b1 of every window row at columns 0..w-1, for a list whose
    // horizontal phase is fractional (the positions that use b or j).
    int16_t* const hb = (int16_t*)(lds + S::BOFF);
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        if (!((dir >> l) & 1) || !(pt.mv[l].x & 3)) continue;
        const uint32_t* lw = (const uint32_t*)(lds + l * S::LWIN);
#pragma unroll
        for (int k = 0; k < S::BPASS; ++k) {
            const int e = tid + NT * k, r = e / S::QN, qd = e - r * S::QN;
            if (r < hh && 4 * qd < w) {
                // Window bytes 4 qd .. 4 qd + 15 of row r; the last dword may
                // lie in the next row (or the chroma windows after the last
                // row), under taps that are 0.
                const uint32_t* src = lw + r * (S::LS / 4) + qd;
                const uint32_t d[4] = {src[0], src[1], src[2], src[3]};
                int o[4];
                tap6x4(d, (uint32_t)shl[l], o);
                *(uint2*)(hb + l * S::BWIN + r * S::MW + 4 * qd) =
                    make_uint2(__builtin_amdgcn_perm((uint32_t)o[1], (uint32_t)o[0], 0x05040100u),
                               __builtin_amdgcn_perm((uint32_t)o[3], (uint32_t)o[2], 0x05040100u));
            }
        }
    }
    __syncthreads();
    if (!valid) return;
    // Lane tid takes samples o = tid + NT k (k < kn) of the partition in
    // raster order over its own width: w h / NT of them (exact for every
    // partition shape a group serves).
    const int lw = 31 - __builtin_clz((unsigned)w), kn = (w * h) / NT;
    int pred[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        if (!((dir >> l) & 1)) continue;
        const int pos = (pt.mv[l].y & 3) * 4 + (pt.mv[l].x & 3);
        const WinB<S::LS, S::MW> q{lds + l * S::LWIN + shl[l], hb + l * S::BWIN};
        // Each term once for the lane's samples; a term no lane of the wave
        // uses is skipped.
        const uint32_t t = qpel_terms(pos);
        int s[4] = {0, 0, 0, 0};
        if (t & kQG) {
            const int dx = (t >> 4) & 1, dy = (t >> 5) & 1;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < kn) s[k] += q.G(((tid + NT * k) & (w - 1)) + dx, ((tid + NT * k) >> lw) + dy);
        }
        if (t & kQB) {
            const int dy = (t >> 6) & 1;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < kn) s[k] += q.b((tid + NT * k) & (w - 1), ((tid + NT * k) >> lw) + dy);
        }
        if (t & kQH) {
            const int dx = (t >> 7) & 1;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < kn) s[k] += q.h(((tid + NT * k) & (w - 1)) + dx, (tid + NT * k) >> lw);
        }
        if (t & kQJ) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < kn) s[k] += q.j((tid + NT * k) & (w - 1), (tid + NT * k) >> lw);
        }
        const bool two = __builtin_popcount(t & 15u) == 2;
#pragma unroll
        for (int k = 0; k < 4; ++k) pred[l][k] = two ? (s[k] + 1) >> 1 : s[k];
    }
    uint8_t* dst = (uint8_t*)P.cur[0];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int o = tid + NT * k, oy = o >> lw, ox = o & (w - 1);
        if (k < kn) dst[(size_t)(y + oy) * P.pitch[0] + x + ox] = (uint8_t)combine(P, 0, dir, pt.ref_idx, pred[0][k], pred[1][k]);
    }
    // Chroma: eighth-sample bilinear (src/h264/inter.rs:243-282), 4:2:0.
    const int ox = tid % (S::MW / 2), oy = tid / (S::MW / 2);
    if (ox < cw && oy < ch) {
        for (int c = 1; c < 3; ++c) {
            int cp[2] = {0, 0};
#pragma unroll
            for (int l = 0; l < 2; ++l) {
                if (!((dir >> l) & 1)) continue;
                const int xf = pt.mv[l].x & 7, yf = pt.mv[l].y & 7;
                const int a = (8 - xf) * (8 - yf), bb = xf * (8 - yf), c2 = (8 - xf) * yf, d = xf * yf;
                const uint8_t* r0 = lds + 2 * S::LWIN + (l * 2 + c - 1) * S::CWIN + oy * S::CS + ox;
                cp[l] = (a * r0[0] + bb * r0[1] + c2 * r0[S::CS] + d * r0[S::CS + 1] + 32) >> 6;
            }
            uint8_t* dc = (uint8_t*)P.cur[c];
            dc[(size_t)((y >> 1) + oy) * P.pitch[c] + (x >> 1) + ox] = (uint8_t)combine(P, c, dir, pt.ref_idx, cp[0], cp[1]);
        }
    }
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) h264_stage0(BatchDev b) {
    __shared__ __attribute__((aligned(16))) uint8_t win[kStage0Lds];
    const int32_t* pre = b.prefix + 0 * (b.n + 1);
    const int it = blockIdx.x;
    const int p = find_pic(pre, b.n, it);
    const H264PicDev& P = *b.h264[p];
    int i = it - pre[p];
    const int tid = threadIdx.x;
    if (i < P.n_intra) {
        // Intra macroblock: no partition; host-supplied prediction copied
        // (device-intra macroblocks are predicted after the transform stage).
        const int a = (int)P.part_order[P.n_parts + i];
        const rv_h264_mb m = P.mbs[a];
        const int mx = (a % P.mbw) * 16, my = (a / P.mbw) * 16;
        if (tid < 16) P.partmap[(size_t)((my >> 2) + (tid >> 2)) * P.w4 + (mx >> 2) + (tid & 3)] = 0xFFFFFFFFu;
        if (m.intra != RV_INTRA_HOST) return;
        const uint16_t* pr = P.intra_pred + m.pred_off;
        for (int e = tid; e < 256; e += 64) {
            int yy = e >> 4, xx = e & 15;
            ((uint8_t*)P.cur[0])[(size_t)(my + yy) * P.pitch[0] + mx + xx] = (uint8_t)pr[e];
        }
        for (int c = 1; c < 3; ++c) {
            const uint16_t* pc = pr + 256 + (c - 1) * 64;
            int yy = tid >> 3, xx = tid & 7;
            ((uint8_t*)P.cur[c])[(size_t)((my >> 1) + yy) * P.pitch[c] + (mx >> 1) + xx] = (uint8_t)pc[tid];
        }
        return;
    }
    i -= P.n_intra;
    // Partitions above 8 x 8 one per workgroup, then the others four per
    // workgroup, 16 lanes each (the host orders them so).
    if (i < P.n_big) {
        mc_part<64>(P, P.part_order[i], tid, win, true);
        return;
    }
    const int g = tid >> 4, k = 4 * (i - P.n_big) + g;
    const bool valid = k < P.n_parts - P.n_big;
    mc_part<16>(P, valid ? P.part_order[P.n_big + k] : 0, tid & 15, win + g * PartShape<16>::LDS, valid);
}

// ---------------------------------------------------------------------
// Stage 1: residual, one block per lane (src/dsp/h264.rs:766-1066)
// ---------------------------------------------------------------------
__device__ __forceinline__ int16_t wa(int16_t a, int16_t b) { return (int16_t)(uint16_t)((uint16_t)a + (uint16_t)b); }
__device__ __forceinline__ int16_t ws(int16_t a, int16_t b) { return (int16_t)(uint16_t)((uint16_t)a - (uint16_t)b); }
__device__ __forceinline__ int16_t sh(int16_t a, int s) { return (int16_t)(a >> s); }

// add_rows_scalar (:803-816).
__device__ __forceinline__ void add_px(uint8_t* p, int16_t v, int max) {
    int16_t r = sh(wa(v, 32), 6);
    *p = (uint8_t)clampi((int)wa((int16_t)*p, r), 0, max);
}

// dc_add_scalar (:984-996), 8-bit: ((dc as i16) + 32) >> 6 in i16, added
// in i32.
__device__ void dc_add16(uint8_t* dst, long pitch, int n, int16_t dc, int max) {
    const int v = (int)sh(wa(dc, 32), 6);
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            uint8_t* p = dst + y * pitch + x;
            *p = (uint8_t)clampi(*p + v, 0, max);
        }
}

// idct4_add_scalar (:774-801) in wrapping i16.
__device__ void idct4_add16(uint8_t* dst, long pitch, const int16_t* d, int max) {
    int16_t t[16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int16_t d0 = d[i * 4], d1 = d[i * 4 + 1], d2 = d[i * 4 + 2], d3 = d[i * 4 + 3];
        int16_t e0 = wa(d0, d2), e1 = ws(d0, d2), e2 = ws(sh(d1, 1), d3), e3 = wa(d1, sh(d3, 1));
        t[i * 4] = wa(e0, e3);
        t[i * 4 + 1] = wa(e1, e2);
        t[i * 4 + 2] = ws(e1, e2);
        t[i * 4 + 3] = ws(e0, e3);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int16_t f0 = t[j], f1 = t[4 + j], f2 = t[8 + j], f3 = t[12 + j];
        int16_t g0 = wa(f0, f2), g1 = ws(f0, f2), g2 = ws(sh(f1, 1), f3), g3 = wa(f1, sh(f3, 1));
        add_px(dst + 0 * pitch + j, wa(g0, g3), max);
        add_px(dst + 1 * pitch + j, wa(g1, g2), max);
        add_px(dst + 2 * pitch + j, ws(g1, g2), max);
        add_px(dst + 3 * pitch + j, ws(g0, g3), max);
    }
}

// residual4_scalar with narrow4 and idct4_add_scalar / dc_add_scalar.
__device__ void residual4(uint8_t* dst, long pitch, const int32_t* co, int has_dc, int dc, int max = 255) {
    int16_t d[16];
    int ac = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int v = co[i];
        d[i] = (int16_t)clampi(v, -32768, 32767);
        if (i) ac |= v;
    }
    if (has_dc) d[0] = (int16_t)dc;
    if (ac == 0) {
        if (d[0] != 0) dc_add16(dst, pitch, 4, d[0], max);
        return;
    }
    idct4_add16(dst, pitch, d, max);
}

// idct8_1d_i16 (:913-958).
__device__ __forceinline__ void idct8_1d(const int16_t* d, int16_t* o) {
    int16_t a0 = wa(d[0], d[4]), a4 = ws(d[0], d[4]);
    int16_t a2 = ws(sh(d[2], 1), d[6]), a6 = wa(d[2], sh(d[6], 1));
    int16_t b0 = wa(a0, a6), b2 = wa(a4, a2), b4 = ws(a4, a2), b6 = ws(a0, a6);
    int16_t a1 = ws(ws(ws(d[5], d[3]), d[7]), sh(d[7], 1));
    int16_t a3 = ws(ws(wa(d[1], d[7]), d[3]), sh(d[3], 1));
    int16_t a5 = wa(wa(ws(d[7], d[1]), d[5]), sh(d[5], 1));
    int16_t a7 = wa(wa(wa(d[3], d[5]), d[1]), sh(d[1], 1));
    int16_t b1 = wa(a1, sh(a7, 2)), b7 = ws(a7, sh(a1, 2));
    int16_t b3 = wa(a3, sh(a5, 2)), b5 = ws(sh(a3, 2), a5);
    o[0] = wa(b0, b7);
    o[1] = wa(b2, b5);
    o[2] = wa(b4, b3);
    o[3] = wa(b6, b1);
    o[4] = ws(b6, b1);
    o[5] = ws(b4, b3);
    o[6] = ws(b2, b5);
    o[7] = ws(b0, b7);
}

// idct8_add_scalar (:961-982) in wrapping i16: rows, then columns.
__device__ void idct8_add16(uint8_t* dst, long pitch, const int16_t* d, int max) {
    int16_t t[64];
    for (int i = 0; i < 8; ++i) idct8_1d(d + i * 8, t + i * 8);
    for (int j = 0; j < 8; ++j) {
        int16_t col[8], o[8];
        for (int k = 0; k < 8; ++k) col[k] = t[k * 8 + j];
        idct8_1d(col, o);
        for (int i = 0; i < 8; ++i) add_px(dst + i * pitch + j, o[i], max);
    }
}

// residual8_scalar with narrow8 and idct8_add_scalar / dc_add_scalar.
__device__ void residual8(uint8_t* dst, long pitch, const int32_t* co, int max = 255) {
    int16_t d[64];
    int ac = 0;
    for (int i = 0; i < 64; ++i) {
        int v = co[i];
        d[i] = (int16_t)clampi(v, -32768, 32767);
        if (i) ac |= v;
    }
    if (ac == 0) {
        if (d[0] != 0) dc_add16(dst, pitch, 8, d[0], max);
        return;
    }
    idct8_add16(dst, pitch, d, max);
}

// Stage 1 reads and writes a block a row at a time: a 4-sample row is one
// dword, an 8-sample row two (block columns are multiples of the block
// size and plane rows 256-byte aligned), each byte updated as add_px /
// dc_add16 do; coefficients load as 16-byte vectors when so aligned.
__device__ __forceinline__ uint32_t add_px_word(uint32_t px, const int16_t* r, int max) {
    uint32_t o = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int16_t v = sh(wa(r[k], 32), 6);
        o |= (uint32_t)clampi((int)wa((int16_t)((px >> (8 * k)) & 0xFF), v), 0, max) << (8 * k);
    }
    return o;
}
__device__ __forceinline__ uint32_t dc_word(uint32_t px, int v, int max) {
    uint32_t o = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) o |= (uint32_t)clampi((int)((px >> (8 * k)) & 0xFF) + v, 0, max) << (8 * k);
    return o;
}
template <int N> __device__ __forceinline__ void load_coefs(const int32_t* co, int (&v)[N]) {
    if (((uintptr_t)co & 15) == 0) {
#pragma unroll
        for (int q = 0; q < N / 4; ++q) {
            const int4 t = ((const int4*)co)[q];
            v[4 * q] = t.x, v[4 * q + 1] = t.y, v[4 * q + 2] = t.z, v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = co[i];
    }
}

// residual4 with idct4_add16 / dc_add16, by rows.
__device__ __forceinline__ void residual4_rows(uint8_t* dst, long pitch, const int (&c)[16], int has_dc, int dc) {
    int16_t d[16];
    int ac = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        d[i] = (int16_t)clampi(c[i], -32768, 32767);
        if (i) ac |= c[i];
    }
    if (has_dc) d[0] = (int16_t)dc;
    if (ac == 0) {
        if (d[0] != 0) {
            const int v = (int)sh(wa(d[0], 32), 6);
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                uint32_t* r = (uint32_t*)(dst + y * pitch);
                *r = dc_word(*r, v, 255);
            }
        }
        return;
    }
    int16_t t[16], res[16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int16_t d0 = d[i * 4], d1 = d[i * 4 + 1], d2 = d[i * 4 + 2], d3 = d[i * 4 + 3];
        int16_t e0 = wa(d0, d2), e1 = ws(d0, d2), e2 = ws(sh(d1, 1), d3), e3 = wa(d1, sh(d3, 1));
        t[i * 4] = wa(e0, e3);
        t[i * 4 + 1] = wa(e1, e2);
        t[i * 4 + 2] = ws(e1, e2);
        t[i * 4 + 3] = ws(e0, e3);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int16_t f0 = t[j], f1 = t[4 + j], f2 = t[8 + j], f3 = t[12 + j];
        int16_t g0 = wa(f0, f2), g1 = ws(f0, f2), g2 = ws(sh(f1, 1), f3), g3 = wa(f1, sh(f3, 1));
        res[j] = wa(g0, g3);
        res[4 + j] = wa(g1, g2);
        res[8 + j] = ws(g1, g2);
        res[12 + j] = ws(g0, g3);
    }
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        uint32_t* r = (uint32_t*)(dst + y * pitch);
        *r = add_px_word(*r, res + 4 * y, 255);
    }
}

// residual8 with idct8_add16 / dc_add16, by rows.
__device__ __forceinline__ void residual8_rows(uint8_t* dst, long pitch, const int (&c)[64]) {
    int16_t d[64];
    int ac = 0;
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        d[i] = (int16_t)clampi(c[i], -32768, 32767);
        if (i) ac |= c[i];
    }
    if (ac == 0) {
        if (d[0] != 0) {
            const int v = (int)sh(wa(d[0], 32), 6);
#pragma unroll
            for (int y = 0; y < 8; ++y) {
                uint2* r = (uint2*)(dst + y * pitch);
                uint2 px = *r;
                px.x = dc_word(px.x, v, 255), px.y = dc_word(px.y, v, 255);
                *r = px;
            }
        }
        return;
    }
    int16_t t[64], res[64];
#pragma unroll
    for (int i = 0; i < 8; ++i) idct8_1d(d + i * 8, t + i * 8);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        int16_t col[8], o[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) col[k] = t[k * 8 + j];
        idct8_1d(col, o);
#pragma unroll
        for (int i = 0; i < 8; ++i) res[i * 8 + j] = o[i];
    }
#pragma unroll
    for (int y = 0; y < 8; ++y) {
        uint2* r = (uint2*)(dst + y * pitch);
        uint2 px = *r;
        px.x = add_px_word(px.x, res + 8 * y, 255), px.y = add_px_word(px.y, res + 8 * y + 4, 255);
        *r = px;
    }
}

// ---------------------------------------------------------------------
// Scaling of levels (8.5.9 - 8.5.13.1), as the reference's parser and
// recon apply it: LevelScale = weightScale * normAdjust (Dequant::new,
// src/h264/transform.rs:24-56), dequant_level (src/h264/mb.rs:1420-1428)
// with shift qP / 6 + 2 (4x4) or qP / 6 (8x8), and the DC transforms
// (transform.rs:179-228).
// ---------------------------------------------------------------------

// normAdjust4x4 / normAdjust8x8 values of one position class for the six
// values of qP % 6, one byte each (m = 0 lowest).
constexpr uint64_t pack6(int a, int b, int c, int d, int e, int f) {
    return (uint64_t)a | (uint64_t)b << 8 | (uint64_t)c << 16 | (uint64_t)d << 24 | (uint64_t)e << 32 |
           (uint64_t)f << 40;
}
// 4x4 classes: both indices even, both odd, mixed.
constexpr uint64_t kNorm4[3] = {pack6(10, 11, 13, 14, 16, 18), pack6(16, 18, 20, 23, 25, 29),
                                pack6(13, 14, 16, 18, 20, 23)};
// 8x8 classes: (i%4, j%4) both 0; i, j odd; both 2; a 0 and an odd; a 0
// and a 2; the rest.
constexpr uint64_t kNorm8[6] = {pack6(20, 22, 26, 28, 32, 36), pack6(18, 19, 23, 25, 28, 32),
                                pack6(32, 35, 42, 45, 51, 58), pack6(19, 21, 24, 26, 30, 34),
                                pack6(25, 28, 33, 35, 40, 46), pack6(24, 26, 31, 33, 38, 43)};
constexpr int norm4_class(int i, int j) { return (i % 2 == 0 && j % 2 == 0) ? 0 : (i % 2 == 1 && j % 2 == 1) ? 1 : 2; }
constexpr int norm8_class(int i, int j) {
    return (i % 4 == 0 && j % 4 == 0)                                  ? 0
           : (i % 2 == 1 && j % 2 == 1)                                ? 1
           : (i % 4 == 2 && j % 4 == 2)                                ? 2
           : ((i % 4 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 4 == 0)) ? 3
           : ((i % 4 == 0 && j % 4 == 2) || (i % 4 == 2 && j % 4 == 0)) ? 4
                                                                        : 5;
}
__device__ __forceinline__ int norm_of(uint64_t packed, int m) { return (int)((packed >> (8 * m)) & 0xFF); }

// (level * table << shift + 32) >> 6 in wrapping 32-bit arithmetic.
__device__ __forceinline__ int dequant_level(int level, int table, int shift) {
    const uint32_t t = ((uint32_t)level * (uint32_t)table) << shift;
    return (int)(t + 32u) >> 6;
}

// The weightScale bytes of a list as N / 4 words (all 16 when flat).
template <int N> __device__ __forceinline__ void load_weights(const uint8_t* list, uint32_t (&w)[N / 4]) {
    if (list) {
#pragma unroll
        for (int q = 0; q < N / 16; ++q) {
            const uint4 u = ((const uint4*)list)[q];
            w[4 * q] = u.x, w[4 * q + 1] = u.y, w[4 * q + 2] = u.z, w[4 * q + 3] = u.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < N / 4; ++q) w[q] = 0x10101010u;
    }
}

__device__ __forceinline__ void dequant4(int (&c)[16], const uint8_t* list, int qp) {
    const int m = qp % 6, shift = qp / 6 + 2;
    const int v0 = norm_of(kNorm4[0], m), v1 = norm_of(kNorm4[1], m), v2 = norm_of(kNorm4[2], m);
    uint32_t w[4];
    load_weights<16>(list, w);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int ws = (int)((w[i >> 2] >> (8 * (i & 3))) & 0xFF);
        const int k = norm4_class(i >> 2, i & 3);
        c[i] = dequant_level(c[i], ws * (k == 0 ? v0 : k == 1 ? v1 : v2), shift);
    }
}

__device__ __forceinline__ void dequant8(int (&c)[64], const uint8_t* list, int qp) {
    const int m = qp % 6, shift = qp / 6;
    const int v0 = norm_of(kNorm8[0], m), v1 = norm_of(kNorm8[1], m), v2 = norm_of(kNorm8[2], m);
    const int v3 = norm_of(kNorm8[3], m), v4 = norm_of(kNorm8[4], m), v5 = norm_of(kNorm8[5], m);
    uint32_t w[16];
    load_weights<64>(list, w);
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        const int ws = (int)((w[i >> 2] >> (8 * (i & 3))) & 0xFF);
        const int k = norm8_class(i >> 3, i & 7);
        const int v = k == 0 ? v0 : k == 1 ? v1 : k == 2 ? v2 : k == 3 ? v3 : k == 4 ? v4 : v5;
        c[i] = dequant_level(c[i], ws * v, shift);
    }
}

__device__ __forceinline__ int sel4(int k, int a, int b, int c, int d) { return k == 0 ? a : k == 1 ? b : k == 2 ? c : d; }

// Entry `idx` (raster) of luma_dc_transform over the 16 DC levels at g:
// each row through the Hadamard butterflies (outputs in the order c0+c1+
// c2+c3, c0+c1-c2-c3, c0-c1-c2+c3, c0-c1+c2-c3), column idx % 4 of the
// result through them again, entry idx / 4 of that; then the scaling at
// qP with LevelScale4x4(qP % 6, 0, 0) = scale00.
__device__ __forceinline__ int luma_dc(const int32_t* g, int idx, int scale00, int qp) {
    int col[4];
    const int j = idx & 3, i = idx >> 2;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int c0 = g[r * 4], c1 = g[r * 4 + 1], c2 = g[r * 4 + 2], c3 = g[r * 4 + 3];
        const int e0 = c0 + c1, e1 = c0 - c1, e2 = c2 - c3, e3 = c2 + c3;
        col[r] = sel4(j, e0 + e3, e0 - e3, e1 - e2, e1 + e2);
    }
    const int e0 = col[0] + col[1], e1 = col[0] - col[1], e2 = col[2] - col[3], e3 = col[2] + col[3];
    const int f = sel4(i, e0 + e3, e0 - e3, e1 - e2, e1 + e2);
    const int q6 = qp / 6;
    const uint32_t p = (uint32_t)f * (uint32_t)scale00;
    return qp >= 36 ? (int)(p << (q6 - 6)) : (int)(p + (1u << (5 - q6))) >> (6 - q6);
}

// Entry `idx` of chroma_dc_transform_420 over the 4 DC levels at g.
__device__ __forceinline__ int chroma_dc(const int32_t* g, int idx, int scale00, int qp) {
    const int c0 = g[0], c1 = g[1], c2 = g[2], c3 = g[3];
    const int f = sel4(idx, c0 + c1 + c2 + c3, c0 - c1 + c2 - c3, c0 + c1 - c2 - c3, c0 - c1 - c2 + c3);
    return (int)(((uint32_t)f * (uint32_t)scale00) << (qp / 6)) >> 5;
}

// The coefficients of block bk as its transform takes them, and its DC:
// scaled values as listed, or levels scaled with the macroblock's QP and
// intra / inter lists (MbDequant::for_mb, src/h264/mb.rs:1384-1414), the
// DC entry of the block's group through its DC transform.
template <int N>
__device__ __forceinline__ void block_input(const H264PicDev& P, const rv_h264_blk& bk, int (&c)[N * N], int& dc) {
    const int32_t* co = P.coefs + bk.coef_off;
    load_coefs<N * N>(co, c);
    dc = bk.dc;
    if (!(bk.flags & RV_BLK_LEVELS)) return;
    const int mb = bk.comp ? (bk.y >> 3) * P.mbw + (bk.x >> 3) : (bk.y >> 4) * P.mbw + (bk.x >> 4);
    const rv_h264_mb m = P.mbs[mb];
    const int qp = bk.comp == 0 ? m.qp : bk.comp == 1 ? m.qpc[0] : m.qpc[1];
    const int list = (m.intra ? 0 : 3) + bk.comp;
    const rv_h264_scaling* sl = P.scaling;
    if constexpr (N == 8) {
        dequant8(c, sl ? sl->list8[list == 0 ? 0 : 1] : nullptr, qp);
    } else {
        dequant4(c, sl ? sl->list4[list] : nullptr, qp);
        if (bk.has_dc) {
            // The DC of this block's position in its group, scaled with
            // the group's list: Intra_16x16 luma uses list 0 (recon.rs:
            // 701-706), chroma the macroblock's (recon.rs:1201-1211).
            const int32_t* g = P.coefs + bk.dc;
            const int dl = bk.comp ? list : 0;
            const int scale00 = (sl ? (int)sl->list4[dl][0] : 16) * norm_of(kNorm4[0], qp % 6);
            dc = bk.comp ? chroma_dc(g, ((bk.y >> 2) & 1) * 2 + ((bk.x >> 2) & 1), scale00, qp)
                         : luma_dc(g, ((bk.y >> 2) & 3) * 4 + ((bk.x >> 2) & 3), scale00, qp);
        }
    }
}

__global__ void __launch_bounds__(64) h264_stage1(BatchDev b) {
    const int32_t* pre = b.prefix + 1 * (b.n + 1);
    const int it = blockIdx.x;
    const int p = find_pic(pre, b.n, it);
    const H264PicDev& P = *b.h264[p];
    const int i = (it - pre[p]) * 64 + threadIdx.x;
    const int n4 = P.n_blks4;
    const int items4 = (n4 + 63) / 64;
    int k;
    if (i < items4 * 64) {
        k = i;
        if (k >= n4) return;
    } else {
        k = n4 + (i - items4 * 64);
        if (k >= n4 + P.n_blks8) return;
    }
    const rv_h264_blk bk = P.blks[k];
    uint8_t* dst = (uint8_t*)P.cur[bk.comp] + (size_t)bk.y * P.pitch[bk.comp] + bk.x;
    int dc;
    if (bk.size == 8) {
        int c[64];
        block_input<8>(P, bk, c, dc);
        residual8_rows(dst, P.pitch[bk.comp], c);
    } else {
        int c[16];
        block_input<4>(P, bk, c, dc);
        residual4_rows(dst, P.pitch[bk.comp], c, bk.has_dc, dc);
    }
}

// ---------------------------------------------------------------------
// Intra prediction of RV_INTRA_DEVICE macroblocks (clause 8.3; the
// predictors of src/h264/intra.rs, the decoding order and availability of
// src/h264/recon.rs:611-828, 936-1128), after stage 1, one launch per
// dependency level, one wave per macroblock. The wave first computes the
// residual of every block of the macroblock into LDS, then predicts luma
// block by block in decoding order in an LDS tile that also holds the
// neighbouring samples (each block's residual added before the next block
// reads it), then chroma; the macroblock is written out at the end.
// ---------------------------------------------------------------------

// The values a residual adds to the prediction (then clipped): the
// transform output r as add_rows_scalar rounds it, (r + 32) >> 6 in
// wrapping i16 (src/dsp/h264.rs:803-816), or the DC-only value
// (:984-996) on every sample.
__device__ __forceinline__ void residual4_vals(const int (&c)[16], int has_dc, int dc, int16_t (&r)[16]) {
    int16_t d[16];
    int ac = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        d[i] = (int16_t)clampi(c[i], -32768, 32767);
        if (i) ac |= c[i];
    }
    if (has_dc) d[0] = (int16_t)dc;
    if (ac == 0) {
        const int16_t v = sh(wa(d[0], 32), 6);
#pragma unroll
        for (int i = 0; i < 16; ++i) r[i] = v;
        return;
    }
    int16_t t[16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int16_t d0 = d[i * 4], d1 = d[i * 4 + 1], d2 = d[i * 4 + 2], d3 = d[i * 4 + 3];
        int16_t e0 = wa(d0, d2), e1 = ws(d0, d2), e2 = ws(sh(d1, 1), d3), e3 = wa(d1, sh(d3, 1));
        t[i * 4] = wa(e0, e3);
        t[i * 4 + 1] = wa(e1, e2);
        t[i * 4 + 2] = ws(e1, e2);
        t[i * 4 + 3] = ws(e0, e3);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int16_t f0 = t[j], f1 = t[4 + j], f2 = t[8 + j], f3 = t[12 + j];
        int16_t g0 = wa(f0, f2), g1 = ws(f0, f2), g2 = ws(sh(f1, 1), f3), g3 = wa(f1, sh(f3, 1));
        r[j] = sh(wa(wa(g0, g3), 32), 6);
        r[4 + j] = sh(wa(wa(g1, g2), 32), 6);
        r[8 + j] = sh(wa(ws(g1, g2), 32), 6);
        r[12 + j] = sh(wa(ws(g0, g3), 32), 6);
    }
}
__device__ __forceinline__ void residual8_vals(const int (&c)[64], int16_t (&r)[64]) {
    int16_t d[64];
    int ac = 0;
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        d[i] = (int16_t)clampi(c[i], -32768, 32767);
        if (i) ac |= c[i];
    }
    if (ac == 0) {
        const int16_t v = sh(wa(d[0], 32), 6);
#pragma unroll
        for (int i = 0; i < 64; ++i) r[i] = v;
        return;
    }
    int16_t t[64];
#pragma unroll
    for (int i = 0; i < 8; ++i) idct8_1d(d + i * 8, t + i * 8);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        int16_t col[8], o[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) col[k] = t[k * 8 + j];
        idct8_1d(col, o);
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i * 8 + j] = sh(wa(o[i], 32), 6);
    }
}

// LDS of one macroblock: residual values (luma 16x16, Cb, Cr 8x8, raster);
// the luma tile y[1 + r][1 + c] = sample (c, r) of the macroblock for
// r, c in -1..15, row 0 extending to c = 23 (the samples above-right);
// the chroma tiles likewise (c in -1..7); the references of the (one or
// two) blocks being predicted (t[0..15], l[0..7], corner); the luma modes.
struct alignas(16) H264IntraLds {
    int16_t res[384];
    uint8_t y[17][32];
    uint8_t c[2][9][16];
    int32_t edge[2][32];
    uint8_t modes[16];
};

__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

struct IAv {
    bool top, left, tl, tr;
};

// Sample (x, y) of an N x N directional mode (pred_4x4 / pred_8x8,
// src/h264/intra.rs:109-325, 427-620; the two sizes share one form) from
// the reference samples t[0..2N), l[0..N), c (unavailable ones zero, the
// top-right substituted).
__device__ __forceinline__ int nxn_sample(const int* t, const int* l, int c, int n, int mode, int x, int y, IAv av) {
    auto T = [&](int i) { return i < 0 ? c : t[i]; };
    auto Lf = [&](int i) { return i < 0 ? c : l[i]; };
    switch (mode) {
        case 0: return t[x];
        case 1: return l[y];
        case 2: {
            int st = 0, sl = 0;
            for (int k = 0; k < n; ++k) st += t[k], sl += l[k];
            const int lg = n == 4 ? 2 : 3;
            return av.top && av.left ? (st + sl + n) >> (lg + 1)
                   : av.top          ? (st + n / 2) >> lg
                   : av.left         ? (sl + n / 2) >> lg
                                     : 128;
        }
        case 3:
            return (x == n - 1 && y == n - 1) ? (t[2 * n - 2] + 3 * t[2 * n - 1] + 2) >> 2
                                              : (t[x + y] + 2 * t[x + y + 1] + t[x + y + 2] + 2) >> 2;
        case 4:
            if (x > y) return (T(x - y - 2) + 2 * T(x - y - 1) + T(x - y) + 2) >> 2;
            if (x < y) return (Lf(y - x - 2) + 2 * Lf(y - x - 1) + Lf(y - x) + 2) >> 2;
            return (T(0) + 2 * c + Lf(0) + 2) >> 2;
        case 5: {
            const int z = 2 * x - y;
            if (z >= 0 && !(z & 1)) return (T(x - (y >> 1) - 1) + T(x - (y >> 1)) + 1) >> 1;
            if (z >= 0) return (T(x - (y >> 1) - 2) + 2 * T(x - (y >> 1) - 1) + T(x - (y >> 1)) + 2) >> 2;
            if (z == -1) return (Lf(0) + 2 * c + T(0) + 2) >> 2;
            return (Lf(y - 2 * x - 1) + 2 * Lf(y - 2 * x - 2) + Lf(y - 2 * x - 3) + 2) >> 2;
        }
        case 6: {
            const int z = 2 * y - x;
            if (z >= 0 && !(z & 1)) return (Lf(y - (x >> 1) - 1) + Lf(y - (x >> 1)) + 1) >> 1;
            if (z >= 0) return (Lf(y - (x >> 1) - 2) + 2 * Lf(y - (x >> 1) - 1) + Lf(y - (x >> 1)) + 2) >> 2;
            if (z == -1) return (Lf(0) + 2 * c + T(0) + 2) >> 2;
            return (T(x - 2 * y - 1) + 2 * T(x - 2 * y - 2) + T(x - 2 * y - 3) + 2) >> 2;
        }
        case 7:
            return (y & 1) ? (t[x + (y >> 1)] + 2 * t[x + (y >> 1) + 1] + t[x + (y >> 1) + 2] + 2) >> 2
                           : (t[x + (y >> 1)] + t[x + (y >> 1) + 1] + 1) >> 1;
        default: {
            const int z = x + 2 * y;
            if (z > 2 * n - 3) return l[n - 1];
            if (z == 2 * n - 3) return (l[n - 2] + 3 * l[n - 1] + 2) >> 2;
            if (!(z & 1)) return (l[y + (x >> 1)] + l[y + (x >> 1) + 1] + 1) >> 1;
            return (l[y + (x >> 1)] + 2 * l[y + (x >> 1) + 1] + l[y + (x >> 1) + 2] + 2) >> 2;
        }
    }
}

// Availability of the block at 4x4 raster (bx, by), size 4 or 8, from the
// macroblock's neighbour bits (intra_avail_4x4 / _8x8, src/h264/recon.rs:
// 936-1011): top-right inside the macroblock only for blocks decoded
// after the block above-right.
__device__ __forceinline__ IAv block_av(uint32_t nb, int bx, int by, int size) {
    const bool A = nb & 1, B = nb & 2, Cc = nb & 4, D = nb & 8;
    IAv av;
    av.left = bx > 0 || A;
    av.top = by > 0 || B;
    av.tl = (bx > 0 && by > 0) || (bx == 0 ? (by == 0 ? D : A) : B);
    if (by == 0) av.tr = bx + size / 4 < 4 ? B : Cc;
    else if (size == 8) av.tr = bx == 0;
    else {
        // 4x4: the block above-right is decoded earlier unless it lies in
        // the 8x8 quadrant to the right (raster (1, 1), (3, 1), (1, 3),
        // (3, 2), (3, 3): luma4x4BlkIdx 3, 7, 11, 13, 15).
        const int o = ((by >> 1) * 2 + (bx >> 1)) * 4 + (by & 1) * 2 + (bx & 1);
        const int bx2 = bx + 1, by2 = by - 1;
        const int o2 = ((by2 >> 1) * 2 + (bx2 >> 1)) * 4 + (by2 & 1) * 2 + (bx2 & 1);
        av.tr = bx < 3 && o2 < o;
    }
    return av;
}

__global__ void __launch_bounds__(64) h264_intra_level(BatchDev b, int level) {
    __shared__ H264IntraLds L;
    const int32_t* pre = b.ipre + (size_t)(level + 1) * (b.n + 1);
    const int it = blockIdx.x;
    const int p = find_pic(pre, b.n, it);
    const H264PicDev& P = *b.h264[p];
    const H264IntraMb im = P.imbs[P.ilevel[level] + (it - pre[p])];
    const rv_h264_mb m = P.mbs[im.addr];
    const int tid = threadIdx.x;
    const int mx = (im.addr % P.mbw) * 16, my = (im.addr / P.mbw) * 16;
    const uint32_t nb = im.avail;
    const bool A = nb & 1, B = nb & 2, Cc = nb & 4, D = nb & 8;
    uint8_t* Yp = (uint8_t*)P.cur[0];
    const int py = P.pitch[0], pc = P.pitch[1];
    // 1. Residual values cleared; the neighbouring samples read; the modes
    // (read per block by a position that is not a constant) into LDS.
    for (int i = tid; i < 192; i += 64) ((uint32_t*)L.res)[i] = 0u;
    if (tid >= 48) L.modes[tid - 48] = m.modes[tid - 48];
    if (tid < 25) {
        const int x = tid - 1;
        const bool ok = x < 0 ? D : x < 16 ? B : Cc;
        L.y[0][tid] = ok ? Yp[(size_t)(my - 1) * py + mx + x] : 0;
    } else if (tid >= 32 && tid < 48) {
        const int y = tid - 32;
        L.y[1 + y][0] = A ? Yp[(size_t)(my + y) * py + mx - 1] : 0;
    }
    {
        const int comp = tid >> 5, k = tid & 31;
        const uint8_t* Cp = (const uint8_t*)P.cur[1 + comp];
        const int cx = mx >> 1, cy = my >> 1;
        if (k < 9) {
            const int x = k - 1;
            const bool ok = x < 0 ? D : B;
            L.c[comp][0][k] = ok ? Cp[(size_t)(cy - 1) * pc + cx + x] : 0;
        } else if (k < 17) {
            const int y = k - 9;
            L.c[comp][1 + y][0] = A ? Cp[(size_t)(cy + y) * pc + cx - 1] : 0;
        }
    }
    lds_sync();
    // 2. The residual of every listed block of the macroblock.
    if (tid < im.nblk) {
        const rv_h264_blk bk = P.iblks[im.blk0 + tid];
        int16_t* out;
        int ostride;
        if (bk.comp == 0) {
            out = L.res + (bk.y - my) * 16 + (bk.x - mx);
            ostride = 16;
        } else {
            out = L.res + 256 + (bk.comp - 1) * 64 + (bk.y - (my >> 1)) * 8 + (bk.x - (mx >> 1));
            ostride = 8;
        }
        int dc;
        if (bk.size == 8) {
            int c[64];
            int16_t r[64];
            block_input<8>(P, bk, c, dc);
            residual8_vals(c, r);
            for (int y = 0; y < 8; ++y)
                for (int x = 0; x < 8; ++x) out[y * ostride + x] = r[y * 8 + x];
        } else {
            int c[16];
            int16_t r[16];
            block_input<4>(P, bk, c, dc);
            residual4_vals(c, bk.has_dc, dc, r);
            for (int y = 0; y < 4; ++y)
                for (int x = 0; x < 4; ++x) out[y * ostride + x] = r[y * 4 + x];
        }
    }
    lds_sync();
    // 3. Luma.
    const int kind = m.intra_kind;
    if (kind == RV_H264_I16X16) {
        // predict_16x16 (src/h264/intra.rs:622-646, 792-923); av top B,
        // left A, top-left D.
        const int mode = m.modes[0];
        const int y = tid >> 2, x0 = (tid & 3) * 4;
        int v[4];
        if (mode == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = L.y[0][1 + x0 + k];
        } else if (mode == 1) {
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = L.y[1 + y][0];
        } else if (mode == 2) {
            int st = 0, sl = 0;
            for (int k = 0; k < 16; ++k) st += L.y[0][1 + k], sl += L.y[1 + k][0];
            const int dcv = A && B ? (st + sl + 16) >> 5 : B ? (st + 8) >> 4 : A ? (sl + 8) >> 4 : 128;
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = dcv;
        } else {
            int hh = 0, vv = 0;
            for (int i = 0; i < 8; ++i) {
                hh += (i + 1) * ((int)L.y[0][1 + 8 + i] - (int)L.y[0][1 + 6 - i]);
                vv += (i + 1) * ((int)L.y[1 + 8 + i][0] - (int)L.y[1 + 6 - i][0]);
            }
            const int bb = (5 * hh + 32) >> 6, cc = (5 * vv + 32) >> 6, a = 16 * (L.y[16][0] + L.y[0][16]);
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = clampi((a + bb * (x0 + k - 7) + cc * (y - 7) + 16) >> 5, 0, 255);
        }
        uint32_t w = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) w |= (uint32_t)clampi(v[k] + L.res[y * 16 + x0 + k], 0, 255) << (8 * k);
        *(uint32_t*)(Yp + (size_t)(my + y) * py + mx + x0) = w;
    } else {
        // Blocks by step: Intra_8x8 one block per step in decoding order;
        // Intra_4x4 block (bx, by) at step bx + 2 by, which follows every
        // block it reads (left, above, above-left, above-right), so the one
        // or two blocks of a step are predicted together (lane groups g).
        const bool n8 = kind == RV_H264_I8X8;
        const int n = n8 ? 8 : 4, steps = n8 ? 4 : 10;
        for (int st = 0; st < steps; ++st) {
            auto block_of = [&](int g, int& bx, int& by) {
                if (n8) {
                    bx = (st & 1) * 2;
                    by = (st >> 1) * 2;
                    return g == 0;
                }
                by = (st >= 2 ? (st - 2) / 2 : 0) + g;
                bx = st - 2 * by;
                return by <= min(3, st >> 1);
            };
            // Reference samples of group g's block into L.edge[g] (t[0..2n),
            // l[0..n) from 16, c at 24) by lanes 32 g .. 32 g + 24:
            // unfiltered for 4x4, filtered (8.3.2.2.1) for 8x8.
            {
                const int g = tid >> 5, e = tid & 31;
                int bx, by;
                if (e < 25 && block_of(g, bx, by)) {
                    const IAv av = block_av(nb, bx, by, n);
                    const int r0 = by * 4, c0 = bx * 4;  // tile row above / column left
                    int o = 0;
                    if (n8) {
                        auto top = [&](int i) {  // p[i, -1] with the substitution
                            return (int)L.y[r0][1 + c0 + (i < 8 || av.tr ? i : 7)];
                        };
                        const int cr = L.y[r0][c0];
                        if (e < 16) {
                            const int i = e;
                            if (av.top) {
                                if (i == 0) o = av.tl ? (cr + 2 * top(0) + top(1) + 2) >> 2 : (3 * top(0) + top(1) + 2) >> 2;
                                else if (i == 15) o = (top(14) + 3 * top(15) + 2) >> 2;
                                else o = (top(i - 1) + 2 * top(i) + top(i + 1) + 2) >> 2;
                            }
                        } else if (e < 24) {
                            const int j = e - 16;
                            auto lf = [&](int q) { return (int)L.y[r0 + 1 + q][c0]; };
                            if (av.left) {
                                if (j == 0) o = av.tl ? (cr + 2 * lf(0) + lf(1) + 2) >> 2 : (3 * lf(0) + lf(1) + 2) >> 2;
                                else if (j == 7) o = (lf(6) + 3 * lf(7) + 2) >> 2;
                                else o = (lf(j - 1) + 2 * lf(j) + lf(j + 1) + 2) >> 2;
                            }
                        } else if (av.tl) {
                            const int t0 = top(0), l0 = L.y[r0 + 1][c0];
                            o = av.top && av.left ? (t0 + 2 * cr + l0 + 2) >> 2
                                : av.top          ? (3 * cr + t0 + 2) >> 2
                                : av.left         ? (3 * cr + l0 + 2) >> 2
                                                  : cr;
                        }
                    } else {
                        // 4x4: t[0..8) (the top-right substituted), l[0..4), c.
                        if (e < 8) o = av.top ? (int)L.y[r0][1 + c0 + (e < 4 || av.tr ? e : 3)] : 0;
                        else if (e >= 16 && e < 20) o = av.left ? (int)L.y[r0 + 1 + e - 16][c0] : 0;
                        else if (e == 24) o = av.tl ? (int)L.y[r0][c0] : 0;
                    }
                    L.edge[g][e] = o;
                }
            }
            lds_sync();
            // Samples: lanes 16 g .. 16 g + 15 (4x4), all 64 (8x8). The lanes
            // read the block's references from LDS (indexed by position:
            // registers would go to scratch).
            {
                const int g = n8 ? 0 : tid >> 4, e = n8 ? tid : tid & 15;
                int bx, by;
                if ((n8 || tid < 32) && block_of(g, bx, by)) {
                    const IAv av = block_av(nb, bx, by, n);
                    const int r0 = by * 4, c0 = bx * 4;
                    const int mode = L.modes[by * 4 + bx];
                    const int x = e % n, y = e / n;
                    const int v = nxn_sample(L.edge[g], L.edge[g] + 16, L.edge[g][24], n, mode, x, y, av);
                    L.y[r0 + 1 + y][c0 + 1 + x] = (uint8_t)clampi(v + L.res[(r0 + y) * 16 + c0 + x], 0, 255);
                }
            }
            lds_sync();
        }
        // The finished macroblock.
        const int y = tid >> 2, x0 = (tid & 3) * 4;
        uint32_t w = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) w |= (uint32_t)L.y[1 + y][1 + x0 + k] << (8 * k);
        *(uint32_t*)(Yp + (size_t)(my + y) * py + mx + x0) = w;
    }
    // 4. Chroma (predict_chroma, src/h264/intra.rs:648-790, 4:2:0): lanes
    // 0..31 Cb, 32..63 Cr, two samples each.
    {
        const int comp = tid >> 5, k = tid & 31, y = k >> 2, x0 = (k & 3) * 2;
        const uint8_t(&T)[9][16] = L.c[comp];
        const int mode = m.chroma_mode;
        int v[2];
        if (mode == 0) {
            const int bx = x0 >> 2, by = y >> 2;
            int st = 0, sl = 0;
            for (int q = 0; q < 4; ++q) st += T[0][1 + bx * 4 + q], sl += T[1 + by * 4 + q][0];
            int dcv;
            if (bx == by) dcv = A && B ? (st + sl + 4) >> 3 : B ? (st + 2) >> 2 : A ? (sl + 2) >> 2 : 128;
            else if (bx) dcv = B ? (st + 2) >> 2 : A ? (sl + 2) >> 2 : 128;
            else dcv = A ? (sl + 2) >> 2 : B ? (st + 2) >> 2 : 128;
            v[0] = v[1] = dcv;
        } else if (mode == 1) {
            v[0] = v[1] = T[1 + y][0];
        } else if (mode == 2) {
            v[0] = T[0][1 + x0];
            v[1] = T[0][2 + x0];
        } else {
            int hh = 0, vv = 0;
            for (int i = 0; i < 4; ++i) {
                hh += (i + 1) * ((int)T[0][1 + 4 + i] - (int)T[0][1 + 2 - i]);
                vv += (i + 1) * ((int)T[1 + 4 + i][0] - (int)T[1 + 2 - i][0]);
            }
            const int bb = (34 * hh + 32) >> 6, cc = (34 * vv + 32) >> 6, a = 16 * (T[8][0] + T[0][8]);
            v[0] = clampi((a + bb * (x0 - 3) + cc * (y - 3) + 16) >> 5, 0, 255);
            v[1] = clampi((a + bb * (x0 + 1 - 3) + cc * (y - 3) + 16) >> 5, 0, 255);
        }
        const int16_t* rs = L.res + 256 + comp * 64 + y * 8 + x0;
        const uint32_t w = (uint32_t)clampi(v[0] + rs[0], 0, 255) | ((uint32_t)clampi(v[1] + rs[1], 0, 255) << 8);
        *(uint16_t*)((uint8_t*)P.cur[1 + comp] + (size_t)((my >> 1) + y) * pc + (mx >> 1) + x0) = (uint16_t)w;
    }
}

// ---------------------------------------------------------------------
// Stage 2: boundary strengths (src/h264/deblock.rs:97-146, 255-282, 640-766)
// ---------------------------------------------------------------------
struct Mot {
    int n;
    int ref[2], mv[2][2];
};

__device__ __forceinline__ Mot mot_of(const H264PicDev& P, int bx, int by) {
    uint32_t pi = P.partmap[(size_t)by * P.w4 + bx];
    Mot m;
    m.n = 0;
    m.ref[0] = m.ref[1] = -1;
    m.mv[0][0] = m.mv[0][1] = m.mv[1][0] = m.mv[1][1] = 0;
    if (pi >= (uint32_t)P.n_parts) return m;
    const rv_h264_part pt = P.parts[pi];
    for (int l = 0; l < 2; ++l)
        if ((pt.dir >> l) & 1) {
            m.ref[l] = P.ref_id[l][pt.ref_idx[l]];
            m.mv[l][0] = pt.mv[l].x;
            m.mv[l][1] = pt.mv[l].y;
            ++m.n;
        }
    return m;
}

__device__ __forceinline__ bool far4(const int* a, const int* b) {
    return abs(a[0] - b[0]) >= 4 || abs(a[1] - b[1]) >= 4;
}

// motion_bs (:97-146), frame macroblocks.
__device__ __forceinline__ int motion_bs(const Mot& p, const Mot& q) {
    if (p.n != q.n) return 1;
    if (p.n == 1) {
        // The one list each uses, by selects (an index that is not a
        // constant would put both records in scratch memory).
        const bool lp = p.ref[0] < 0, lq = q.ref[0] < 0;
        if ((lp ? p.ref[1] : p.ref[0]) != (lq ? q.ref[1] : q.ref[0])) return 1;
        const int a[2] = {lp ? p.mv[1][0] : p.mv[0][0], lp ? p.mv[1][1] : p.mv[0][1]};
        const int b[2] = {lq ? q.mv[1][0] : q.mv[0][0], lq ? q.mv[1][1] : q.mv[0][1]};
        return far4(a, b);
    }
    bool straight = p.ref[0] == q.ref[0] && p.ref[1] == q.ref[1];
    bool crossed = p.ref[0] == q.ref[1] && p.ref[1] == q.ref[0];
    if (!straight && !crossed) return 1;
    if (p.ref[0] == p.ref[1]) {
        bool a = far4(p.mv[0], q.mv[0]) || far4(p.mv[1], q.mv[1]);
        bool b = far4(p.mv[0], q.mv[1]) || far4(p.mv[1], q.mv[0]);
        return a && b;
    }
    if (straight) return far4(p.mv[0], q.mv[0]) || far4(p.mv[1], q.mv[1]);
    return far4(p.mv[0], q.mv[1]) || far4(p.mv[1], q.mv[0]);
}

__device__ __forceinline__ uint32_t partof(const H264PicDev& P, int bx, int by) {
    return P.partmap[(size_t)by * P.w4 + bx];
}

// The parameters of one edge with packed strengths `bs` and QP average `qa`
// (ThrTable::get, src/h264/deblock.rs:311-344; filter_edge, :192-213):
// x = alpha | beta << 8 | bS-4 flag << 16; y = one byte per 4-line
// segment, 0 where the segment is left alone, else tC0 + 1 (1 under bS 4,
// which byte 0 of the edge decides for all four).
__device__ uint2 edge_words(const H264PicDev& P, uint32_t bs, int qa) {
    if (bs == 0) return make_uint2(0, 0);
    const int ia = clampi(qa + P.offset_a, 0, 51), ib = clampi(qa + P.offset_b, 0, 51);
    const bool bs4 = (bs & 0xFF) == 4;
    uint32_t segs = 0;
    for (int k = 0; k < 4; ++k) {
        const int s = (bs >> (8 * k)) & 3;
        const uint32_t v = bs4 ? 1u : s ? (uint32_t)kTc0[ia][s - 1] + 1 : 0u;
        segs |= v << (8 * k);
    }
    return make_uint2((uint32_t)kAlpha[ia] | (uint32_t)kBeta[ib] << 8 | (bs4 ? 1u << 16 : 0u), segs);
}

// Packed strengths (one byte per 4-line segment) of edge e (0..3) of
// macroblock addr, vertical or horizontal: bS 4 on an intra macroblock's
// macroblock edges and 3 inside it; else 2 where a side has coefficients,
// else motion, only across a partition boundary inside the macroblock.
__device__ uint32_t edge_bs(const H264PicDev& P, int addr, const rv_h264_mb& m, bool vert, int e) {
    const int mbx = addr % P.mbw, mby = addr / P.mbw;
    const int bx0 = mbx * 4, by0 = mby * 4;
    const bool has = vert ? mbx > 0 : mby > 0;  // the macroblock edge exists
    if (m.intra) return e == 0 ? (has ? 0x04040404u : 0u) : 0x03030303u;
    const uint32_t nz = m.nz_mask;
    uint32_t w = 0;
    if (e > 0) {
        if (m.transform_8x8 && (e & 1)) return 0u;
        for (int k = 0; k < 4; ++k) {
            // Vertical: block row k across columns e - 1 | e; horizontal: the transpose.
            const int pb = vert ? k * 4 + e - 1 : (e - 1) * 4 + k, qb = vert ? k * 4 + e : e * 4 + k;
            const int px = vert ? bx0 + e - 1 : bx0 + k, py = vert ? by0 + k : by0 + e - 1;
            const int qx = vert ? bx0 + e : bx0 + k, qy = vert ? by0 + k : by0 + e;
            uint32_t v;
            if (((nz >> pb) | (nz >> qb)) & 1) v = 2;
            else if (partof(P, px, py) != partof(P, qx, qy)) v = motion_bs(mot_of(P, px, py), mot_of(P, qx, qy));
            else v = 0;
            w |= v << (8 * k);
        }
        return w;
    }
    if (!has) return 0u;
    const rv_h264_mb n = P.mbs[vert ? addr - 1 : addr - P.mbw];
    if (n.intra) return 0x04040404u;
    for (int k = 0; k < 4; ++k) {
        uint32_t v;
        const bool coef = vert ? (((n.nz_mask >> (k * 4 + 3)) | (nz >> (k * 4))) & 1)
                               : (((n.nz_mask >> (12 + k)) | (nz >> k)) & 1);
        if (coef) v = 2;
        else if (vert) v = motion_bs(mot_of(P, bx0 - 1, by0 + k), mot_of(P, bx0, by0 + k));
        else v = motion_bs(mot_of(P, bx0 + k, by0 - 1), mot_of(P, bx0 + k, by0));
        w |= v << (8 * k);
    }
    return w;
}

// One edge's parameter words (layout of kH264DbkWords): its luma words and,
// for luma edges 0 and 2 (chroma edges 0 and 4), both chroma planes'.
__device__ void edge_params(const H264PicDev& P, int addr, bool vert, int e) {
    const rv_h264_mb m = P.mbs[addr];
    const uint32_t bs = edge_bs(P, addr, m, vert, e);
    const int mbx = addr % P.mbw, mby = addr / P.mbw;
    const bool has = vert ? mbx > 0 : mby > 0;
    const rv_h264_mb n = P.mbs[has ? (vert ? addr - 1 : addr - P.mbw) : addr];  // QP neighbour of edge 0
    uint32_t* w = P.dbk + (size_t)addr * kH264DbkWords + (vert ? 0 : 16);
    const bool on = e == 0 || (e & 1) == 0 || !m.transform_8x8;  // 8x8 transform: no luma edges 1, 3
    const uint2 lw = edge_words(P, on ? bs : 0u, e == 0 ? (n.qp + m.qp + 1) >> 1 : m.qp);
    for (int k = 0; k < 4; ++k) w[k * 4 + e] = lw.x | ((lw.y >> (8 * k)) & 0xFF) << 24;
    if (e & 1) return;
    for (int c = 0; c < 2; ++c) {
        const uint2 cw = edge_words(P, bs, e == 0 ? (n.qpc[c] + m.qpc[c] + 1) >> 1 : m.qpc[c]);
        for (int k = 0; k < 4; ++k) w[32 + c * 8 + k * 2 + (e >> 1)] = cw.x | ((cw.y >> (8 * k)) & 0xFF) << 24;
    }
}

__global__ void __launch_bounds__(64) h264_stage2(BatchDev b) {
    const int32_t* pre = b.prefix + 2 * (b.n + 1);
    const int it = blockIdx.x;
    const int p = find_pic(pre, b.n, it);
    const H264PicDev& P = *b.h264[p];
    const int i = (it - pre[p]) * 64 + threadIdx.x;
    // The deblocking exchange and band ticket start at zero every launch.
    for (size_t k = i; k < (size_t)P.n_xch; k += (size_t)P.items[2] * 64) P.xch[k] = 0;
    // One thread per macroblock edge: bits 0..1 the edge, bit 2 horizontal.
    if (i < P.n_mbs * 8) edge_params(P, i >> 3, !((i >> 2) & 1), i & 3);
}

// ---------------------------------------------------------------------
// Stage 3: deblocking wavefront (src/h264/deblock.rs:567-815)
// ---------------------------------------------------------------------

// deblock_line (src/dsp/h264.rs:603-656) on one line, in place through
// `d` (q0) with `across` crossing the edge.
__device__ void line_filter(uint8_t* d, long across, bool bs4, int tc0, int alpha, int beta, bool chroma,
                            int max = 255) {
    int p0 = d[-across], p1 = d[-2 * across], q0 = d[0], q1 = d[across];
    if (!(abs(p0 - q0) < alpha && abs(p1 - p0) < beta && abs(q1 - q0) < beta)) return;
    if (chroma) {
        if (!bs4) {
            int tc = tc0 + 1;
            int delta = clampi((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
            d[-across] = (uint8_t)clampi(p0 + delta, 0, max);
            d[0] = (uint8_t)clampi(q0 - delta, 0, max);
        } else {
            d[-across] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
            d[0] = (uint8_t)((2 * q1 + q0 + p1 + 2) >> 2);
        }
        return;
    }
    int p2 = d[-3 * across], q2 = d[2 * across];
    int ap = abs(p2 - p0), aq = abs(q2 - q0);
    if (!bs4) {
        int tc = tc0 + (ap < beta) + (aq < beta);
        int delta = clampi((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
        d[-across] = (uint8_t)clampi(p0 + delta, 0, max);
        d[0] = (uint8_t)clampi(q0 - delta, 0, max);
        if (ap < beta) d[-2 * across] = (uint8_t)(p1 + clampi((p2 + ((p0 + q0 + 1) >> 1) - (p1 << 1)) >> 1, -tc0, tc0));
        if (aq < beta) d[across] = (uint8_t)(q1 + clampi((q2 + ((p0 + q0 + 1) >> 1) - (q1 << 1)) >> 1, -tc0, tc0));
    } else {
        int p3 = d[-4 * across], q3 = d[3 * across];
        bool strong = abs(p0 - q0) < ((alpha >> 2) + 2);
        if (ap < beta && strong) {
            d[-across] = (uint8_t)((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3);
            d[-2 * across] = (uint8_t)((p2 + p1 + p0 + q0 + 2) >> 2);
            d[-3 * across] = (uint8_t)((2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3);
        } else {
            d[-across] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
        }
        if (aq < beta && strong) {
            d[0] = (uint8_t)((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3);
            d[across] = (uint8_t)((p0 + q0 + q1 + q2 + 2) >> 2);
            d[2 * across] = (uint8_t)((2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3);
        } else {
            d[0] = (uint8_t)((2 * q1 + q0 + p1 + 2) >> 2);
        }
    }
}

// Stage 3 cuts each picture into bands of kH264DbkRows (16) macroblock
// rows, two workgroups per band, one for luma and one for chroma (the two
// are independent, and apart their waves each have a SIMD to themselves):
// a slot per row, slot s filtering row kH264DbkRows * band + s with 16
// lanes, so each lane's chain of edges per step is one plane's. Each row takes one macroblock per step, with one
// barrier per step, and each step filters the vertical edges before the
// horizontal ones. MB (x, y) needs MB (x - 1, y) complete and, for its top
// edge only, MB (x + 1, y - 1)'s vertical edges (the left edge of that MB
// filters the bottom-right corner of MB (x, y - 1)). So row s starts one
// step after row s - 1 when both are in one wave (the wave's program order
// puts the row above's vertical edges first) and two steps after it across
// waves (behind the barrier): dbk_lag(). Bands take their index from a
// per-picture ticket of their plane kind, so a band's predecessor of the
// same kind is always running. The first
// row of a band takes the strips of the row above from the band before
// through the exchange below.
//
// Dependencies through the samples. MB (x, y) reads its own unfiltered
// samples, columns 12..15 of MB (x - 1, y) as that MB left them, and rows
// 12..15 of MB (x, y - 1) after MB (x + 1, y - 1) filtered its left edge.
// It changes columns 13..15 of MB (x - 1, y), rows 13..15 of MB (x, y - 1)
// and itself. So every sample is read once from the picture (before any
// filtering of it) and written once (when final): its macroblock's rows
// 0..12 x columns 0..11 by the macroblock itself, columns 12..15 by the
// next macroblock of the row (its left edge), rows 13..15 by the
// macroblock below (its top edge) — or sooner at the right and bottom
// borders. Chroma likewise with columns / rows 7 and 6..7. The carried
// samples stay on chip:
//   tile  the macroblock after its vertical edges (luma 16 x 16, Cb and Cr
//         8 x 8), rewritten after the horizontal edges in the columns the
//         next macroblock's left edge reads (luma 12..15, chroma 6..7);
//   ring  per macroblock of the slot's row, rows 12..15 of each luma
//         column and rows 6..7 of each chroma column, final but for the
//         row below's top edge, which reads entry x at its step x (3
//         entries: across waves the row below runs 2 macroblocks behind,
//         while this row fills entries x + 1 and x + 2).
//   exchange  the last row of a band publishes each ring entry, once
//         complete, as 24 64-bit words (data | tag << 32, tag = x + 1),
//         relaxed atomics at device scope: each word is coherent on its own,
//         so the first row of the next band, which polls its words until the
//         tag matches, needs no cache-wide release or acquire. The strength
//         stage zeroes the exchange and the ticket before every launch.
// Lane r of a slot filters luma row r (luma waves), or chroma row r & 7 of
// Cb (r < 8) or Cr (chroma waves), across the vertical edges, then the
// column of the same index across the horizontal edges, the samples of a
// line in registers.

constexpr int kDbkSlots = kH264DbkRows;  // rows per band
constexpr int kDbkThreads = 16 * kDbkSlots;  // 16 lanes per row (luma, or Cb and Cr)
constexpr int kDbkTile = 96;   // words: luma 16 x 16 bytes, then Cb and Cr 8 x 8
constexpr int kDbkRing = 24;   // words per entry: 16 luma columns, 16 chroma (16-bit)
constexpr int kDbkRingN = 3;   // entries per slot
constexpr int kDbkPer = kDbkTile + kDbkRing * kDbkRingN;
// The exchange's poll bound (H264PicDev::xch_spin, 2^18 polls) is far beyond
// any real wait: the band before is running and publishes within a few
// steps. A poll that reaches it counts an error on the device, which the
// host reports at its next synchronisation as RV_E_DEVICE, and goes on
// with what it read rather than hang the workgroup.
static_assert(kDbkSlots % 4 == 0 && kDbkThreads <= 1024, "four slots of 16 lanes per 64-lane wave");

// The step at which slot s reaches macroblock 0: one step after slot s - 1
// in the same wave, two across a wave boundary.
__device__ __forceinline__ int dbk_lag(int s) { return s + (s >> 2); }

// |a - b| of two samples (0..255) in one v_sad_u8.
__device__ __forceinline__ int absd(int a, int b) { return (int)__builtin_amdgcn_sad_u8((uint32_t)a, (uint32_t)b, 0u); }

__device__ __forceinline__ uint32_t pk4(int a, int b, int c, int d) {
    const uint32_t lo = __builtin_amdgcn_perm((uint32_t)b, (uint32_t)a, 0x0c0c0400u);
    const uint32_t hi = __builtin_amdgcn_perm((uint32_t)d, (uint32_t)c, 0x04000c0cu);
    return lo | hi;
}

template <int O, int N> __device__ __forceinline__ void unpk4(int (&v)[N], uint32_t w) {
    v[O] = w & 0xFF;
    v[O + 1] = (w >> 8) & 0xFF;
    v[O + 2] = (w >> 16) & 0xFF;
    v[O + 3] = w >> 24;
}

// deblock_line (src/dsp/h264.rs:603-656) on one line held in registers:
// p3..q3 = v[Q - 4 .. Q + 3]. `a` is the line's parameter word (alpha |
// beta << 8 | bS-4 flag << 16 | tcb << 24; tcb 0: leave the line alone,
// else tC0 + 1). kLuma: the luma filter (p2, q2 read, p2..q2 written);
// else chroma (p1..q1 read, p0, q0 written). kBs4: a macroblock edge,
// which may carry bS 4.
template <int Q, bool kLuma, bool kBs4, int N>
__device__ __forceinline__ void dbk_line(int (&v)[N], uint32_t a) {
    const int alpha = a & 0xFF, beta = (a >> 8) & 0xFF;
    const uint32_t tcb = a >> 24;
    const int p0 = v[Q - 1], p1 = v[Q - 2], q0 = v[Q], q1 = v[Q + 1];
    const int d0 = absd(p0, q0);
    const bool on = tcb != 0 && d0 < alpha && absd(p1, p0) < beta && absd(q1, q0) < beta;
    const int tc0 = (int)tcb - 1;
    const int delta0 = (((q0 - p0) << 2) + (p1 - q1) + 4) >> 3;
    if constexpr (kLuma) {
        const int p2 = v[Q - 3], q2 = v[Q + 2];
        const bool lp = absd(p2, p0) < beta, lq = absd(q2, q0) < beta;
        const int tc = tc0 + (int)lp + (int)lq;
        const int delta = clampi(delta0, -tc, tc);
        int P0 = clampi(p0 + delta, 0, 255), Q0 = clampi(q0 - delta, 0, 255);
        const int avg = (p0 + q0 + 1) >> 1;
        int P1 = lp ? p1 + clampi((p2 + avg - (p1 << 1)) >> 1, -tc0, tc0) : p1;
        int Q1 = lq ? q1 + clampi((q2 + avg - (q1 << 1)) >> 1, -tc0, tc0) : q1;
        int P2 = p2, Q2 = q2;
        if constexpr (kBs4) {
            if ((a >> 16) & 1) {
                const int p3 = v[Q - 4], q3 = v[Q + 3];
                const bool strong = d0 < ((alpha >> 2) + 2);
                if (lp && strong) {
                    P0 = (p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3;
                    P1 = (p2 + p1 + p0 + q0 + 2) >> 2;
                    P2 = (2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3;
                } else {
                    P0 = (2 * p1 + p0 + q1 + 2) >> 2;
                    P1 = p1;
                }
                if (lq && strong) {
                    Q0 = (p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3;
                    Q1 = (p0 + q0 + q1 + q2 + 2) >> 2;
                    Q2 = (2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3;
                } else {
                    Q0 = (2 * q1 + q0 + p1 + 2) >> 2;
                    Q1 = q1;
                }
            }
        }
        if (on) {
            v[Q - 3] = P2, v[Q - 2] = P1, v[Q - 1] = P0;
            v[Q] = Q0, v[Q + 1] = Q1, v[Q + 2] = Q2;
        }
    } else {
        const int tc = tc0 + 1;
        const int delta = clampi(delta0, -tc, tc);
        int P0 = clampi(p0 + delta, 0, 255), Q0 = clampi(q0 - delta, 0, 255);
        if constexpr (kBs4) {
            if ((a >> 16) & 1) {
                P0 = (2 * p1 + p0 + q1 + 2) >> 2;
                Q0 = (2 * q1 + q0 + p1 + 2) >> 2;
            }
        }
        if (on) v[Q - 1] = P0, v[Q] = Q0;
    }
}

// One lane's inputs for one macroblock, fetched a step ahead: its line's
// samples (luma row r, or chroma row r & 7 of the lane's plane in .x, .y)
// and the parameter words of the line's segments across the vertical and
// the horizontal edges (luma edges 0..3, chroma edges 0 and 4 in .x, .y);
// in a band's first row, also its exchange word (luma column r; the chroma
// pair holding column r & 7 of its plane).
struct DbkIn {
    uint4 s, pv, ph;
    uint64_t ix;
};

// Global-memory pointers: pointers read from a descriptor are generic, and
// generic (flat) accesses count against the LDS wait counter too.
template <typename T> using gptr = __attribute__((address_space(1))) T*;
template <typename T> __device__ __forceinline__ gptr<T> to_global(T* p) { return (gptr<T>)p; }
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint4 gld4(gptr<const void> p) {
    const u32x4 v = *(gptr<const u32x4>)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 gld2(gptr<const void> p) {
    const u32x2 v = *(gptr<const u32x2>)p;
    return make_uint2(v.x, v.y);
}

// The picture fields stage 3 uses, read once: its byte stores could alias
// the descriptor as far as the compiler knows.
struct DbkPic {
    gptr<uint8_t> y;  // uniform: luma addresses are 32-bit offsets from it
    gptr<uint8_t> c;  // the lane's chroma plane
    uint32_t py;
    size_t pc;
    gptr<const uint32_t> prm;
    int mbw, mbh;
    // The exchange's poll bound, tag mask and error counter.
    int spin;
    uint32_t tag_xor;
    uint32_t* err;
};

__device__ __forceinline__ void dbk_fetch(const DbkPic& D, bool lu, int x, int y, int r, DbkIn& in) {
    const gptr<const uint32_t> pb = D.prm + ((size_t)y * D.mbw + x) * kH264DbkWords;
    if (lu) {
        const int sl = r >> 2;  // segment of the line
        in.s = gld4(D.y + ((uint32_t)(16 * y + r) * D.py + 16 * x));
        in.pv = gld4(pb + sl * 4), in.ph = gld4(pb + 16 + sl * 4);
    } else {
        const int pl = r >> 3, rr = r & 7, sc = rr >> 1;  // plane, line, segment
        const uint2 c = gld2(D.c + (size_t)(8 * y + rr) * D.pc + 8 * x);
        const uint2 v = gld2(pb + 32 + pl * 8 + sc * 2), h = gld2(pb + 48 + pl * 8 + sc * 2);
        // Whole vectors on both paths keep DbkIn in registers.
        in.s = make_uint4(c.x, c.y, 0u, 0u), in.pv = make_uint4(v.x, v.y, 0u, 0u), in.ph = make_uint4(h.x, h.y, 0u, 0u);
    }
}

// Workgroup barrier that orders LDS only: the next step's prefetched
// loads and this step's picture stores stay in flight across it (every
// picture sample is stored once and read only before it is stored).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ uint64_t xch_load(gptr<const uint64_t> p) {
    return __hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xch_store(gptr<uint64_t> p, uint64_t v) {
    __hip_atomic_store((uint64_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Poll an exchange word until it carries `tag`, at most D.spin times; a
// wait that gives up is counted in *D.err. The fields come from registers:
// re-reading the descriptor inside the step would wait on LDS as well.
__device__ __forceinline__ uint64_t xch_wait(const DbkPic& D, gptr<const uint64_t> p, uint64_t v, uint32_t tag) {
    tag ^= D.tag_xor;
    for (int n = 0; (uint32_t)(v >> 32) != tag && n < D.spin; ++n) {
        __builtin_amdgcn_s_sleep(1);
        v = xch_load(p);
    }
    if ((uint32_t)(v >> 32) != tag) __hip_atomic_fetch_add(D.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v;
}

__global__ void __launch_bounds__(kDbkThreads) h264_stage3(BatchDev b) {
    __shared__ uint32_t dl[kDbkSlots * kDbkPer];
    __shared__ int band_s;
    const int32_t* pre = b.prefix + 3 * (b.n + 1);
    const int p = find_pic(pre, b.n, (int)blockIdx.x);
    const H264PicDev& P = *b.h264[p];
    if (!P.deblock) return;
    // [0] luma ticket, [1] chroma ticket, then the boundaries. A picture's
    // workgroups alternate luma, chroma.
    const gptr<uint64_t> xch = to_global(P.xch);
    const bool lu = ((blockIdx.x - pre[p]) & 1) == 0;
    if (threadIdx.x == 0)
        band_s = (int)__hip_atomic_fetch_add((uint64_t*)xch + (lu ? 0 : 1), (uint64_t)1, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int band = band_s;
    if (band >= h264_deblock_bands(P.mbh)) return;  // only with a stale ticket
    const int t = (int)threadIdx.x;
    const int slot = t >> 4, r = t & 15;
    const int pl = r >> 3, rr = r & 7;  // chroma lanes: plane, row / column
    const DbkPic D{to_global((uint8_t*)P.cur[0]), to_global((uint8_t*)P.cur[1 + pl]), (uint32_t)P.pitch[0],
                   (size_t)P.pitch[1 + pl], to_global((const uint32_t*)P.dbk), P.mbw, P.mbh,
                   P.xch_spin, P.xch_tag_xor, P.err};
    const int mbw = D.mbw, mbh = D.mbh;
    const int y0 = band * kDbkSlots, rows = min(kDbkSlots, mbh - y0);
    const bool first = slot == 0 && band > 0;  // imports its strips
    // Publishes its strips: the band's last row, one word per luma lane and
    // per chroma lane 0..7.
    const bool last = slot == rows - 1 && y0 + rows < mbh && (lu || r < 8);
    // Exchange words of this lane: imported (boundary band - 1) and
    // published (boundary band), macroblock x at + 24 x.
    const gptr<const uint64_t> imp = xch + 2 + (size_t)max(band - 1, 0) * mbw * kDbkRing + (lu ? r : 16 + (r >> 1));
    const int pw = lu ? r : 16 + r;  // ring word this lane publishes
    const gptr<uint64_t> pub = xch + 2 + (size_t)band * mbw * kDbkRing + pw;
    uint32_t* const tile = dl + slot * kDbkPer;
    uint8_t* const tb = (uint8_t*)tile;
    uint32_t* const ring = tile + kDbkTile;
    const uint32_t* const up = dl + max(slot - 1, 0) * kDbkPer + kDbkTile;  // read by slots > 0
    const gptr<uint8_t> Y = D.y;
    const gptr<uint8_t> Cp = D.c;
    const uint32_t py = D.py;
    const size_t pc = D.pc;
    const int y = y0 + slot;
    const int steps = dbk_lag(rows - 1) + mbw + (y0 + rows < mbh ? 1 : 0);

    int x = -dbk_lag(slot), ex = 0;  // the slot's macroblock at the next step, x % 3
    bool act = slot < rows && x == 0;
    DbkIn ba, bb;  // this step's inputs and the next step's, alternating
    if (act) {
        dbk_fetch(D, lu, 0, y, r, ba);
        if (first) ba.ix = xch_load(imp);
    }
    // One step: filter the macroblock `in` holds, fetch the next into `nx`.
    auto step = [&](const DbkIn& in, DbkIn& nx) __attribute__((always_inline)) {
        // Wait here for this step's inputs (fetched a step ago) and the last
        // step's stores, before the next fetch: later waits then never hold
        // the next step's loads (the vector memory counter retires in order).
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        const int cx = x, cy = y, cex = ex;
        const bool cact = act;
        ++x;
        act = slot < rows && x >= 0 && x < mbw;
        ex = x <= 0 ? 0 : (ex + 1 == kDbkRingN ? 0 : ex + 1);  // x % 3
        if (act) {
            dbk_fetch(D, lu, x, y, r, nx);
            if (first) nx.ix = xch_load(imp + (size_t)x * kDbkRing);
        }
        const bool lastc = cx == mbw - 1, lastr = cy == mbh - 1;
        if (cact) {
            uint8_t* const le = (uint8_t*)(ring + (cex == 0 ? kDbkRingN - 1 : cex - 1) * kDbkRing);  // entry x - 1
            // Vertical edges: luma row r or chroma row rr. Then the left
            // neighbour's columns 12..15 (chroma 7) are final but for rows
            // 13..15 (chroma 7) outside the last row, which go to its ring
            // entry.
            if (lu) {
                int v[20];
                unpk4<0>(v, tile[r * 4 + 3]);  // previous macroblock, columns 12..15
                unpk4<4>(v, in.s.x), unpk4<8>(v, in.s.y), unpk4<12>(v, in.s.z), unpk4<16>(v, in.s.w);
                dbk_line<4, true, true>(v, in.pv.x);
                dbk_line<8, true, false>(v, in.pv.y);
                dbk_line<12, true, false>(v, in.pv.z);
                dbk_line<16, true, false>(v, in.pv.w);
                *(uint4*)&tile[r * 4] =
                    make_uint4(pk4(v[4], v[5], v[6], v[7]), pk4(v[8], v[9], v[10], v[11]),
                               pk4(v[12], v[13], v[14], v[15]), pk4(v[16], v[17], v[18], v[19]));
                if (cx > 0) {
                    if (r <= 12 || lastr)
                        *(gptr<uint32_t>)(Y + ((uint32_t)(16 * cy + r) * py + 16 * cx - 4)) = pk4(v[0], v[1], v[2], v[3]);
                    if (!lastr && r >= 12)
                        for (int k = 0; k < 4; ++k) le[(12 + k) * 4 + r - 12] = (uint8_t)v[k];
                }
            } else {
                int w[10];
                const uint32_t lc = tile[64 + pl * 16 + rr * 2 + 1];  // previous macroblock, columns 4..7
                w[0] = (lc >> 16) & 0xFF, w[1] = lc >> 24;
                unpk4<2>(w, in.s.x), unpk4<6>(w, in.s.y);
                dbk_line<2, false, true>(w, in.pv.x);
                dbk_line<6, false, false>(w, in.pv.y);
                *(uint2*)&tile[64 + pl * 16 + rr * 2] = make_uint2(pk4(w[2], w[3], w[4], w[5]), pk4(w[6], w[7], w[8], w[9]));
                if (cx > 0) {
                    if (rr <= 6 || lastr) Cp[(size_t)(8 * cy + rr) * pc + 8 * cx - 1] = (uint8_t)w[1];
                    if (!lastr && rr >= 6) le[64 + (pl * 8 + 7) * 2 + rr - 6] = (uint8_t)w[1];
                }
            }
        }
        // A slot's lanes, and the rows one step apart, are one wave's: LDS
        // order within the wave.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        // The band's last row publishes ring entry x - 1: its columns
        // 12..15 are final since this step's vertical edges, the others
        // since the last step's horizontal ones (all of them at the row's
        // end, which takes one more step).
        if (last) {
            const int xp = cx - 1;
            if (xp >= 0 && xp < mbw) {
                const uint32_t* e = ring + (cex == 0 ? kDbkRingN - 1 : cex - 1) * kDbkRing;
                xch_store(pub + (size_t)xp * kDbkRing, ((uint64_t)(xp + 1) << 32) | e[pw]);
            }
        }
        if (cact) {
            // Horizontal edges: luma column r or chroma column rr; rows
            // above from the ring of the row above, or the exchange in a
            // band's first row (unused in row 0: no top edge).
            uint32_t u = 0;
            if (slot > 0) {
                const uint32_t* ue = up + cex * kDbkRing;
                u = lu ? ue[r] : ((const uint16_t*)(ue + 16))[pl * 8 + rr];
            } else if (first) {
                u = (uint32_t)xch_wait(D, imp + (size_t)cx * kDbkRing, in.ix, (uint32_t)cx + 1);
                if (!lu) u = (u >> (16 * (r & 1))) & 0xFFFF;
            }
            if (lu) {
                int h[20];
                unpk4<0>(h, u);
                for (int k = 0; k < 16; ++k) h[4 + k] = tb[k * 16 + r];
                dbk_line<4, true, true>(h, in.ph.x);
                dbk_line<8, true, false>(h, in.ph.y);
                dbk_line<12, true, false>(h, in.ph.z);
                dbk_line<16, true, false>(h, in.ph.w);
                const uint32_t col = (uint32_t)(16 * cy) * py + 16 * cx + r;
                if (cy > 0) {  // rows 13..15 of the macroblock above are final
                    Y[col - 3 * py] = (uint8_t)h[1];
                    Y[col - 2 * py] = (uint8_t)h[2];
                    Y[col - py] = (uint8_t)h[3];
                }
                if (r <= 11 || lastc) {
                    for (int k = 0; k < 13; ++k) Y[col + k * py] = (uint8_t)h[4 + k];
                    if (lastr) {
                        for (int k = 13; k < 16; ++k) Y[col + k * py] = (uint8_t)h[4 + k];
                    } else {
                        ring[cex * kDbkRing + r] = pk4(h[16], h[17], h[18], h[19]);
                    }
                }
                if (!lastc && r >= 12)  // the columns the next macroblock's left edge reads
                    for (int k = 0; k < 16; ++k) tb[k * 16 + r] = (uint8_t)h[4 + k];
            } else {
                int g[10];
                g[0] = u & 0xFF, g[1] = u >> 8;
                for (int k = 0; k < 8; ++k) g[2 + k] = tb[256 + pl * 64 + k * 8 + rr];
                dbk_line<2, false, true>(g, in.ph.x);
                dbk_line<6, false, false>(g, in.ph.y);
                gptr<uint8_t> ccol = Cp + (size_t)(8 * cy) * pc + 8 * cx + rr;
                if (cy > 0) ccol[-(long)pc] = (uint8_t)g[1];
                if (rr <= 6 || lastc) {
                    for (int k = 0; k < 7; ++k) ccol[k * pc] = (uint8_t)g[2 + k];
                    if (lastr) ccol[7 * pc] = (uint8_t)g[9];
                    else ((uint16_t*)(ring + cex * kDbkRing + 16))[pl * 8 + rr] = (uint16_t)(g[8] | g[9] << 8);
                }
                if (!lastc && rr >= 6)
                    for (int k = 0; k < 8; ++k) tb[256 + pl * 64 + k * 8 + rr] = (uint8_t)g[2 + k];
            }
        }
        lds_barrier();
    };
    for (int s2 = 0; s2 < steps; s2 += 2) {
        step(ba, bb);
        if (s2 + 1 < steps) step(bb, ba);
    }
}

// ---------------------------------------------------------------------
// Per-entry kernel: one call of one H264Dsp<u8> entry (src/dsp/h264.rs:
// 123-170) on one workgroup of 256 threads, through the device functions
// of the batched stages. Prediction blocks are 16 samples per row
// (PRED_STRIDE, :22-24).
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256) h264_entry(H264Entry e) {
    __shared__ int16_t win[21 * 32];
    const int tid = threadIdx.x;
    const int n = e.w * e.h;
    switch (e.op) {
        case HA_QPEL: {
            // The (w+5) x (h+5) six-tap window, stride 32 as stage 0 keeps it.
            for (int i = tid; i < (e.h + 5) * 32; i += 256) {
                const int r = i >> 5, c = i & 31;
                if (c < e.w + 5) win[i] = e.src[r * e.src_stride + c];
            }
            __syncthreads();
            const Win q{win, e.max};
            for (int i = tid; i < n; i += 256) {
                const int y = i / e.w, x = i - y * e.w;
                e.dst[y * 16 + x] = (uint8_t)qpel_at(q, e.pos, x, y);
            }
            break;
        }
        case HA_CHROMA: {  // chroma_scalar (:499-520)
            const int a = (8 - e.xf) * (8 - e.yf), b = e.xf * (8 - e.yf), c = (8 - e.xf) * e.yf, d = e.xf * e.yf;
            for (int i = tid; i < n; i += 256) {
                const int y = i / e.w, x = i - y * e.w;
                const uint8_t* r0 = e.src + y * e.src_stride + x;
                const uint8_t* r1 = r0 + e.src_stride;
                e.dst[y * 16 + x] = (uint8_t)((a * r0[0] + b * r0[1] + c * r1[0] + d * r1[1] + 32) >> 6);
            }
            break;
        }
        case HA_COPY:
        case HA_AVG:
        case HA_WUNI:
        case HA_WBI:  // copy / avg / weighted_uni / weighted_bi (:522-597)
            for (int i = tid; i < n; i += 256) {
                const int y = i / e.w, x = i - y * e.w;
                const int va = e.src[y * 16 + x];
                int v;
                if (e.op == HA_COPY) v = va;
                else if (e.op == HA_AVG) v = (va + e.b[y * 16 + x] + 1) >> 1;
                else if (e.op == HA_WUNI)
                    v = e.log_wd >= 1 ? clampi(((va * e.w0 + (1 << (e.log_wd - 1))) >> e.log_wd) + e.o0, 0, e.max)
                                      : clampi(va * e.w0 + e.o0, 0, e.max);
                else
                    v = clampi(((va * e.w0 + e.b[y * 16 + x] * e.w1 + (1 << e.log_wd)) >> (e.log_wd + 1)) +
                                   ((e.o0 + e.o1 + 1) >> 1),
                               0, e.max);
                e.dst[y * e.stride + x] = (uint8_t)v;
            }
            break;
        case HA_DBK: {  // deblock_luma_lines<16> / deblock_chroma_scalar (:658-763)
            const int lines = e.chroma ? 8 : 16;
            if (tid < lines) {
                const int t = e.intra ? 0 : e.tc0[tid / (e.chroma ? 2 : 4)];
                if (e.intra || t >= 0) {
                    const long along = e.vertical ? e.stride : 1, across = e.vertical ? 1 : e.stride;
                    line_filter(e.dst + e.off + tid * along, across, e.intra != 0, t, e.alpha, e.beta, e.chroma != 0,
                                e.max);
                }
            }
            break;
        }
        case HA_IDCT4:
            if (tid == 0) idct4_add16(e.dst, e.stride, (const int16_t*)e.coefs, e.max);
            break;
        case HA_IDCT8:
            if (tid == 0) idct8_add16(e.dst, e.stride, (const int16_t*)e.coefs, e.max);
            break;
        case HA_DC:
            if (tid == 0) dc_add16(e.dst, e.stride, e.n, (int16_t)e.dc, e.max);
            break;
        case HA_RES4:
            if (tid == 0) residual4(e.dst, e.stride, (const int32_t*)e.coefs, e.has_dc, e.dc, e.max);
            break;
        case HA_RES8:
            if (tid == 0) residual8(e.dst, e.stride, (const int32_t*)e.coefs, e.max);
            break;
        default: break;
    }
}

}  // namespace

hipError_t h264_entry_launch(const H264Entry& e, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(h264_entry, dim3(blocks), dim3(256), 0, st, e);
    return hipGetLastError();
}

void h264_stage_items(H264PicDev& P) {
    P.items[0] = P.n_intra + P.n_big + (P.n_parts - P.n_big + 3) / 4;
    P.items[1] = (P.n_blks4 + 63) / 64 + (P.n_blks8 + 63) / 64;
    P.items[2] = (P.n_mbs * 8 + 63) / 64;
    P.items[3] = P.deblock ? 2 * h264_deblock_bands(P.mbh) : 0;  // a luma and a chroma workgroup per band
    P.items[4] = 0;
}

size_t h264_deblock_xch_words(int mbw, int mbh) {
    return 2 + (size_t)(h264_deblock_bands(mbh) - 1) * mbw * kDbkRing;
}

hipError_t h264_launch_stage(int stage, const BatchDev& b, int total, hipStream_t st) {
    if (stage == 1) {
        // The transforms, then the device-intra macroblocks level by level.
        if (total > 0) hipLaunchKernelGGL(h264_stage1, dim3(total), dim3(64), 0, st, b);
        for (int l = 0; l < b.ilevels; ++l)
            if (b.itotal[l] > 0) hipLaunchKernelGGL(h264_intra_level, dim3(b.itotal[l]), dim3(64), 0, st, b, l);
        return hipGetLastError();
    }
    if (total <= 0) return hipSuccess;
    switch (stage) {
        case 0: hipLaunchKernelGGL(h264_stage0, dim3(total), dim3(64), 0, st, b); break;
        case 2: hipLaunchKernelGGL(h264_stage2, dim3(total), dim3(64), 0, st, b); break;
        case 3: hipLaunchKernelGGL(h264_stage3, dim3(total), dim3(kDbkThreads), 0, st, b); break;
        default: return hipSuccess;
    }
    return hipGetLastError();
}

}  // namespace rv
