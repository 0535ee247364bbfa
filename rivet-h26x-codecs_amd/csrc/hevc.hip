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

#include <algorithm>
#include <cstddef>

#include "rv_device.h"
#include "rv_kernels.h"

namespace rv {
namespace {

// Table 8-12 / 8-13 (src/hevc/tables_gen.rs:173-191); row 0 unused.
struct FilterTables {
    int8_t q[4][8];  // luma taps by quarter fraction
    int8_t e[8][4];  // chroma taps by eighth fraction
    // The same taps as packed i16 pairs for v_dot2_i32_i16: luma even pairs
    // (f0 f1)(f2 f3)(f4 f5)(f6 f7), odd pairs (0 f0)(f1 f2)(f3 f4)(f5 f6)
    // (f7 0); chroma even (f0 f1)(f2 f3), odd (0 f0)(f1 f2)(f3 0). Odd pairs
    // serve outputs whose inputs start one sample after a pair boundary.
    // Row 0 holds the identity (64 at the centre tap) instead of zeros, so
    // every fraction takes the two-pass path: for bd <= 12 a whole-sample
    // first stage gives x << (6 - shift1) exactly, and the second stage's
    // >> 6 then equals the one-stage filters' >> shift1 and the copy's
    // << (14 - bd) (src/dsp/hevc.rs:823-882).
    alignas(16) uint32_t qpk[4][12];
    alignas(16) uint32_t epk[8][8];
    // Signed-byte taps for v_dot4_i32_i8 (the 8-bit first stage): luma taps
    // 0..3 then 4..7, chroma taps 0..3, one dword each; row 0 the identity.
    alignas(16) uint32_t qb[4][2];
    uint32_t eb[8];
};
constexpr uint32_t pair16(int lo, int hi) { return (uint32_t)(lo & 0xFFFF) | ((uint32_t)(hi & 0xFFFF) << 16); }
constexpr uint32_t bytes4(int a, int b, int c, int d) {
    return (uint32_t)(a & 0xFF) | ((uint32_t)(b & 0xFF) << 8) | ((uint32_t)(c & 0xFF) << 16) | ((uint32_t)(d & 0xFF) << 24);
}
constexpr FilterTables make_filter_tables() {
    FilterTables t{{{0, 0, 0, 0, 0, 0, 0, 0},
                    {-1, 4, -10, 58, 17, -5, 1, 0},
                    {-1, 4, -11, 40, 40, -11, 4, -1},
                    {0, 1, -5, 17, 58, -10, 4, -1}},
                   {{0, 0, 0, 0}, {-2, 58, 10, -2}, {-4, 54, 16, -2}, {-6, 46, 28, -4},
                    {-4, 36, 36, -4}, {-4, 28, 46, -6}, {-2, 16, 54, -4}, {-2, 10, 58, -2}},
                   {},
                   {}};
    for (int f = 0; f < 4; ++f) {
        int g[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // g[k + 1] = tap k, g[0] = g[9] = 0
        for (int k = 0; k < 8; ++k) g[k + 1] = f ? t.q[f][k] : (k == 3 ? 64 : 0);
        for (int i = 0; i < 4; ++i) t.qpk[f][i] = pair16(g[2 * i + 1], g[2 * i + 2]);
        for (int i = 0; i < 5; ++i) t.qpk[f][4 + i] = pair16(g[2 * i], g[2 * i + 1]);
        t.qb[f][0] = bytes4(g[1], g[2], g[3], g[4]);
        t.qb[f][1] = bytes4(g[5], g[6], g[7], g[8]);
    }
    for (int f = 0; f < 8; ++f) {
        int g[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) g[k + 1] = f ? t.e[f][k] : (k == 1 ? 64 : 0);
        for (int i = 0; i < 2; ++i) t.epk[f][i] = pair16(g[2 * i + 1], g[2 * i + 2]);
        for (int i = 0; i < 3; ++i) t.epk[f][2 + i] = pair16(g[2 * i], g[2 * i + 1]);
        t.eb[f] = bytes4(g[1], g[2], g[3], g[4]);
    }
    return t;
}
__constant__ FilterTables kF = make_filter_tables();
// tC' / beta' (src/hevc/tables_gen.rs:195-199).
__constant__ uint8_t kTc[54] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4,
                                4, 4, 5, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 22, 24};
__constant__ uint8_t kBeta[52] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  6,  7,
                                  8,  9,  10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 22, 24, 26, 28, 30, 32,
                                  34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64};
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }
__device__ __forceinline__ int sgn(int v) { return (v > 0) - (v < 0); }

// The batch tables and picture descriptors are written before the launch
// and only read while the stage kernels run. Reading them through the
// constant address space lets uniform reads use scalar loads (the compiler
// cannot otherwise rule out that the kernel's own stores alias them).
typedef const HevcPicDev __attribute__((address_space(4))) ConstPic;
typedef ConstPic& PicRef;
typedef const int32_t __attribute__((address_space(4))) ConstI32;

__device__ __forceinline__ PicRef picture(const BatchDev& b, int p) {
    return *(ConstPic*)b.hevc[p];
}
__device__ __forceinline__ ConstI32* prefix(const BatchDev& b, int stage) {
    return (ConstI32*)(b.prefix + stage * (b.n + 1));
}

__device__ __forceinline__ int find_pic(ConstI32* pre, int n, int it) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= it) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Kernel variants by sample depth: D = 0 serves any picture, D = 1 only
// 8-bit pictures and D = 2 only deeper ones, so the specialised variants
// compile one sample path and need fewer registers. A batch picks among
// the variants its pictures allow (the autotuner times them).
template <int D>
__device__ __forceinline__ bool eight_bit(int bd) {
    return D == 1 ? true : D == 2 ? false : bd == 8;
}

// ---------------------------------------------------------------------
// Stage 0: side data, intra blocks, motion compensation
// ---------------------------------------------------------------------

template <typename S>
__device__ __forceinline__ void intra_copy(PicRef P, const rv_hevc_cu& cu, int tid) {
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

typedef short short2v __attribute__((ext_vector_type(2)));

// Weighting of one component of one block: the explicit weights and
// offsets of its references (Weighting::Explicit, src/hevc/inter.rs:71-84),
// chosen per lane from the picture's (uniform) tables.
struct Wp {
    int w[2], o[2];
};
__device__ __forceinline__ int pick4(int k, int a, int b, int c, int d) { return k == 0 ? a : k == 1 ? b : k == 2 ? c : d; }
__device__ __forceinline__ Wp weights(PicRef P, int c, const int (&ri)[2]) {
    Wp r;
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        r.w[l] = pick4(ri[l], P.wp_w[l][0][c], P.wp_w[l][1][c], P.wp_w[l][2][c], P.wp_w[l][3][c]);
        r.o[l] = pick4(ri[l], P.wp_o[l][0][c], P.wp_o[l][1][c], P.wp_o[l][2][c], P.wp_o[l][3][c]);
    }
    return r;
}

// The combine of predict_block (src/hevc/inter.rs:712-741) with the
// scalar combiners (src/dsp/hevc.rs:997-1087): default uni / bi, or
// explicit weights (lwd = log2_wd of the component).
__device__ __forceinline__ int combine(bool weighted, int lwd, const Wp& wp, int dir, int a, int b, int bd) {
    const int max = (1 << bd) - 1;
    const bool both = dir == 3;
    if (!weighted) {
        if (!both) {
            int shift = 14 - bd;
            int v = dir == 1 ? a : b;
            return clampi((v + (1 << (shift - 1))) >> shift, 0, max);
        }
        int shift = 15 - bd;
        return clampi((a + b + (1 << (shift - 1))) >> shift, 0, max);
    }
    if (!both) {
        const int l = dir == 1 ? 0 : 1;
        const int v = l == 0 ? a : b, wt = wp.w[l], o = wp.o[l];
        if (lwd >= 1) return clampi(((v * wt + (1 << (lwd - 1))) >> lwd) + o, 0, max);
        return clampi(v * wt + o, 0, max);
    }
    return clampi((a * wp.w[0] + b * wp.w[1] + ((wp.o[0] + wp.o[1] + 1) << lwd)) >> (lwd + 1), 0, max);
}

__device__ __forceinline__ short2v pk(int lo, int hi) {
    short2v r;
    r.x = (short)lo;
    r.y = (short)hi;
    return r;
}
__device__ __forceinline__ short2v as2(uint32_t v) { return __builtin_bit_cast(short2v, v); }
__device__ __forceinline__ int dot2(short2v a, short2v b, int c) { return __builtin_amdgcn_sdot2(a, b, c, false); }

// The taps of one phase as packed pairs (FilterTables::qpk / epk) or signed
// bytes (qb / eb), per lane (the blocks of one wave have their own
// vectors), read from a copy of the packed tables in LDS: `tab` holds qpk
// (4 x 12 words), epk (8 x 8), qb (4 x 2) and eb (8).
constexpr int kTapWords = 4 * 12 + 8 * 8 + 4 * 2 + 8;
constexpr int kQbWord = 4 * 12 + 8 * 8, kEbWord = kQbWord + 4 * 2;  // offsets of qb, eb
static_assert(offsetof(FilterTables, epk) == offsetof(FilterTables, qpk) + sizeof(uint32_t) * 48 &&
                  offsetof(FilterTables, qb) == offsetof(FilterTables, qpk) + sizeof(uint32_t) * kQbWord &&
                  offsetof(FilterTables, eb) == offsetof(FilterTables, qpk) + sizeof(uint32_t) * kEbWord,
              "the packed tap tables are contiguous");
struct Taps8 {
    short2v e[4], o[5];
};
__device__ __forceinline__ Taps8 luma_taps(const uint32_t* tab, int frac) {
    const uint4* p = (const uint4*)(tab + 12 * frac);
    const uint4 a = p[0], b = p[1];
    const uint32_t c = tab[12 * frac + 8];
    Taps8 t;
    t.e[0] = as2(a.x), t.e[1] = as2(a.y), t.e[2] = as2(a.z), t.e[3] = as2(a.w);
    t.o[0] = as2(b.x), t.o[1] = as2(b.y), t.o[2] = as2(b.z), t.o[3] = as2(b.w), t.o[4] = as2(c);
    return t;
}
struct Taps4 {
    short2v e[2], o[3];
};
__device__ __forceinline__ Taps4 chroma_taps(const uint32_t* tab, int frac) {
    const uint4 a = *(const uint4*)(tab + 48 + 8 * frac);
    const uint32_t b = tab[48 + 8 * frac + 4];
    Taps4 t;
    t.e[0] = as2(a.x), t.e[1] = as2(a.y);
    t.o[0] = as2(a.z), t.o[1] = as2(a.w), t.o[2] = as2(b);
    return t;
}
__device__ __forceinline__ uint2 luma_taps_b(const uint32_t* tab, int frac) {
    return *(const uint2*)(tab + kQbWord + 2 * frac);
}
__device__ __forceinline__ uint32_t chroma_taps_b(const uint32_t* tab, int frac) { return tab[kEbWord + frac]; }

// 8-bit first stage with v_dot4_i32_i8: the span's bytes x are biased to
// x - 128 (top bit flipped) to make them signed; every filter's taps sum to
// 64, so starting each sum at 128 * 64 restores sum f x exactly. w holds
// the span's samples in byte order, 4 per dword.
__device__ __forceinline__ int dot4(uint32_t a, uint32_t b, int c) {
    return __builtin_amdgcn_sdot4((int)a, (int)b, c, false);
}
__device__ __forceinline__ void filt4_b(const uint32_t (&w)[3], uint2 t, int (&acc)[4]) {
    const uint32_t x0 = w[0] ^ 0x80808080u, x1 = w[1] ^ 0x80808080u, x2 = w[2] ^ 0x80808080u;
    acc[0] = dot4(x1, t.y, dot4(x0, t.x, 8192));
#pragma unroll
    for (int j = 1; j < 4; ++j)
        acc[j] = dot4(__builtin_amdgcn_alignbyte(x2, x1, j), t.y, dot4(__builtin_amdgcn_alignbyte(x1, x0, j), t.x, 8192));
}
__device__ __forceinline__ void filt4c_b(const uint32_t (&w)[2], uint32_t t, int (&acc)[4]) {
    const uint32_t x0 = w[0] ^ 0x80808080u, x1 = w[1] ^ 0x80808080u;
    acc[0] = dot4(x0, t, 8192);
#pragma unroll
    for (int j = 1; j < 4; ++j) acc[j] = dot4(__builtin_amdgcn_alignbyte(x1, x0, j), t, 8192);
}

// Four consecutive 8-tap outputs from twelve inputs held as six packed
// pairs p[0..5] (inputs 2i, 2i+1): output j = sum_k f[k] in[j + k], exact
// in i32 (v_dot2_i32_i16).
__device__ __forceinline__ void filt4(const short2v (&p)[6], const Taps8& t, int (&acc)[4]) {
    acc[0] = dot2(p[3], t.e[3], dot2(p[2], t.e[2], dot2(p[1], t.e[1], dot2(p[0], t.e[0], 0))));
    acc[1] = dot2(p[4], t.o[4], dot2(p[3], t.o[3], dot2(p[2], t.o[2], dot2(p[1], t.o[1], dot2(p[0], t.o[0], 0)))));
    acc[2] = dot2(p[4], t.e[3], dot2(p[3], t.e[2], dot2(p[2], t.e[1], dot2(p[1], t.e[0], 0))));
    acc[3] = dot2(p[5], t.o[4], dot2(p[4], t.o[3], dot2(p[3], t.o[2], dot2(p[2], t.o[1], dot2(p[1], t.o[0], 0)))));
}

// The same for four 4-tap outputs from eight inputs (four pairs).
__device__ __forceinline__ void filt4c(const short2v (&p)[4], const Taps4& t, int (&acc)[4]) {
    acc[0] = dot2(p[1], t.e[1], dot2(p[0], t.e[0], 0));
    acc[1] = dot2(p[2], t.o[2], dot2(p[1], t.o[1], dot2(p[0], t.o[0], 0)));
    acc[2] = dot2(p[2], t.e[1], dot2(p[1], t.e[0], 0));
    acc[3] = dot2(p[3], t.o[2], dot2(p[2], t.o[1], dot2(p[1], t.o[0], 0)));
}

__device__ __forceinline__ void load_pairs(const int16_t* src, short2v (&p)[6]) {
    // Three 8-byte LDS reads; src is 8-byte aligned.
    const uint2 a = *(const uint2*)src, b = *(const uint2*)(src + 4), c = *(const uint2*)(src + 8);
    p[0] = as2(a.x), p[1] = as2(a.y), p[2] = as2(b.x), p[3] = as2(b.y), p[4] = as2(c.x), p[5] = as2(c.y);
}

// N (8 or 12) samples of one reference row from column gx on, as N / 2
// packed i16 pairs, in two steps so that a wave can issue the loads of all
// its spans before it uses any: span_at() clamps the row to the plane and
// places a window of N columns inside it (at gx when the span lies inside
// the plane), read as aligned dwords; span_pairs() extracts the samples
// with alignbyte or, for a span crossing the plane's left or right edge,
// picks column clamp(gx + i, 0, pw - 1) of the window for each sample.
// Both give the samples of the reference's replicated border
// (src/hevc/frame.rs:169-214) and clamped gather (src/hevc/inter.rs:116-123).
// A span keeps only its first dword and its column (the window's start
// column, byte shift and whether it crosses an edge follow from the
// column), so little stays live while a wave's loads are in flight.
template <typename S, int N>
struct Span {
    static constexpr int ND = N * (int)sizeof(S) / 4 + 1;  // dwords covering the span
    const uint32_t* w;  // first dword
    int gx;             // column of the span's first sample
};
// The window's first column: gx when the span lies inside the plane, else
// the in-plane start that holds every clamped column. Reads may run past
// the row's last sample by less than N + 2 samples, inside the row pitch
// or the plane's slack.
template <int N>
__device__ __forceinline__ int span_cx(int gx, int pw) {
    return max(min(gx, pw - N), 0);
}
template <typename S, int N>
__device__ __forceinline__ Span<S, N> span_at(const S* pl, int pitch, int pw, int ph, int gx, int gy) {
    Span<S, N> sp;
    const S* row = pl + (size_t)clampi(gy, 0, ph - 1) * pitch;
    const uint32_t off = (uint32_t)span_cx<N>(gx, pw) * sizeof(S);
    sp.w = (const uint32_t*)((const char*)row + (off & ~3u));
    sp.gx = gx;
    return sp;
}
// Sample i of a span crossing the plane's left or right edge: column
// clamp(gx + i, 0, pw - 1) of the span's (clamped) row, read again from
// memory. The row starts cx samples, rounded down to a dword, before the
// span's first dword. Such spans are rare (the picture's outer columns).
template <typename S, int N>
__device__ __forceinline__ uint32_t span_clamped(const Span<S, N>& sp, int pw, int cx, int i) {
    const S* row = (const S*)((const char*)sp.w - (((uint32_t)cx * sizeof(S)) & ~3u));
    return row[clampi(sp.gx + i, 0, pw - 1)];
}
template <typename S, int N>
__device__ __forceinline__ bool span_inside(const Span<S, N>& sp, int pw) {
    return sp.gx >= 0 && sp.gx + N - 1 < pw;
}
template <typename S, int N>
__device__ __forceinline__ void span_pairs(const Span<S, N>& sp, const uint32_t (&d)[Span<S, N>::ND], int pw,
                                           short2v (&p)[N / 2]) {
    const int cx = span_cx<N>(sp.gx, pw);
    const uint32_t sh = ((uint32_t)cx * sizeof(S)) & 3u;  // byte offset of column cx in d[0]
    if (!span_inside(sp, pw)) {
#pragma unroll
        for (int i = 0; i < N / 2; ++i) p[i] = as2(span_clamped(sp, pw, cx, 2 * i) | (span_clamped(sp, pw, cx, 2 * i + 1) << 16));
        return;
    }
    if constexpr (sizeof(S) == 1) {
#pragma unroll
        for (int i = 0; i < N / 4; ++i) {
            const uint32_t v = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
            p[2 * i] = as2(__builtin_amdgcn_perm(0u, v, 0x0c010c00u));
            p[2 * i + 1] = as2(__builtin_amdgcn_perm(0u, v, 0x0c030c02u));
        }
    } else {
#pragma unroll
        for (int i = 0; i < N / 2; ++i) p[i] = as2(__builtin_amdgcn_alignbyte(d[i + 1], d[i], sh));
    }
}
// The 8-bit span's N samples as bytes, 4 per dword.
template <int N>
__device__ __forceinline__ void span_bytes(const Span<uint8_t, N>& sp, const uint32_t (&d)[Span<uint8_t, N>::ND],
                                           int pw, uint32_t (&w)[N / 4]) {
    const int cx = span_cx<N>(sp.gx, pw);
    const uint32_t sh = (uint32_t)cx & 3u;
    if (!span_inside(sp, pw)) {
#pragma unroll
        for (int i = 0; i < N / 4; ++i)
            w[i] = span_clamped(sp, pw, cx, 4 * i) | (span_clamped(sp, pw, cx, 4 * i + 1) << 8) |
                   (span_clamped(sp, pw, cx, 4 * i + 2) << 16) | (span_clamped(sp, pw, cx, 4 * i + 3) << 24);
        return;
    }
#pragma unroll
    for (int i = 0; i < N / 4; ++i) w[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
}

// First-stage outputs of one task: 4 luma outputs from a 12-sample span or
// 4 chroma outputs from an 8-sample span, at phase `frac`; dot4 for 8-bit
// samples, dot2 on i16 pairs above.
template <typename S>
__device__ __forceinline__ void first_luma(const Span<S, 12>& sp, const uint32_t (&d)[Span<S, 12>::ND], int pw,
                                           const uint32_t* tab, int frac, int (&acc)[4]) {
    if constexpr (sizeof(S) == 1) {
        uint32_t w[3];
        span_bytes<12>(sp, d, pw, w);
        filt4_b(w, luma_taps_b(tab, frac), acc);
    } else {
        short2v p[6];
        span_pairs<S, 12>(sp, d, pw, p);
        filt4(p, luma_taps(tab, frac), acc);
    }
}
template <typename S>
__device__ __forceinline__ void first_chroma(const Span<S, 8>& sp, const uint32_t (&d)[Span<S, 8>::ND], int pw,
                                             const uint32_t* tab, int frac, int (&acc)[4]) {
    if constexpr (sizeof(S) == 1) {
        uint32_t w[2];
        span_bytes<8>(sp, d, pw, w);
        filt4c_b(w, chroma_taps_b(tab, frac), acc);
    } else {
        short2v p[4];
        span_pairs<S, 8>(sp, d, pw, p);
        filt4c(p, chroma_taps(tab, frac), acc);
    }
}

// Plane c of reference ri of list l, chosen per lane among the picture's
// (uniform) reference table.
__device__ __forceinline__ const void* ref_plane(PicRef P, int l, int ri, int c) {
    const void* p0 = P.ref[l][0][c];
    const void* p1 = P.ref[l][1][c];
    const void* p2 = P.ref[l][2][c];
    const void* p3 = P.ref[l][3][c];
    return ri == 0 ? p0 : ri == 1 ? p1 : ri == 2 ? p2 : p3;
}

// LDS of one MC work item (four blocks, both lists): first-stage outputs,
// transposed ([column][row]) so the second stage reads a column as 8-byte
// (luma) or 4-byte (chroma) vectors.
struct alignas(16) McLds {
    uint32_t taps[kTapWords];  // packed filter pairs, copied from kF
    int16_t ltmp[2][4][8 * 16];
    int16_t ctmp[2][4][2][4 * 8];
};

// Motion compensation of four blocks of at most 8x8 (predict_block,
// src/hevc/inter.rs:555-742, with interp_scalar and the combiners,
// src/dsp/hevc.rs:823-1087). Lanes 16q..16q+15 serve block first + q.
// First stage: lane task (row, 4-column group) of the luma support (bh + 7
// rows) and (component, row) of the chroma support (ch + 3 rows), read
// straight from the reference planes. Second stage: luma lane (x = s & 7,
// rows 4 (s >> 3) ..+3), chroma lane (component s >> 3, x = s & 3, rows
// 2 ((s >> 2) & 1) ..+1). Every lane reaches the barrier.
// QUAD: the four blocks are the 8x8 quadrants of 16x16 tile `first` (a
// window of 15 x 15 luma samples per quadrant instead of one shared
// 23 x 23 window).
template <typename S, int LB, bool QUAD = false>
__device__ __forceinline__ void mc_item(PicRef P, int first, int tid, McLds& L) {
    first = __builtin_amdgcn_readfirstlane(first);
    const int slot = tid >> 4, s = tid & 15;
    McTile t;
    if constexpr (QUAD) {
        t = __builtin_bit_cast(McTile, *(const uint4*)(P.tiles16 + first));
        t.x = (uint16_t)(t.x + (slot & 1) * 8);
        t.y = (uint16_t)(t.y + (slot >> 1) * 8);
        t.tw = t.th = 8;
    } else {
        // The wave's four records (uniform loads: the list is padded to a
        // multiple of four with empty records, dir 0).
        const uint4* rec = (const uint4*)(P.tiles + first);
        const uint4 q0 = rec[0], q1 = rec[1], q2 = rec[2], q3 = rec[3];
        const uint4 q = slot == 0 ? q0 : slot == 1 ? q1 : slot == 2 ? q2 : q3;
        t = __builtin_bit_cast(McTile, q);
    }
    // The packed taps, read per lane below: an LDS copy costs two coalesced
    // loads per wave instead of one per lane and tap set. Issued first,
    // written to LDS once the reference loads are in flight.
    const uint32_t* kt = &kF.qpk[0][0];
    const uint32_t tap0 = kt[tid], tap1 = tid + 64 < kTapWords ? kt[tid + 64] : 0u;
    const int bw = t.tw, bh = t.th, dir = t.dir, bd = P.bd;
    const int shift1 = min(bd, 12) - 8;
    int ri[2], mvx[2], mvy[2];
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        ri[l] = (t.refs >> (4 * l)) & 15;
        mvx[l] = t.mv[l].x;
        mvy[l] = t.mv[l].y;
    }
    // 1. First (horizontal) stage. The spans of LB lists at a time (both,
    // or one when registers are short) are located and their loads issued
    // before any is filtered: one memory round trip per group of lists.
    const int pw0 = P.pw[0], pw1 = P.pw[1];
    const int nt = (bh + 7) * 2, ng = bw >> 2, nr = (bh >> 1) + 3;
    const int cc = s >= nr ? 1 : 0, cr = s - nr * cc;  // chroma task: component, row
#pragma unroll
    for (int g0 = 0; g0 < 2; g0 += LB) {
        Span<S, 12> ls[LB][2];
        Span<S, 8> cs[LB];
        bool lact[LB][2], cact[LB];
        uint32_t ld[LB][2][Span<S, 12>::ND], cd[LB][Span<S, 8>::ND];
#pragma unroll
        for (int k = 0; k < LB; ++k) {
            const int l = g0 + k;
            const bool use = (dir >> l) & 1;
            const S* ref = (const S*)ref_plane(P, l, ri[l], 0);
            const int x0 = t.x + (mvx[l] >> 2) - 3, y0 = t.y + (mvy[l] >> 2) - 3;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int e = s + 16 * j, r = e >> 1, g = e & 1;
                lact[k][j] = use && e < nt && g < ng;
                ls[k][j] = span_at<S, 12>(ref, P.pitch[0], pw0, P.ph[0], x0 + 4 * g, y0 + r);
            }
            const S* cref = (const S*)(cc ? ref_plane(P, l, ri[l], 2) : ref_plane(P, l, ri[l], 1));
            cact[k] = use && s < 2 * nr;
            cs[k] = span_at<S, 8>(cref, P.pitch[1], pw1, P.ph[1], (t.x >> 1) + (mvx[l] >> 3) - 1,
                                  (t.y >> 1) + (mvy[l] >> 3) - 1 + cr);
        }
        // Left unset for inactive tasks (never read): a zero default would
        // need a register copy of each loaded value, i.e. a wait.
#pragma unroll
        for (int k = 0; k < LB; ++k) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (lact[k][j]) {
#pragma unroll
                    for (int i = 0; i < Span<S, 12>::ND; ++i) ld[k][j][i] = ls[k][j].w[i];
                }
            }
            if (cact[k]) {
#pragma unroll
                for (int i = 0; i < Span<S, 8>::ND; ++i) cd[k][i] = cs[k].w[i];
            }
        }
        if (g0 == 0) {
            L.taps[tid] = tap0;
            if (tid + 64 < kTapWords) L.taps[tid + 64] = tap1;
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < LB; ++k) {
            const int l = g0 + k;
            if (!((dir >> l) & 1)) continue;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (lact[k][j]) {
                    const int e = s + 16 * j, r = e >> 1, g = e & 1;
                    int acc[4];
                    first_luma<S>(ls[k][j], ld[k][j], pw0, L.taps, mvx[l] & 3, acc);
#pragma unroll
                    for (int i = 0; i < 4; ++i) L.ltmp[l][slot][(4 * g + i) * 16 + r] = (int16_t)(acc[i] >> shift1);
                }
            }
            if (cact[k]) {
                int acc[4];
                first_chroma<S>(cs[k], cd[k], pw1, L.taps, mvx[l] & 7, acc);
#pragma unroll
                for (int i = 0; i < 4; ++i) L.ctmp[l][slot][cc][i * 8 + cr] = (int16_t)(acc[i] >> shift1);
            }
        }
    }
    __syncthreads();
    // 2. Second (vertical) stage, combine, store.
    const bool weighted = P.weighted != 0;
    {
        const int x = s & 7, y0 = (s >> 3) * 4;
        int pred[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
        for (int l = 0; l < 2; ++l) {
            if (!((dir >> l) & 1)) continue;
            const Taps8 tp = luma_taps(L.taps, mvy[l] & 3);
            short2v p[6];
            load_pairs(L.ltmp[l][slot] + x * 16 + y0, p);
            int acc[4];
            filt4(p, tp, acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) pred[l][i] = (int16_t)(acc[i] >> 6);
        }
        if (dir && x < bw && y0 < bh) {
            Wp wp = {{1, 1}, {0, 0}};
            if (weighted) wp = weights(P, 0, ri);
            const int lwd = P.log2_wd[0];
            S* dst = (S*)P.cur[0] + (size_t)(t.y + y0) * P.pitch[0] + t.x + x;
#pragma unroll
            for (int k = 0; k < 4; ++k) dst[k * P.pitch[0]] = (S)combine(weighted, lwd, wp, dir, pred[0][k], pred[1][k], bd);
        }
    }
    // Chroma (4:2:0): eighth-sample vectors in chroma units
    // (src/hevc/inter.rs:581-594).
    {
        const int c = s >> 3, x = s & 3, y0 = ((s >> 2) & 1) * 2;
        int pred[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
        for (int l = 0; l < 2; ++l) {
            if (!((dir >> l) & 1)) continue;
            const Taps4 tp = chroma_taps(L.taps, mvy[l] & 7);
            const uint32_t* T = (const uint32_t*)(L.ctmp[l][slot][c] + x * 8 + y0);
            const short2v p0 = as2(T[0]), p1 = as2(T[1]), p2 = as2(T[2]);
            pred[l][0] = (int16_t)(dot2(p1, tp.e[1], dot2(p0, tp.e[0], 0)) >> 6);
            pred[l][1] = (int16_t)(dot2(p2, tp.o[2], dot2(p1, tp.o[1], dot2(p0, tp.o[0], 0))) >> 6);
        }
        if (dir && x < (bw >> 1) && y0 < (bh >> 1)) {
            Wp wp = {{1, 1}, {0, 0}};
            if (weighted) {
                const Wp cb = weights(P, 1, ri), cr = weights(P, 2, ri);
                wp = c ? cr : cb;
            }
            const int lwd = c ? P.log2_wd[2] : P.log2_wd[1];
            const int pitch = P.pitch[1];  // Cb and Cr share the geometry
            S* dst = (S*)(c ? P.cur[2] : P.cur[1]) + (size_t)((t.y >> 1) + y0) * pitch + (t.x >> 1) + x;
#pragma unroll
            for (int k = 0; k < 2; ++k) dst[k * pitch] = (S)combine(weighted, lwd, wp, dir, pred[0][k], pred[1][k], bd);
        }
    }
}

// LDS of one 16x16 tile: first-stage outputs of its shared window,
// transposed: luma [list][column 0..15][row 0..22, one pad], chroma
// [list][component][column 0..7][row 0..10, one pad].
struct alignas(16) McLds16 {
    uint32_t taps[kTapWords];
    int16_t lt[2][16 * 24];
    int16_t ct[2][2][8 * 12];
};

typedef const uint4 __attribute__((address_space(4))) ConstU4;

// Motion compensation of one 16x16 tile of a PU (predict_block,
// src/hevc/inter.rs:555-742, with interp_scalar and the combiners,
// src/dsp/hevc.rs:823-1087), the same arithmetic as mc_item over one
// window: luma first stage 23 rows x 4 groups of 4 columns (task e = tid,
// tid + 64), chroma 2 components x 11 rows x 2 groups (task e = tid); the
// second stage gives lane (x = tid & 15, rows 4 (tid >> 4) ..+3) of luma
// and lane (component tid >> 5, x = tid & 7, rows 2 ((tid >> 3) & 3) ..+1)
// of chroma. Every lane reaches the barriers.
template <typename S>
__device__ __forceinline__ void mc_tile16(PicRef P, int k, int tid, McLds16& L) {
    k = __builtin_amdgcn_readfirstlane(k);
    const McTile t = __builtin_bit_cast(McTile, *(ConstU4*)(P.tiles16 + k));
    const uint32_t* kt = &kF.qpk[0][0];
    const uint32_t tap0 = kt[tid], tap1 = tid + 64 < kTapWords ? kt[tid + 64] : 0u;
    const int dir = t.dir, bd = P.bd;
    const int shift1 = min(bd, 12) - 8;
    int ri[2], mvx[2], mvy[2];
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        ri[l] = (t.refs >> (4 * l)) & 15;
        mvx[l] = t.mv[l].x;
        mvy[l] = t.mv[l].y;
    }
    const int pw0 = P.pw[0], pw1 = P.pw[1];
    const int cc = tid >= 22 ? 1 : 0, crem = tid - 22 * cc, cr = crem >> 1, cg = crem & 1;
    Span<S, 12> ls[2][2];
    Span<S, 8> cs[2];
    bool lact[2][2], cact[2];
    uint32_t ld[2][2][Span<S, 12>::ND], cd[2][Span<S, 8>::ND];
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        const bool use = (dir >> l) & 1;
        const S* ref = (const S*)ref_plane(P, l, ri[l], 0);
        const int x0 = t.x + (mvx[l] >> 2) - 3, y0 = t.y + (mvy[l] >> 2) - 3;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int e = tid + 64 * j, r = e >> 2, g = e & 3;
            lact[l][j] = use && e < 92;
            ls[l][j] = span_at<S, 12>(ref, P.pitch[0], pw0, P.ph[0], x0 + 4 * g, y0 + r);
        }
        const S* cref = (const S*)(cc ? ref_plane(P, l, ri[l], 2) : ref_plane(P, l, ri[l], 1));
        cact[l] = use && tid < 44;
        cs[l] = span_at<S, 8>(cref, P.pitch[1], pw1, P.ph[1], (t.x >> 1) + (mvx[l] >> 3) - 1 + 4 * cg,
                              (t.y >> 1) + (mvy[l] >> 3) - 1 + cr);
    }
    // All loads issued before any is used; unset for inactive tasks.
#pragma unroll
    for (int l = 0; l < 2; ++l) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (lact[l][j]) {
#pragma unroll
                for (int i = 0; i < Span<S, 12>::ND; ++i) ld[l][j][i] = ls[l][j].w[i];
            }
        }
        if (cact[l]) {
#pragma unroll
            for (int i = 0; i < Span<S, 8>::ND; ++i) cd[l][i] = cs[l].w[i];
        }
    }
    L.taps[tid] = tap0;
    if (tid + 64 < kTapWords) L.taps[tid + 64] = tap1;
    __syncthreads();
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        if (!((dir >> l) & 1)) continue;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (lact[l][j]) {
                const int e = tid + 64 * j, r = e >> 2, g = e & 3;
                int acc[4];
                first_luma<S>(ls[l][j], ld[l][j], pw0, L.taps, mvx[l] & 3, acc);
#pragma unroll
                for (int i = 0; i < 4; ++i) L.lt[l][(4 * g + i) * 24 + r] = (int16_t)(acc[i] >> shift1);
            }
        }
        if (cact[l]) {
            int acc[4];
            first_chroma<S>(cs[l], cd[l], pw1, L.taps, mvx[l] & 7, acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) L.ct[l][cc][(4 * cg + i) * 12 + cr] = (int16_t)(acc[i] >> shift1);
        }
    }
    __syncthreads();
    const bool weighted = P.weighted != 0;
    {
        const int x = tid & 15, y0 = (tid >> 4) * 4;
        int pred[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
        for (int l = 0; l < 2; ++l) {
            if (!((dir >> l) & 1)) continue;
            const Taps8 tp = luma_taps(L.taps, mvy[l] & 3);
            short2v p[6];
            load_pairs(L.lt[l] + x * 24 + y0, p);
            int acc[4];
            filt4(p, tp, acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) pred[l][i] = (int16_t)(acc[i] >> 6);
        }
        if (dir) {
            Wp wp = {{1, 1}, {0, 0}};
            if (weighted) wp = weights(P, 0, ri);
            const int lwd = P.log2_wd[0];
            S* dst = (S*)P.cur[0] + (size_t)(t.y + y0) * P.pitch[0] + t.x + x;
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2)
                dst[k2 * P.pitch[0]] = (S)combine(weighted, lwd, wp, dir, pred[0][k2], pred[1][k2], bd);
        }
    }
    {
        const int c = tid >> 5, x = tid & 7, y0 = ((tid >> 3) & 3) * 2;
        int pred[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
        for (int l = 0; l < 2; ++l) {
            if (!((dir >> l) & 1)) continue;
            const Taps4 tp = chroma_taps(L.taps, mvy[l] & 7);
            const uint32_t* T = (const uint32_t*)(L.ct[l][c] + x * 12 + y0);
            const short2v p0 = as2(T[0]), p1 = as2(T[1]), p2 = as2(T[2]);
            pred[l][0] = (int16_t)(dot2(p1, tp.e[1], dot2(p0, tp.e[0], 0)) >> 6);
            pred[l][1] = (int16_t)(dot2(p2, tp.o[2], dot2(p1, tp.o[1], dot2(p0, tp.o[0], 0))) >> 6);
        }
        if (dir) {
            Wp wp = {{1, 1}, {0, 0}};
            if (weighted) {
                const Wp cb = weights(P, 1, ri), cr2 = weights(P, 2, ri);
                wp = c ? cr2 : cb;
            }
            const int lwd = c ? P.log2_wd[2] : P.log2_wd[1];
            const int pitch = P.pitch[1];
            S* dst = (S*)(c ? P.cur[2] : P.cur[1]) + (size_t)((t.y >> 1) + y0) * pitch + (t.x >> 1) + x;
#pragma unroll
            for (int k2 = 0; k2 < 2; ++k2) dst[k2 * pitch] = (S)combine(weighted, lwd, wp, dir, pred[0][k2], pred[1][k2], bd);
        }
    }
}

// Side data (src/hevc/ctu.rs:751-800 for CUs and PUs, :1493-1510 for
// TUs): per 4x4 cell, a CU writes QpY and its intra / edge flags (and, if
// intra, the no-PU marker with edge flags into pumap), a PU its index and
// edge flags into pumap, a luma TU its cbf and edge flags. One work item
// writes kSideCellsPerItem consecutive cells of one class, 64 per round;
// each round's entry records are requested before any cell is written.
typedef const SideClass __attribute__((address_space(4))) ConstSide;

__device__ __forceinline__ void side_item(PicRef P, int item, int tid) {
    ConstSide* cl = (ConstSide*)P.sclass;
    const int lo = ((const uint16_t __attribute__((address_space(4)))*)P.sidx)[item];
    const int kind = cl[lo].kind, first = cl[lo].first, count = cl[lo].count, cw = cl[lo].cw, ch = cl[lo].ch;
    const int cells = cw * ch, base = (item - cl[lo].item0) * kSideCellsPerItem;
    constexpr int R = kSideCellsPerItem / 64;
    int e[R], cx[R], cy[R];
    bool on[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int g = base + r * 64 + tid;
        e[r] = g / cells;
        on[r] = e[r] < count;
        const int cell = g - e[r] * cells;
        cy[r] = cell / cw;
        cx[r] = cell - cy[r] * cw;
    }
    // Entry records as dwords: CU {x | y << 16, log2 | qp << 8 | intra << 16},
    // PU {x | y << 16}, TU {x | y << 16, comp | log2 << 8 | flags << 16}.
    const uint32_t* recs = kind == SIDE_CU ? (const uint32_t*)P.cus
                           : kind == SIDE_PU ? (const uint32_t*)P.pus : (const uint32_t*)P.ltus;
    const int stride = kind == SIDE_CU ? 5 : kind == SIDE_PU ? 5 : 4;  // dwords per record
    uint32_t w0[R], w1[R], wx[R];
    const bool lf = kind == SIDE_CU && P.ctbw != nullptr;  // exemptions in rv_hevc_cu dword 4
#pragma unroll
    for (int r = 0; r < R; ++r) {
        wx[r] = 0;
        if (on[r]) {
            const uint32_t* rp = recs + (size_t)(first + e[r]) * stride;
            w0[r] = rp[0];
            w1[r] = kind == SIDE_PU ? 0u : rp[1];
            if (lf) wx[r] = rp[4];
        }
    }
    const int w4 = P.w4;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!on[r]) continue;
        const int bx = (int)(w0[r] & 0xFFFF) / 4 + cx[r], by = (int)(w0[r] >> 16) / 4 + cy[r];
        const size_t q = (size_t)by * w4 + bx;
        const uint32_t left = cx[r] == 0, top = cy[r] == 0;
        if (kind == SIDE_CU) {
            const uint32_t intra = ((w1[r] >> 16) & 0xFF) ? 1u : 0u;
            P.qp[q] = (int8_t)((w1[r] >> 8) & 0xFF);
            const uint32_t ex = (wx[r] >> 8) & 0xFF;  // rv_hevc_cu.exempt
            const uint32_t exb = ((ex & RV_CU_LF_EXEMPT) ? kCuSaoExempt : 0u) |
                                 ((ex & (RV_CU_LF_EXEMPT | RV_CU_DBK_EXEMPT)) ? kCuDbkExempt : 0u);
            P.cuinfo[q] = (uint8_t)(intra | (left << 1) | (top << 2) | exb);
            if (intra) P.pumap[q] = 0x3FFFFFFFu | (left << 30) | (top << 31);
        } else if (kind == SIDE_PU) {
            P.pumap[q] = (uint32_t)(first + e[r]) | (left << 30) | (top << 31);
        } else {
            const uint32_t cbf = ((w1[r] >> 16) & RV_TU_CODED) ? 1u : 0u;
            P.tuinfo[q] = (uint8_t)(cbf | (left << 1) | (top << 2));
        }
    }
}

// QUAD: 16x16 tiles as four 8x8 quadrants through mc_item (a smaller
// window per block, no shared one): a tile and window shape the tuner
// weighs against the shared 16x16 window.
template <int D, int W, int LB = 2, bool QUAD = false>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W))) hevc_stage0(BatchDev b) {
    __shared__ union {
        McLds blocks;
        McLds16 tile;
    } lds;
    ConstI32* pre = prefix(b, 0);
    const int it = xcd_item();
    if (it >= pre[b.n]) return;
    // Uniform by construction; readfirstlane keeps the compiler's
    // uniformity analysis (and scalar loads of the descriptor) on it.
    const int p = __builtin_amdgcn_readfirstlane(find_pic(pre, b.n, it));
    PicRef P = picture(b, p);
    int i = __builtin_amdgcn_readfirstlane(it - pre[p]);
    const int tid = threadIdx.x;
    if (i < P.n_side) {
        side_item(P, i, tid);
        return;
    }
    i -= P.n_side;
    if (i < P.n_intra) {
        const rv_hevc_cu cu = P.cus[P.intra_idx[i]];
        if (eight_bit<D>(P.bd)) intra_copy<uint8_t>(P, cu, tid);
        else intra_copy<uint16_t>(P, cu, tid);
        return;
    }
    i -= P.n_intra;
    if (i < P.n_tiles16) {
        if constexpr (QUAD) {
            if (eight_bit<D>(P.bd)) mc_item<uint8_t, LB, true>(P, i, tid, lds.blocks);
            else mc_item<uint16_t, LB, true>(P, i, tid, lds.blocks);
        } else {
            if (eight_bit<D>(P.bd)) mc_tile16<uint8_t>(P, i, tid, lds.tile);
            else mc_tile16<uint16_t>(P, i, tid, lds.tile);
        }
        return;
    }
    i -= P.n_tiles16;
    if (eight_bit<D>(P.bd)) mc_item<uint8_t, LB>(P, 4 * i, tid, lds.blocks);
    else mc_item<uint16_t, LB>(P, 4 * i, tid, lds.blocks);
}

// ---------------------------------------------------------------------
// Stage 0, strips: motion compensation of the blocks of 8-bit pictures
// whose reference windows lie inside the reference planes, one lane per
// strip of 4 output columns (McStrip). The lane walks its window row by row: each row's 4
// first-stage outputs (interp_scalar's horizontal pass, src/dsp/hevc.rs:
// 823-882) come from one aligned load, and rows are kept in registers as
// i16 pairs (row 2k, row 2k + 1) of each column, so the vertical pass is
// v_dot2 over pairs: output row 2m takes pairs m..m+3 with the even taps
// (f0 f1)..(f6 f7), row 2m + 1 pairs m..m+4 with the odd taps (0 f0)..
// (f7 0) (chroma: 2 and 3 pairs). An output row is finished, combined
// over both lists and stored as one dword (8-bit) as soon as its pairs
// exist, so only the last 5 (chroma 3) pairs of each list stay live.
// ---------------------------------------------------------------------

// Reference window of one list of one strip lane (8-bit samples): the
// dword holding its first sample (row 0), the sample's byte offset in it,
// the row pitch in dwords and the two phases.
struct StripSrc {
    const uint32_t* w;
    uint32_t sh;
    int pitch4;
    int fx, fy;
};
__device__ __forceinline__ StripSrc strip_src(PicRef P, int l, int ri, int c, int x0, int y0, int fx, int fy) {
    const char* base = (const char*)ref_plane(P, l, ri, c);
    const int pitch_b = P.pitch[c];
    const size_t off = (size_t)y0 * pitch_b + (size_t)x0;
    StripSrc s;
    s.w = (const uint32_t*)(base + (off & ~(size_t)3));
    s.sh = (uint32_t)off & 3u;
    s.pitch4 = pitch_b >> 2;
    s.fx = fx;
    s.fy = fy;
    return s;
}

// The dwords of one window row of a strip: 11 (luma) or 7 (chroma) samples
// after a byte shift of at most 3.
template <bool LUMA>
struct StripRow {
    static constexpr int ND = LUMA ? 4 : 3;
    uint32_t d[ND];
};
template <bool LUMA>
__device__ __forceinline__ StripRow<LUMA> strip_load(const StripSrc& s, int r) {
    const uint32_t* w = s.w + (size_t)r * s.pitch4;
    StripRow<LUMA> v;
#pragma unroll
    for (int i = 0; i < StripRow<LUMA>::ND; ++i) v.d[i] = w[i];
    return v;
}

// First stage of one loaded row of a strip: 4 outputs (exact at 8 bits,
// shift1 = 0) from luma (8-tap) or chroma (4-tap) samples.
template <bool LUMA>
__device__ __forceinline__ void strip_row(const StripSrc& s, const StripRow<LUMA>& v, const uint32_t* tab,
                                          int (&acc)[4]) {
    const uint32_t* d = v.d;
    if constexpr (LUMA) {
        const uint32_t x[3] = {__builtin_amdgcn_alignbyte(d[1], d[0], s.sh),
                               __builtin_amdgcn_alignbyte(d[2], d[1], s.sh),
                               __builtin_amdgcn_alignbyte(d[3], d[2], s.sh)};
        filt4_b(x, luma_taps_b(tab, s.fx), acc);
    } else {
        const uint32_t x[2] = {__builtin_amdgcn_alignbyte(d[1], d[0], s.sh),
                               __builtin_amdgcn_alignbyte(d[2], d[1], s.sh)};
        filt4c_b(x, chroma_taps_b(tab, s.fx), acc);
    }
}

// Low halves of lo (row 2k) and hi (row 2k + 1) as one i16 pair.
__device__ __forceinline__ uint32_t pair_rows(int lo, int hi) {
    return __builtin_amdgcn_perm((uint32_t)hi, (uint32_t)lo, 0x05040100u);
}

// One strip (LUMA: H rows of 8-tap; else H rows of 4-tap chroma), both
// lists when BI, the one in st.dir otherwise.
template <bool LUMA, int H, bool BI, int PF>
__device__ __forceinline__ void mc_strip(PicRef P, const McStrip& st, const uint32_t* tab) {
    constexpr int T = LUMA ? 8 : 4;       // taps
    constexpr int R = H + T - 1;          // window rows (odd)
    constexpr int RING = T / 2 + 1;       // live row pairs
    constexpr int NL = BI ? 2 : 1;
    const int bd = 8;
    const int c = LUMA ? 0 : st.comp;
    const int dir = st.dir;
    int ri[2] = {st.refs & 15, st.refs >> 4};
    StripSrc src[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int l = BI ? k : (dir >> 1);
        const int mvx = st.mv[l].x, mvy = st.mv[l].y;
        if constexpr (LUMA)
            src[k] = strip_src(P, l, ri[l], 0, st.x + (mvx >> 2) - 3, st.y + (mvy >> 2) - 3, mvx & 3, mvy & 3);
        else
            src[k] = strip_src(P, l, ri[l], c, st.x + (mvx >> 3) - 1, st.y + (mvy >> 3) - 1, mvx & 7, mvy & 7);
    }
    const bool weighted = P.weighted != 0;
    Wp wp = {{1, 1}, {0, 0}};
    if (weighted) wp = weights(P, c, ri);
    const int lwd = P.log2_wd[c];
    const int pitch = P.pitch[c];
    uint8_t* dst = (uint8_t*)P.cur[c] + (size_t)st.y * pitch + st.x;
    const int cols = st.cols;

    uint32_t ring[NL][RING][4];
    // Output row y of list k from the pairs in the ring, as an i16 value;
    // the second-stage taps (even pairs e, odd pairs o) are read from LDS
    // per row rather than held in registers.
    auto vert = [&](int k, int y, int (&out)[4]) {
        const int m = y >> 1;
        short2v tp[T / 2 + 1];
        if constexpr (LUMA) {
            const Taps8 t = luma_taps(tab, src[k].fy);
#pragma unroll
            for (int j = 0; j <= T / 2; ++j) tp[j] = (y & 1) ? t.o[j] : (j < T / 2 ? t.e[j] : t.o[0]);
        } else {
            const Taps4 t = chroma_taps(tab, src[k].fy);
#pragma unroll
            for (int j = 0; j <= T / 2; ++j) tp[j] = (y & 1) ? t.o[j] : (j < T / 2 ? t.e[j] : t.o[0]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int s = 0;
            if ((y & 1) == 0) {
#pragma unroll
                for (int j = 0; j < T / 2; ++j) s = dot2(as2(ring[k][(m + j) % RING][i]), tp[j], s);
            } else {
#pragma unroll
                for (int j = 0; j <= T / 2; ++j) s = dot2(as2(ring[k][(m + j) % RING][i]), tp[j], s);
            }
            out[i] = (int16_t)(s >> 6);
        }
    };
    auto emit = [&](int y) {
        int a[4], bq[4] = {0, 0, 0, 0};
        vert(0, y, a);
        if constexpr (BI) vert(1, y, bq);
        int v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = BI ? combine(weighted, lwd, wp, 3, a[i], bq[i], bd)
                                              : combine(weighted, lwd, wp, dir, a[i], a[i], bd);
        uint8_t* d = dst + (size_t)y * pitch;
        const uint32_t pk = (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
        if (cols == 4) *(uint32_t*)d = pk;
        else if (cols == 2) *(uint16_t*)d = (uint16_t)pk;
    };
    // Row pairs in a software pipeline: the loads of pair kp + PF are
    // issued before pair kp is filtered (the scheduling barriers keep the
    // compiler from sinking them to their use), so each wave keeps PF pairs
    // of every list in flight.
    constexpr int NPAIR = (R + 1) / 2;
    StripRow<LUMA> buf[PF + 1][NL][2];
    auto load_pair = [&](int kp, StripRow<LUMA> (&into)[NL][2]) {
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            into[k][0] = strip_load<LUMA>(src[k], 2 * kp);
            if (2 * kp + 1 < R) into[k][1] = strip_load<LUMA>(src[k], 2 * kp + 1);
        }
    };
#pragma unroll
    for (int kp = 0; kp < PF && kp < NPAIR; ++kp) load_pair(kp, buf[kp]);
#pragma unroll
    for (int kp = 0; kp < NPAIR; ++kp) {
        if (kp + PF < NPAIR) load_pair(kp + PF, buf[(kp + PF) % (PF + 1)]);
        __builtin_amdgcn_sched_barrier(0);
        StripRow<LUMA>(&cur)[NL][2] = buf[kp % (PF + 1)];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            int a0[4], a1[4] = {0, 0, 0, 0};
            strip_row<LUMA>(src[k], cur[k][0], tab, a0);
            if (2 * kp + 1 < R) strip_row<LUMA>(src[k], cur[k][1], tab, a1);
#pragma unroll
            for (int i = 0; i < 4; ++i) ring[k][kp % RING][i] = pair_rows(a0[i], a1[i]);
        }
        // Pair kp is complete: output rows 2kp - T + 1 (odd) and 2kp - T + 2
        // (even) have all their pairs.
        const int yo = 2 * kp - T + 1, ye = 2 * kp - T + 2;
        if (yo >= 1 && yo < H) emit(yo);
        if (ye >= 0 && ye < H) emit(ye);
        __builtin_amdgcn_sched_barrier(0);
    }
}

template <int PF>
__device__ __forceinline__ void strip_group(PicRef P, int g, const McStrip& st, const uint32_t* tab) {
    switch (g) {
        case 0: mc_strip<true, 16, true, PF>(P, st, tab); break;
        case 1: mc_strip<true, 16, false, PF>(P, st, tab); break;
        case 2: mc_strip<true, 8, true, PF>(P, st, tab); break;
        case 3: mc_strip<true, 8, false, PF>(P, st, tab); break;
        case 4: mc_strip<true, 4, true, PF>(P, st, tab); break;
        case 5: mc_strip<true, 4, false, PF>(P, st, tab); break;
        case 6: mc_strip<false, 8, true, PF>(P, st, tab); break;
        case 7: mc_strip<false, 8, false, PF>(P, st, tab); break;
        case 8: mc_strip<false, 4, true, PF>(P, st, tab); break;
        case 9: mc_strip<false, 4, false, PF>(P, st, tab); break;
        case 10: mc_strip<false, 2, true, PF>(P, st, tab); break;
        default: mc_strip<false, 2, false, PF>(P, st, tab); break;
    }
}

// One wave per 64 strips of one group of one 8-bit picture (b.spre: prefix
// over pictures of their strip waves). W waves per SIMD, PF row pairs of
// each list in flight per lane (PF 2 at W 3 fits in 155 VGPRs; PF 3 needs
// W 2, which leaves too few waves to cover the loads).
template <int W, int PF = 2>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W))) hevc_mc_strips(BatchDev b) {
    __shared__ uint32_t taps[kTapWords];
    ConstI32* spre = (ConstI32*)b.spre;
    const int it = xcd_item();
    if (it >= spre[b.n]) return;
    const int tid = threadIdx.x;
    const uint32_t* kt = &kF.qpk[0][0];
    taps[tid] = kt[tid];
    if (tid + 64 < kTapWords) taps[tid + 64] = kt[tid + 64];
    const int p = __builtin_amdgcn_readfirstlane(find_pic(spre, b.n, it));
    PicRef P = picture(b, p);
    const int w = __builtin_amdgcn_readfirstlane(it - spre[p]);
    const int g = __builtin_amdgcn_readfirstlane((int)((const uint8_t __attribute__((address_space(4)))*)P.swg)[w]);
    const McStrip st = __builtin_bit_cast(McStrip, *(const uint4*)(P.strips + (size_t)w * 64 + tid));
    __syncthreads();
    strip_group<PF>(P, g, st, taps);
}

// ---------------------------------------------------------------------
// Stage 1: inverse transform + residual add; boundary strengths
// ---------------------------------------------------------------------

// Magnitudes of the 8.6.4.2 matrix (its first column) and entry (k, n):
// +-mag(m), m = (2n + 1) k mod 128 folded by the cosine symmetries.
constexpr int mag32(int m) {
    return m == 0    ? 64
           : m <= 3  ? 90
           : m == 4  ? 89
           : m == 5  ? 88
           : m == 6  ? 87
           : m == 7  ? 85
           : m == 8  ? 83
           : m == 9  ? 82
           : m == 10 ? 80
           : m == 11 ? 78
           : m == 12 ? 75
           : m == 13 ? 73
           : m == 14 ? 70
           : m == 15 ? 67
           : m == 16 ? 64
           : m == 17 ? 61
           : m == 18 ? 57
           : m == 19 ? 54
           : m == 20 ? 50
           : m == 21 ? 46
           : m == 22 ? 43
           : m == 23 ? 38
           : m == 24 ? 36
           : m == 25 ? 31
           : m == 26 ? 25
           : m == 27 ? 22
           : m == 28 ? 18
           : m == 29 ? 13
           : m == 30 ? 9
           : m == 31 ? 4
                     : 0;
}
constexpr int tcoef(int k, int n) {
    return (((2 * n + 1) * k) & 127) <= 32   ? mag32(((2 * n + 1) * k) & 127)
           : (((2 * n + 1) * k) & 127) <= 64 ? -mag32(64 - (((2 * n + 1) * k) & 127))
           : (((2 * n + 1) * k) & 127) <= 96 ? -mag32((((2 * n + 1) * k) & 127) - 64)
                                             : mag32(128 - (((2 * n + 1) * k) & 127));
}

// idct1 (src/dsp/hevc.rs:474-497): the even/odd partial butterfly over all
// N inputs. The reference stops at the nonzero bound; zeros beyond it add
// nothing, so the dense butterfly gives the same sums. Unrolled, every
// matrix entry is an immediate.
template <int N>
__device__ __forceinline__ void idct1d(const int (&s)[N], int (&o)[N]) {
    if constexpr (N == 1) {
        o[0] = 64 * s[0];
    } else {
        constexpr int H = N / 2, STEP = 32 / N;
        int ev[H], eo[H];
#pragma unroll
        for (int k = 0; k < H; ++k) ev[k] = s[2 * k];
        idct1d<H>(ev, eo);
#pragma unroll
        for (int k = 0; k < H; ++k) {
            int acc = 0;
#pragma unroll
            for (int j = 1; j < N; j += 2) acc += tcoef(j * STEP, k) * s[j];
            o[k] = eo[k] + acc;
            o[N - 1 - k] = eo[k] - acc;
        }
    }
}

// idst4_scalar's 1-D step (src/dsp/hevc.rs:529-556): o[i] = sum_j M[j][i] s[j].
__device__ __forceinline__ void idst1d(const int (&s)[4], int (&o)[4]) {
    o[0] = 29 * s[0] + 74 * s[1] + 84 * s[2] + 55 * s[3];
    o[1] = 55 * s[0] + 74 * s[1] - 29 * s[2] - 84 * s[3];
    o[2] = 74 * s[0] + 0 * s[1] - 74 * s[2] + 74 * s[3];
    o[3] = 84 * s[0] - 74 * s[1] + 55 * s[2] - 29 * s[3];
}

// Add one row of N residuals to N samples at d (16 * k or 8 bytes,
// aligned), clipping to [0, max] (add_residual_scalar, src/dsp/hevc.rs:558-566).
template <typename S, int N>
__device__ __forceinline__ void add_row(S* d, const int (&r)[N], int max) {
    constexpr int BYTES = N * (int)sizeof(S), W = BYTES / 4, SPW = 4 / (int)sizeof(S);
    constexpr uint32_t MASK = sizeof(S) == 1 ? 0xFFu : 0xFFFFu;
    uint32_t w[W];
    if constexpr (BYTES == 8) {
        const uint2 u = *(const uint2*)d;
        w[0] = u.x;
        w[1] = u.y;
    } else {
#pragma unroll
        for (int v = 0; v < W / 4; ++v) {
            const uint4 u = ((const uint4*)d)[v];
            w[4 * v] = u.x;
            w[4 * v + 1] = u.y;
            w[4 * v + 2] = u.z;
            w[4 * v + 3] = u.w;
        }
    }
    uint32_t o[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        o[k] = 0;
#pragma unroll
        for (int e = 0; e < SPW; ++e) {
            const int sh = e * 8 * (int)sizeof(S);
            o[k] |= (uint32_t)clampi((int)((w[k] >> sh) & MASK) + r[k * SPW + e], 0, max) << sh;
        }
    }
    if constexpr (BYTES == 8) {
        *(uint2*)d = make_uint2(o[0], o[1]);
    } else {
#pragma unroll
        for (int v = 0; v < W / 4; ++v) ((uint4*)d)[v] = make_uint4(o[4 * v], o[4 * v + 1], o[4 * v + 2], o[4 * v + 3]);
    }
}

// scale_coefficients (src/hevc/residual.rs:1223-1283) for the levels of
// one TU: d = Clip3(-32768, 32767, ((c * m * levelScale[qP % 6]) << (qP / 6)
// + (1 << (bdShift - 1))) >> bdShift), bdShift = BitDepth + log2 - 5. m is
// 16 without scaling lists, else the TU's list entry (src/hevc/ctu.rs:
// 2016-2029): raster for 4x4 and 8x8; for 16x16 / 32x32 the 8x8 grid entry
// covering (x, y), and dc at (0, 0). |c * m * levelScale| < 2^31 (|c| <=
// 2^15, m <= 255, levelScale <= 72): only the shift needs 64 bits.
struct LevelScale {
    int ls, q6, bd_shift, dc;
    const uint8_t* list;  // null: flat
};

__device__ __forceinline__ LevelScale level_scale(PicRef P, const rv_hevc_tu& tu, int log2) {
    LevelScale s;
    const int qp = tu.qp, q6 = qp / 6;
    // levelScale = {40, 45, 51, 57, 64, 72}, one byte each.
    s.ls = (int)((0x484039332D28ull >> (8 * (qp - 6 * q6))) & 0xFF);
    s.q6 = q6;
    s.bd_shift = P.bd + log2 - 5;
    s.list = nullptr;
    s.dc = 16;
    if (P.scaling) {
        const int mid = ((tu.flags & RV_TU_INTRA) ? 0 : 3) + tu.comp;
        s.list = log2 == 2 ? P.scaling->list4[mid] : P.scaling->list8[log2 - 3][mid];
        if (log2 >= 4) s.dc = P.scaling->dc[log2 - 4][mid];
    }
    return s;
}

template <int LOG2> __device__ __forceinline__ int16_t scale_level(const LevelScale& s, int c, int pos) {
    int m = 16;
    if (s.list) {
        if constexpr (LOG2 <= 3) {
            m = s.list[pos];
        } else {
            constexpr int R = LOG2 - 3;
            const int x = pos & ((1 << LOG2) - 1), y = pos >> LOG2;
            m = pos == 0 ? s.dc : s.list[((y >> R) << 3) + (x >> R)];
        }
    }
    const int64_t v = ((int64_t)(c * m * s.ls) << s.q6) + ((int64_t)1 << (s.bd_shift - 1));
    const int64_t d = v >> s.bd_shift;
    return (int16_t)(d < -32768 ? -32768 : (d > 32767 ? 32767 : d));
}

// One wave reconstructs 64 / L coded TUs of size N (N >= 8), L lanes per
// TU: idct_scalar (src/dsp/hevc.rs:499-527) — columns, round by 7 and clip
// to i16, rows, round by bdShift = 20 - BitDepth and clip — then the
// residual add. The TU's dense raster lives in LDS (rows of N, TUs padded
// apart so the column reads of one half-wave hit distinct banks): each
// lane zeroes 1/L of it, scatters every L-th nonzero, transforms C = N / L
// adjacent columns (one C-sample LDS read per row), then C whole rows
// (16-byte LDS reads), adding each row to the picture with 16-byte
// accesses.
template <int N, int L, typename S>
__device__ __forceinline__ void tu_lanes(PicRef P, int first, int count, int tid, int16_t* lds) {
    constexpr int C = N / L, PAD = N == 8 ? 8 : N == 16 ? 16 : 32, STRIDE = N * N + PAD;
    constexpr int ZV = N * N * 2 / L / 16;  // 16-byte zero stores per lane
    static_assert(C >= 2 && C <= 4 && ZV >= 1, "lane split");
    const int slot = tid / L, sub = tid % L;
    const bool active = slot < count;
    const rv_hevc_tu tu = P.ctus[first + (active ? slot : 0)];
    int16_t* t = lds + slot * STRIDE;
    uint4* tz = (uint4*)t + sub * ZV;
#pragma unroll
    for (int i = 0; i < ZV; ++i) tz[i] = make_uint4(0, 0, 0, 0);
    __syncthreads();
    if (active) {
        // Lane `sub` scatters its own contiguous run of the TU's entries, a
        // multiple of 4 long, 8 entries (two 16-byte loads) per round trip;
        // a chunk may read up to 3 entries past the run, still inside the
        // work-list arena.
        const uint32_t* z = (const uint32_t*)(P.nz + tu.coef_off);
        const int n = tu.n_nz, per = ((n + L - 1) / L + 3) & ~3;
        const int lo = sub * per, hi = min(lo + per, n);
        const bool levels = tu.flags & RV_TU_LEVELS;
        constexpr int LOG2 = N == 8 ? 3 : N == 16 ? 4 : 5;
        LevelScale sc;
        if (levels) sc = level_scale(P, tu, LOG2);
        for (int k = lo; k < hi; k += 8) {
            uint32_t v[8];
            if (k < hi) {
#pragma unroll
                for (int i = 0; i < 4; ++i) v[i] = z[k + i];
            }
            if (k + 4 < hi) {
#pragma unroll
                for (int i = 4; i < 8; ++i) v[i] = z[k + i];
            }
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (k + i < hi) {
                    const int pos = (v[i] & 0xFFFF) & (N * N - 1);
                    const int c = (int)(int16_t)(v[i] >> 16);
                    t[pos] = levels ? scale_level<LOG2>(sc, c, pos) : (int16_t)c;
                }
        }
    }
    __syncthreads();
    if (active) {
        int col[C][N];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const int16_t* src = t + j * N + sub * C;
            if constexpr (C == 4) {
                const uint2 u = *(const uint2*)src;
                col[0][j] = (int)(int16_t)(u.x & 0xFFFF);
                col[1][j] = (int)(int16_t)(u.x >> 16);
                col[2][j] = (int)(int16_t)(u.y & 0xFFFF);
                col[3][j] = (int)(int16_t)(u.y >> 16);
            } else {
                const uint32_t u = *(const uint32_t*)src;
                col[0][j] = (int)(int16_t)(u & 0xFFFF);
                col[1][j] = (int)(int16_t)(u >> 16);
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            int o[N];
            idct1d<N>(col[c], o);
#pragma unroll
            for (int y = 0; y < N; ++y) col[c][y] = clampi((o[y] + 64) >> 7, -32768, 32767);
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            int16_t* dst = t + j * N + sub * C;
            if constexpr (C == 4) {
                *(uint2*)dst = make_uint2((uint32_t)(col[0][j] & 0xFFFF) | ((uint32_t)col[1][j] << 16),
                                          (uint32_t)(col[2][j] & 0xFFFF) | ((uint32_t)col[3][j] << 16));
            } else {
                *(uint32_t*)dst = (uint32_t)(col[0][j] & 0xFFFF) | ((uint32_t)col[1][j] << 16);
            }
        }
    }
    __syncthreads();
    if (active) {
        const int bd_shift = 20 - P.bd, round2 = 1 << (bd_shift - 1);
        const int max = (1 << P.bd) - 1, pitch = P.pitch[tu.comp];
        S* d = (S*)P.cur[tu.comp] + (size_t)tu.y * pitch + tu.x;
        for (int r = 0; r < C; ++r) {
            const int y = sub * C + r;
            const uint4* src = (const uint4*)(t + y * N);
            int s_[N], o[N];
#pragma unroll
            for (int v = 0; v < N / 8; ++v) {
                const uint4 u = src[v];
                const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    s_[8 * v + 2 * e] = (int)(int16_t)(w[e] & 0xFFFF);
                    s_[8 * v + 2 * e + 1] = (int)(int16_t)(w[e] >> 16);
                }
            }
            idct1d<N>(s_, o);
#pragma unroll
            for (int x = 0; x < N; ++x) o[x] = clampi((o[x] + round2) >> bd_shift, -32768, 32767);
            add_row<S, N>(d + (size_t)y * pitch, o, max);
        }
    }
}

// 4x4 TUs, one per lane (64 per wave): idct_scalar / idst4_scalar
// (src/dsp/hevc.rs:499-556) and the residual add. The lane scatters its
// nonzero list into its own 32 bytes of LDS (no other lane touches them,
// so no barrier), reads the dense raster back as two 16-byte loads, runs
// both passes in registers and adds the residual one
// 4-sample row (one 4- or 8-byte access) at a time.
template <typename S>
__device__ __forceinline__ void tu4_lane(PicRef P, int first, int count, int tid, int16_t* lds) {
    if (tid >= count) return;
    const rv_hevc_tu tu = P.ctus[first + tid];
    int16_t* c = lds + tid * 16;
    uint4* c4 = (uint4*)c;
    c4[0] = make_uint4(0, 0, 0, 0);
    c4[1] = make_uint4(0, 0, 0, 0);
    // rv_nz is {uint16 pos, int16 val}: one dword, 4-byte aligned. All of
    // the TU's (at most 16) entries are requested at once, in 4-entry
    // chunks (one 16-byte load each; the last may read up to 3 entries past
    // the run, still inside the work-list arena), before any is stored.
    const uint32_t* z = (const uint32_t*)(P.nz + tu.coef_off);
    const int n = tu.n_nz;
    uint32_t v[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (4 * q < n) {
#pragma unroll
            for (int i = 0; i < 4; ++i) v[4 * q + i] = z[4 * q + i];
        }
    }
    if (tu.flags & RV_TU_LEVELS) {
        const LevelScale sc = level_scale(P, tu, 2);
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (i < n) c[v[i] & 15] = scale_level<2>(sc, (int)(int16_t)(v[i] >> 16), v[i] & 15);
    } else {
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (i < n) c[v[i] & 15] = (int16_t)(v[i] >> 16);
    }
    const uint4 r0 = c4[0], r1 = c4[1];
    const uint32_t w[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
    int m[16];  // raster: m[y * 4 + x]
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        m[2 * j] = (int)(int16_t)(w[j] & 0xFFFF);
        m[2 * j + 1] = (int)(int16_t)(w[j] >> 16);
    }
    const bool dst = tu.flags & RV_TU_DST;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        const int s_[4] = {m[x], m[4 + x], m[8 + x], m[12 + x]};
        int o[4];
        if (dst) idst1d(s_, o);
        else idct1d<4>(s_, o);
#pragma unroll
        for (int y = 0; y < 4; ++y) m[y * 4 + x] = clampi((o[y] + 64) >> 7, -32768, 32767);
    }
    const int bd_shift = 20 - P.bd, round2 = 1 << (bd_shift - 1);
    const int max = (1 << P.bd) - 1, pitch = P.pitch[tu.comp];
    S* d = (S*)P.cur[tu.comp] + (size_t)tu.y * pitch + tu.x;
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        const int s_[4] = {m[y * 4], m[y * 4 + 1], m[y * 4 + 2], m[y * 4 + 3]};
        int o[4];
        if (dst) idst1d(s_, o);
        else idct1d<4>(s_, o);
        int r[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) r[x] = clampi((o[x] + round2) >> bd_shift, -32768, 32767);
        S* row = d + (size_t)y * pitch;
        if constexpr (sizeof(S) == 1) {
            const uint32_t old = *(const uint32_t*)row;
            uint32_t nw = 0;
#pragma unroll
            for (int x = 0; x < 4; ++x) nw |= (uint32_t)clampi((int)((old >> (8 * x)) & 0xFF) + r[x], 0, max) << (8 * x);
            *(uint32_t*)row = nw;
        } else {
            const uint2 old = *(const uint2*)row;
            const uint32_t ow[2] = {old.x, old.y};
            uint32_t nw[2] = {0, 0};
#pragma unroll
            for (int x = 0; x < 4; ++x)
                nw[x >> 1] |= (uint32_t)clampi((int)((ow[x >> 1] >> (16 * (x & 1))) & 0xFFFF) + r[x], 0, max)
                              << (16 * (x & 1));
            *(uint2*)row = make_uint2(nw[0], nw[1]);
        }
    }
}

template <typename S>
__device__ __forceinline__ void tx_item(PicRef P, int i, int tid, int16_t* lds) {
    const uint32_t w = ((const uint32_t __attribute__((address_space(4)))*)P.txw)[i];
    const int cls = (int)(w >> 30), first = (int)(w & 0xFFFFFF), count = (int)((w >> 24) & 63) + 1;
    switch (cls) {
        case 0: tu4_lane<S>(P, first, count, tid, lds); break;
        case 1: tu_lanes<8, 64 / tus_per_item(1), S>(P, first, count, tid, lds); break;
        case 2: tu_lanes<16, 64 / tus_per_item(2), S>(P, first, count, tid, lds); break;
        default: tu_lanes<32, 64 / tus_per_item(3), S>(P, first, count, tid, lds); break;
    }
}

struct MotionRef {
    int uses[2];
    int delta[2], lt[2];
    int mv[2][2];
};

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

// The fields of a PU that motion_bs reads, from its dwords 1..4 (dir,
// ref_idx[2], mv[0], mv[1]); rv_hevc_pu is 20 bytes, so every record is
// 4-byte aligned. No PU (index past the list): zeros, i.e. no list used.
__device__ __forceinline__ uint4 pu_motion_words(PicRef P, uint32_t pm) {
    const uint32_t idx = pm & 0x3FFFFFFFu;
    if (idx >= (uint32_t)P.n_pus) return make_uint4(0, 0, 0, 0);
    const uint32_t* w = (const uint32_t*)P.pus + (size_t)idx * 5 + 1;
    return make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ __forceinline__ MotionRef motion_from(PicRef P, uint4 w) {
    MotionRef m;
    const int dir = (w.x >> 16) & 0xFF;
    for (int l = 0; l < 2; ++l) {
        m.uses[l] = (dir >> l) & 1;
        const int ri = m.uses[l] ? (int)(int8_t)((w.y >> (8 * l)) & 0xFF) : 0;
        const int d = pick4(ri, P.ref_delta[l][0], P.ref_delta[l][1], P.ref_delta[l][2], P.ref_delta[l][3]);
        const int lt = pick4(ri, P.ref_lt[l][0], P.ref_lt[l][1], P.ref_lt[l][2], P.ref_lt[l][3]);
        m.delta[l] = m.uses[l] ? d : 0;
        m.lt[l] = m.uses[l] ? lt : 0;
        const uint32_t mv = l ? w.w : w.z;
        m.mv[l][0] = (int)(int16_t)(mv & 0xFFFF);
        m.mv[l][1] = (int)(int16_t)(mv >> 16);
    }
    return m;
}

// One edge of boundary_strengths (src/hevc/deblock.rs:62-180): block q
// with edge flags (tu_edge, pu_edge) against neighbour p. Returns -1 when
// the motion comparison decides (both PUs inter, no intra or coded-TU rule
// applies, different PUs); a PU against itself has bS 0 (motion_bs of
// equal motion).
__device__ __forceinline__ int edge_rule(bool on, bool tu_edge, bool pu_edge, int ci_p, int ci_q, int ti_p, int ti_q,
                                         uint32_t pm_p, uint32_t pm_q) {
    if (!on || !(tu_edge || pu_edge)) return 0;
    if ((ci_q | ci_p) & 1) return 2;
    if (tu_edge && ((ti_p | ti_q) & 1)) return 1;
    if ((pm_p & 0x3FFFFFFFu) == (pm_q & 0x3FFFFFFFu)) return 0;
    return -1;
}

// bS of one 8x8 area (blocks (bx, by), (bx+1, by), (bx, by+1), (bx+1,
// by+1) of 4x4, bx and by even) per thread: the left edges of the two left
// blocks and the top edges of the two top blocks lie on the 8x8 grid; the
// other edges carry no bS. Map bytes of horizontally adjacent block pairs
// are read together (w4 is even, so every pair starts at an even index);
// the PU records of the edges that need a motion comparison are all
// requested before any is used. Output byte per 4x4 block: bits 0..1 bS of
// its left edge, bits 2..3 of its top edge.
__device__ __forceinline__ void area_bs(PicRef P, int a) {
    const int w4 = P.w4, aw = w4 >> 1;
    const int bx = (a % aw) * 2, by = (a / aw) * 2;
    const size_t q = (size_t)by * w4 + bx;
    uint32_t top_pair = 0, bottom_pair = 0;
    if (P.deblock) {
        const uint16_t* cu = (const uint16_t*)P.cuinfo;
        const uint16_t* tu = (const uint16_t*)P.tuinfo;
        const uint2* pmp = (const uint2*)P.pumap;
        const bool has_l = bx > 0, has_t = by > 0;
        const size_t h0 = q >> 1, h1 = (q + w4) >> 1;  // pair indices of rows by, by+1
        const uint32_t c0 = cu[h0], c1 = cu[h1], t0 = tu[h0], t1 = tu[h1];
        const uint2 m0 = pmp[h0], m1 = pmp[h1];
        uint32_t cl0 = 0, cl1 = 0, tl0 = 0, tl1 = 0, ct = 0, tt = 0;
        uint2 ml0 = make_uint2(0, 0), ml1 = ml0, mt = ml0;
        if (has_l) {
            cl0 = cu[h0 - 1];
            cl1 = cu[h1 - 1];
            tl0 = tu[h0 - 1];
            tl1 = tu[h1 - 1];
            ml0 = pmp[h0 - 1];
            ml1 = pmp[h1 - 1];
        }
        if (has_t) {
            const size_t ht = (q - w4) >> 1;
            ct = cu[ht];
            tt = tu[ht];
            mt = pmp[ht];
        }
        // Edges: 0 = left of (bx, by), 1 = left of (bx, by+1), 2 = top of
        // (bx, by), 3 = top of (bx+1, by).
        int ci_q[4] = {(int)(c0 & 0xFF), (int)(c1 & 0xFF), (int)(c0 & 0xFF), (int)(c0 >> 8)};
        int ti_q[4] = {(int)(t0 & 0xFF), (int)(t1 & 0xFF), (int)(t0 & 0xFF), (int)(t0 >> 8)};
        uint32_t pm_q[4] = {m0.x, m1.x, m0.x, m0.y};
        int ci_p[4] = {(int)(cl0 >> 8), (int)(cl1 >> 8), (int)(ct & 0xFF), (int)(ct >> 8)};
        int ti_p[4] = {(int)(tl0 >> 8), (int)(tl1 >> 8), (int)(tt & 0xFF), (int)(tt >> 8)};
        uint32_t pm_p[4] = {ml0.y, ml1.y, mt.x, mt.y};
        int bs[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const bool vert = e < 2;
            const int sh = vert ? 1 : 2;  // CU / TU edge-flag bit
            const bool tu_edge = ((ci_q[e] >> sh) | (ti_q[e] >> sh)) & 1;
            const bool pu_edge = (pm_q[e] >> (vert ? 30 : 31)) & 1;
            bs[e] = edge_rule(vert ? has_l : has_t, tu_edge, pu_edge, ci_p[e], ci_q[e], ti_p[e], ti_q[e], pm_p[e],
                              pm_q[e]);
        }
        if (P.ctbw) {
            // The four q blocks lie in one CTB: its slice's deblocking
            // switch, and the CTB boundaries it may not filter across
            // (src/hevc/deblock.rs:105-165).
            const uint32_t cw = P.ctbw[(size_t)((by * 4) >> P.log2_ctb) * P.wc + ((bx * 4) >> P.log2_ctb)];
            const int m = (1 << (P.log2_ctb - 2)) - 1;
            const bool v_off = (cw & kCtbDbkOff) || ((bx & m) == 0 && (cw & kCtbLeftOff));
            const bool h_off = (cw & kCtbDbkOff) || ((by & m) == 0 && (cw & kCtbTopOff));
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (e < 2 ? v_off : h_off) bs[e] = 0;
        }
        uint4 wq[4], wp[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            wq[e] = wp[e] = make_uint4(0, 0, 0, 0);
            if (bs[e] < 0) {
                wq[e] = pu_motion_words(P, pm_q[e]);
                wp[e] = pu_motion_words(P, pm_p[e]);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (bs[e] < 0) bs[e] = motion_bs(motion_from(P, wp[e]), motion_from(P, wq[e]));
        top_pair = (uint32_t)(bs[0] | (bs[2] << 2)) | ((uint32_t)(bs[3] << 2) << 8);
        bottom_pair = (uint32_t)bs[1];
    }
    uint16_t* out = (uint16_t*)P.bs;
    out[q >> 1] = (uint16_t)top_pair;
    out[(q + w4) >> 1] = (uint16_t)bottom_pair;
}

template <int D, int W>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W))) hevc_stage1(BatchDev b) {
    // The largest class layout: 16 TUs of 16x16, each padded by 16 entries.
    __shared__ __align__(16) int16_t lds[tus_per_item(2) * (16 * 16 + 16)];
    ConstI32* pre = prefix(b, 1);
    const int it = xcd_item();
    if (it >= pre[b.n]) return;
    // Uniform by construction; readfirstlane keeps the compiler's
    // uniformity analysis (and scalar loads of the descriptor) on it.
    const int p = __builtin_amdgcn_readfirstlane(find_pic(pre, b.n, it));
    PicRef P = picture(b, p);
    int i = __builtin_amdgcn_readfirstlane(it - pre[p]);
    const int tid = threadIdx.x;
    if (i < P.n_txw) {
        if (eight_bit<D>(P.bd)) tx_item<uint8_t>(P, i, tid, lds);
        else tx_item<uint16_t>(P, i, tid, lds);
        return;
    }
    i -= P.n_txw;
    const int a = i * 64 + tid;
    if (a < (P.w4 >> 1) * (P.h4 >> 1)) area_bs(P, a);
}

// ---------------------------------------------------------------------
// Stage 1, after the transform kernel: intra prediction of RV_INTRA_DEVICE
// CUs (src/hevc/intra.rs, availability as src/hevc/ctu.rs:1727-1833). A
// preparation kernel computes the residual of every TU at once (no
// residual depends on predicted samples) and files each TU under its
// dependency level; then one launch per level (the host derives the levels
// from decoding order) predicts the level's TUs and adds their residuals.
// ---------------------------------------------------------------------
// One wave's reference samples in LDS, in substitution (scan) order from
// s[7] on — scan index j: 0 .. 2n-1 = p[-1][2n-1] .. p[-1][0], 2n =
// p[-1][-1], 2n+1 .. 4n = p[0][-1] .. p[2n-1][-1] — their smoothed copy
// (likewise) and the angular ref[]
// (biased by n). Every array 16-byte aligned, so that the wide LDS reads
// the compiler forms from consecutive elements are naturally aligned.
struct alignas(16) IntraLds {
    alignas(16) uint16_t s[136];
    alignas(16) uint16_t f[136];
    alignas(16) uint16_t ref[104];
};
struct alignas(16) ResidualLds {
    alignas(16) int16_t res[32 * 32];
};

// intraPredAngle, modes 2..34, and invAngle, modes 11..25 (Tables 8-4,
// 8-5 of H.265).
__constant__ int8_t kIntraAngle[33] = {32,  26,  21,  17,  13,  9,  5,  2,  0,  -2, -5, -9, -13, -17, -21, -26, -32,
                                       -26, -21, -17, -13, -9,  -5, -2, 0,  2,  5,  9,  13,  17,  21,  26,  32};
__constant__ int16_t kInvAngle[15] = {-4096, -1638, -910, -630, -482, -390, -315, -256,
                                      -315,  -390,  -482, -630, -910, -1638, -4096};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// MinTbAddrZs (6.5.2) of 4x4 luma block (bx, by): CTB raster address, then
// the block's z order inside the CTB (sh = log2 CTB size - 2).
__device__ __forceinline__ uint32_t zs4(int bx, int by, int sh, int wc) {
    const uint32_t m = (1u << sh) - 1;
    uint32_t ix = (uint32_t)bx & m, iy = (uint32_t)by & m;
    ix = (ix | (ix << 2)) & 0x33u, ix = (ix | (ix << 1)) & 0x55u;
    iy = (iy | (iy << 2)) & 0x33u, iy = (iy | (iy << 1)) & 0x55u;
    return ((uint32_t)((by >> sh) * wc + (bx >> sh)) << (2 * sh)) | ix | (iy << 1);
}

// The nearest available scan position below j (-1 if none), from the
// wave's three availability masks.
__device__ __forceinline__ int prev_avail(int j, uint64_t m0, uint64_t m1, uint64_t m2) {
    const int seg = j >> 6;
    const uint64_t below = (1ull << (j & 63)) - 1;
    const uint64_t ms[3] = {m0, m1, m2};
    for (int s2 = seg; s2 >= 0; --s2) {
        const uint64_t mm = (s2 == 0 ? m0 : s2 == 1 ? m1 : m2) & (s2 == seg ? below : ~0ull);
        if (mm) return s2 * 64 + 63 - __builtin_clzll(mm);
    }
    (void)ms;
    return -1;
}

template <typename S> __device__ __forceinline__ void store_row4(S* d, const int* v, int count) {
    // count samples (a multiple of 4) as dwords; rows start at multiples of
    // 4 samples, so every dword is aligned.
    if constexpr (sizeof(S) == 1) {
        for (int k = 0; k < count; k += 4)
            *(uint32_t*)(d + k) = (uint32_t)v[k] | (uint32_t)v[k + 1] << 8 | (uint32_t)v[k + 2] << 16 |
                                  (uint32_t)v[k + 3] << 24;
    } else {
        for (int k = 0; k < count; k += 2) *(uint32_t*)(d + k) = (uint32_t)v[k] | (uint32_t)v[k + 1] << 16;
    }
}

// The residual of one coded transform block, by the stage-1 arithmetic
// (levels scaled, columns rounded by 7 and clipped, rows by 20 - BitDepth
// and clipped to 16 bits), stored row-major at ires + res.
template <int LOG2>
__device__ __forceinline__ void intra_residual(PicRef P, const IntraTu it, int lane, ResidualLds& L) {
    constexpr int N = 1 << LOG2;
    const rv_hevc_tu tu = it.t;
    for (int o = lane; o < N * N; o += 64) L.res[o] = 0;
    wave_sync();
    {
        const uint32_t* z = (const uint32_t*)(P.nz + tu.coef_off);
        const bool levels = tu.flags & RV_TU_LEVELS;
        LevelScale scl;
        if (levels) scl = level_scale(P, tu, LOG2);
        for (int e = lane; e < tu.n_nz; e += 64) {
            const uint32_t w = z[e];
            const int pos = (int)(w & 0xFFFF) & (N * N - 1), cv = (int)(int16_t)(w >> 16);
            L.res[pos] = levels ? scale_level<LOG2>(scl, cv, pos) : (int16_t)cv;
        }
    }
    wave_sync();
    const bool dstx = LOG2 == 2 && (tu.flags & RV_TU_DST);
    if (lane < N) {
        int col[N], o[N];
#pragma unroll
        for (int j = 0; j < N; ++j) col[j] = L.res[j * N + lane];
        if constexpr (N == 4) {
            if (dstx) idst1d(col, o);
            else idct1d<4>(col, o);
        } else {
            idct1d<N>(col, o);
        }
#pragma unroll
        for (int j = 0; j < N; ++j) L.res[j * N + lane] = (int16_t)clampi((o[j] + 64) >> 7, -32768, 32767);
    }
    wave_sync();
    if (lane < N) {
        const int bd_shift = 20 - P.bd, round2 = 1 << (bd_shift - 1);
        int row[N], o[N];
#pragma unroll
        for (int x = 0; x < N; ++x) row[x] = L.res[lane * N + x];
        if constexpr (N == 4) {
            if (dstx) idst1d(row, o);
            else idct1d<4>(row, o);
        } else {
            idct1d<N>(row, o);
        }
        // Row starts are multiples of 4 samples: whole dwords.
        uint32_t* d = (uint32_t*)(P.ires + it.res + lane * N);
#pragma unroll
        for (int x = 0; x < N; x += 2) {
            const int u = clampi((o[x] + round2) >> bd_shift, -32768, 32767);
            const int v = clampi((o[x + 1] + round2) >> bd_shift, -32768, 32767);
            d[x >> 1] = (uint32_t)(u & 0xFFFF) | ((uint32_t)v << 16);
        }
    }
}

// One transform block: reference samples, substitution, smoothing, then
// per lane one run of a row predicted (planar, DC, angular, boundary
// filters), its residual added and the run stored.
// Inlined into the kernel: its LDS stays in the LDS address space (DS
// instructions, LDS alignment rules). Reached through a generic pointer in
// a called function, the 16-bit LDS arrays were read with merged 4-byte
// FLAT loads at 2-byte-aligned addresses, which the hardware refuses.
template <typename S, int LOG2>
__device__ __forceinline__ void intra_tu(PicRef P, const IntraTu it, int lane, IntraLds& L) {
    constexpr int N = 1 << LOG2, M = 4 * N + 1;
    // A lane's run: RUN samples of one row (a multiple of 4: whole dwords);
    // LANES lanes cover the block.
    constexpr int RUN = N * N / 64 < 4 ? 4 : N * N / 64, PER_ROW = N / RUN, LANES = N * N / RUN;
    const rv_hevc_tu tu = it.t;
    const int mode = it.mode, c = tu.comp, sc = c ? 1 : 0, bd = P.bd, max = (1 << bd) - 1;
    const int x0 = tu.x, y0 = tu.y, pitch = P.pitch[c];
    S* plane = (S*)P.cur[c];
    // The scan arrays start 7 samples in: the top row, p[x][-1] at scan
    // index 2n + 1 + x, then lies on 16-byte boundaries for x a multiple
    // of 8, where the compiler reads runs of it as single wide loads.
    uint16_t* const Ls = L.s + 7;
    uint16_t* const Lf = L.f + 7;
    const int ry = lane / PER_ROW, rx = (lane % PER_ROW) * RUN;
    const bool active = lane < LANES;
    // The run's residual, loaded first so that the load overlaps the rest.
    // Residuals start at multiples of 16 samples (32 bytes), rows at
    // multiples of 4: 8-byte loads, or 16-byte ones for 32x32 runs.
    int res[RUN];
#pragma unroll
    for (int k = 0; k < RUN; ++k) res[k] = 0;
    if (active && it.res >= 0) {
        const int16_t* r = P.ires + it.res + ry * N + rx;
        if constexpr (RUN == 4) {
            const uint2 v = *(const uint2*)r;
            const uint32_t w[2] = {v.x, v.y};
#pragma unroll
            for (int k = 0; k < 4; ++k) res[k] = (int)(int16_t)(w[k >> 1] >> (16 * (k & 1)));
        } else {
#pragma unroll
            for (int q = 0; q < RUN / 8; ++q) {
                const uint4 v = ((const uint4*)r)[q];
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int k = 0; k < 8; ++k) res[8 * q + k] = (int)(int16_t)(w[k >> 1] >> (16 * (k & 1)));
            }
        }
    }
    // 1. Availability per sample (per 4x4 luma block: in the picture,
    // earlier in z-scan order, intra under constrained intra prediction)
    // and the available samples.
    const int sh = P.log2_ctb - 2;
    const uint32_t zc = zs4((x0 << sc) >> 2, (y0 << sc) >> 2, sh, P.wc);
    uint64_t mk[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int j = lane + 64 * k;
        const int xs = j < 2 * N + 1 ? x0 - 1 : x0 + (j - 2 * N - 1);
        const int ys = j < 2 * N ? y0 + 2 * N - 1 - j : y0 - 1;
        bool ok = j < M && xs >= 0 && ys >= 0;
        if (ok) {
            const int xl = xs << sc, yl = ys << sc;
            ok = xl < P.w && yl < P.h;
            if (ok) {
                const int bx = xl >> 2, by = yl >> 2;
                ok = !(zs4(bx, by, sh, P.wc) > zc) && (!P.cip || (P.cuinfo[(size_t)by * P.w4 + bx] & 1));
            }
        }
        if (ok) Ls[j] = (uint16_t)plane[(size_t)ys * pitch + xs];
        mk[k] = __ballot(ok);
    }
    wave_sync();
    // 2. Substitution (8.4.4.2.2).
    if (!(mk[0] | mk[1] | mk[2])) {
        for (int j = lane; j < M; j += 64) Ls[j] = (uint16_t)(1 << (bd - 1));
    } else {
        const int first = mk[0] ? __builtin_ctzll(mk[0]) : mk[1] ? 64 + __builtin_ctzll(mk[1]) : 128 + __builtin_ctzll(mk[2]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int j = lane + 64 * k;
            if (j < M && !((mk[k] >> lane) & 1)) {
                const int src = prev_avail(j, mk[0], mk[1], mk[2]);
                Ls[j] = Ls[src >= 0 ? src : first];
            }
        }
    }
    wave_sync();
    // 3. Smoothing (8.4.4.2.3): luma only in 4:2:0; not DC, not 4x4.
    bool smoothed = false;
    if (c == 0 && mode != 1 && N != 4) {
        const int dist = min(abs(mode - 26), abs(mode - 10));
        smoothed = dist > (N == 8 ? 7 : N == 16 ? 1 : 0);
    }
    const uint16_t* R = Ls;
    if (smoothed) {
        const int cr = Ls[2 * N];
        const bool bi = N == 32 && P.strong && abs(cr + (int)Ls[4 * N] - 2 * (int)Ls[3 * N]) < (1 << (bd - 5)) &&
                        abs(cr + (int)Ls[0] - 2 * (int)Ls[N]) < (1 << (bd - 5));
        for (int j = lane; j < M; j += 64) {
            int v = Ls[j];
            if (bi) {
                if (j >= 1 && j < 2 * N) {  // p[-1][i], i = 2n - 1 - j <= 62
                    const int i = 2 * N - 1 - j;
                    v = ((63 - i) * cr + (i + 1) * (int)Ls[0] + 32) >> 6;
                } else if (j > 2 * N && j < 4 * N) {
                    const int i = j - 2 * N - 1;
                    v = ((63 - i) * cr + (i + 1) * (int)Ls[4 * N] + 32) >> 6;
                }
            } else if (j > 0 && j < M - 1) {
                v = ((int)Ls[j - 1] + 2 * v + (int)Ls[j + 1] + 2) >> 2;
            }
            Lf[j] = (uint16_t)v;
        }
        wave_sync();
        R = Lf;
    }
    // 4. Prediction of the lane's run; DC's sum and the angular ref[]
    // by the whole wave first.
    const bool edge = c == 0 && N < 32;
    int dc = 0;
    if (mode == 1) {
        int v = lane < N ? (int)R[2 * N + 1 + lane] : lane < 2 * N ? (int)R[2 * N - 1 - (lane - N)] : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) v += __shfl_xor(v, d);
        dc = (v + N) >> (LOG2 + 1);
    }
    const int angle = mode >= 2 ? kIntraAngle[mode - 2] : 0;
    const bool vert = mode >= 18;
    if (mode >= 2) {
        // ref[k] at L.ref[N + k], k = -N .. 2N (8.4.4.2.6).
        const int last = (N * angle) >> 5;
        const int inv = angle < 0 ? (int)kInvAngle[mode - 11] : 0;
        for (int q = lane; q <= 3 * N; q += 64) {
            const int k = q - N;
            int v = 0;
            bool set = true;
            if (k == 0) {
                v = R[2 * N];
            } else if (k > 0 && (k <= N || angle >= 0)) {
                v = vert ? R[2 * N + k] : R[2 * N - k];  // main[k - 1]
            } else if (k < 0 && angle < 0 && last < -1 && k >= last) {
                const int idx = -1 + ((k * inv + 128) >> 8);
                v = idx < 0 ? R[2 * N] : vert ? R[2 * N - 1 - idx] : R[2 * N + 1 + idx];
            } else {
                set = false;
            }
            if (set) L.ref[q] = (uint16_t)v;
        }
        wave_sync();
    }
    if (!active) return;
    int v[RUN];
    const int y = ry;
    if (mode == 0) {
        const int tn = R[3 * N + 1], ln = R[N - 1], left = R[2 * N - 1 - y];
#pragma unroll
        for (int k = 0; k < RUN; ++k) {
            const int x = rx + k;
            v[k] = ((N - 1 - x) * left + (x + 1) * tn + (N - 1 - y) * (int)R[2 * N + 1 + x] + (y + 1) * ln + N) >>
                   (LOG2 + 1);
        }
    } else if (mode == 1) {
#pragma unroll
        for (int k = 0; k < RUN; ++k) {
            const int x = rx + k;
            int p = dc;
            if (edge) {
                if (x == 0 && y == 0) p = ((int)R[2 * N - 1] + 2 * dc + (int)R[2 * N + 1] + 2) >> 2;
                else if (y == 0) p = ((int)R[2 * N + 1 + x] + 3 * dc + 2) >> 2;
                else if (x == 0) p = ((int)R[2 * N - 1 - y] + 3 * dc + 2) >> 2;
            }
            v[k] = p;
        }
    } else {
        const int cr = R[2 * N];
#pragma unroll
        for (int k = 0; k < RUN; ++k) {
            const int px = rx + k, py = y;
            const int kx = vert ? px : py, ky = vert ? py : px;
            const int i_idx = ((ky + 1) * angle) >> 5, i_fact = ((ky + 1) * angle) & 31;
            const int a = L.ref[N + kx + i_idx + 1];
            int w = i_fact ? ((32 - i_fact) * a + i_fact * (int)L.ref[N + kx + i_idx + 2] + 16) >> 5 : a;
            if (edge) {
                if (mode == 26 && px == 0) w = clampi((int)R[2 * N + 1] + (((int)R[2 * N - 1 - py] - cr) >> 1), 0, max);
                else if (mode == 10 && py == 0) w = clampi((int)R[2 * N - 1] + (((int)R[2 * N + 1 + px] - cr) >> 1), 0, max);
            }
            v[k] = w;
        }
    }
#pragma unroll
    for (int k = 0; k < RUN; ++k) v[k] = clampi(v[k] + res[k], 0, max);
    store_row4<S>(plane + (size_t)(y0 + ry) * pitch + x0 + rx, v, RUN);
}

template <typename S> __device__ __forceinline__ void intra_tu_any(PicRef P, const IntraTu it, int lane, IntraLds& L) {
    switch (it.t.log2) {
        case 2: intra_tu<S, 2>(P, it, lane, L); break;
        case 3: intra_tu<S, 3>(P, it, lane, L); break;
        case 4: intra_tu<S, 4>(P, it, lane, L); break;
        default: intra_tu<S, 5>(P, it, lane, L); break;
    }
}

__device__ __forceinline__ void intra_residual_any(PicRef P, const IntraTu it, int lane, ResidualLds& L) {
    switch (it.t.log2) {
        case 2: intra_residual<2>(P, it, lane, L); break;
        case 3: intra_residual<3>(P, it, lane, L); break;
        case 4: intra_residual<4>(P, it, lane, L); break;
        default: intra_residual<5>(P, it, lane, L); break;
    }
}

// Preparation of the device-intra levels, one TU of the batch per 64-thread
// workgroup: the TU's task in its level's run of the task table and, if
// coded, its residual.
__global__ void __launch_bounds__(64) hevc_intra_prep(BatchDev b) {
    __shared__ ResidualLds lds;
    ConstI32* pre = (ConstI32*)b.ipre;
    const int it = blockIdx.x;
    if (it >= pre[b.n]) return;
    const int p = __builtin_amdgcn_readfirstlane(find_pic(pre, b.n, it));
    PicRef P = picture(b, p);
    const int k = __builtin_amdgcn_readfirstlane(it - pre[p]);
    const IntraTu t = P.itus[k];
    if (threadIdx.x == 0) {
        const int l = t.level;
        ConstI32* lp = (ConstI32*)(b.ipre + (size_t)(l + 1) * (b.n + 1));
        b.itask[((ConstI32*)b.ibase)[l] + lp[p] + (k - P.ilevel[l])] = IntraTask{b.hevc[p], P.itus + k};
    }
    if (t.res >= 0) intra_residual_any(P, t, (int)threadIdx.x, lds);
}

// The task and its TU read as dword vectors through the constant address
// space (scalar loads: both are uniform and written before the launch).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef const u32x4 __attribute__((address_space(4))) ConstU32x4;

// One dependency level: one TU per 64-thread workgroup, from the level's
// run of the task table. The kernel boundary between levels orders each
// TU after the TUs whose samples it reads.
__global__ void __launch_bounds__(64) hevc_intra_level(const IntraTask* tasks) {
    __shared__ IntraLds lds;
    const u32x4 w = ((ConstU32x4*)tasks)[blockIdx.x];
    IntraTask tk;
    __builtin_memcpy(&tk, &w, sizeof(tk));
    PicRef P = *(ConstPic*)tk.pic;
    const u32x4 u0 = ((ConstU32x4*)tk.tu)[0], u1 = ((ConstU32x4*)tk.tu)[1];
    IntraTu t;
    __builtin_memcpy(&t, &u0, 16);
    __builtin_memcpy((char*)&t + 16, &u1, 16);
    if (P.bd == 8) intra_tu_any<uint8_t>(P, t, (int)threadIdx.x, lds);
    else intra_tu_any<uint16_t>(P, t, (int)threadIdx.x, lds);
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
__device__ __forceinline__ void luma_thresholds(PicRef P, int b, int qp_p, int qp_q, int bo, int to, int* beta,
                                                int* tc) {
    int qp = (qp_p + qp_q + 1) >> 1, sh = P.bd - 8;
    *beta = (int)kBeta[clampi(qp + bo, 0, 51)] << sh;
    *tc = (int)kTc[clampi(qp + 2 * (b - 1) + to, 0, 53)] << sh;
}

// The slice offsets of the q block at luma (x, y) and the exemptions of
// blocks p and q (src/hevc/deblock.rs:186-201, 264-279): the picture's
// offsets and none without per-CTB words.
struct SegRules {
    int bo, to;
    bool np, nq;
};
__device__ __forceinline__ SegRules seg_rules(PicRef P, int x, int y, size_t p, size_t q) {
    SegRules r{P.beta_offset, P.tc_offset, false, false};
    if (P.ctbw) {
        const uint32_t cw = P.ctbw[(size_t)(y >> P.log2_ctb) * P.wc + (x >> P.log2_ctb)];
        r.bo = (int8_t)(cw >> 16);
        r.to = (int8_t)(cw >> 24);
        r.np = (P.cuinfo[p] & kCuDbkExempt) != 0;
        r.nq = (P.cuinfo[q] & kCuDbkExempt) != 0;
    }
    return r;
}

// chroma_qp (src/hevc/ctu.rs:3072-3084), 4:2:0.
__device__ __forceinline__ int chroma_qp420(int qpi) {
    if (qpi < 30) return qpi;
    if (qpi >= 43) return qpi - 6;
    const int t[14] = {29, 30, 31, 32, 33, 33, 34, 34, 35, 35, 36, 36, 37, 37};
    return t[qpi - 30];
}

// Chroma tc for Cb (c = 0) / Cr (c = 1) (src/hevc/deblock.rs:281-294).
__device__ __forceinline__ int chroma_tc(PicRef P, int qp_p, int qp_q, int c, int to) {
    int qpi = ((qp_p + qp_q + 1) >> 1) + (c == 0 ? P.cb_qp_offset : P.cr_qp_offset);
    int qpc = qpi < 0 ? qpi : chroma_qp420(qpi);
    return (int)kTc[clampi(qpc + 2 + to, 0, 53)] << (P.bd - 8);
}

// chroma_edge_scalar (src/dsp/hevc.rs:1326-1353), two lines. d points at
// q0 of line 0; `step` crosses the edge, `along` moves between lines;
// no_p / no_q leave that side unwritten (filter exemptions).
template <typename S>
__device__ void chroma_lines(S* d, long step, long along, int tc, int max, bool no_p = false, bool no_q = false) {
    if (tc == 0) return;
#pragma unroll
    for (int line = 0; line < 2; ++line) {
        S* b = d + line * along;
        int p0 = b[-step], p1 = b[-2 * step], q0 = b[0], q1 = b[step];
        int delta = clampi((((q0 - p0) << 2) + p1 - q1 + 4) >> 3, -tc, tc);
        if (!no_p) b[-step] = (S)clampi(p0 + delta, 0, max);
        if (!no_q) b[0] = (S)clampi(q0 - delta, 0, max);
    }
}

template <typename S>
__device__ void deblock_v_seg(PicRef P, int t) {
    const int nx = P.w >> 3;
    const int kx = t % nx, j = t / nx;
    if (kx == 0 || j >= P.h4) return;
    const int q = j * P.w4 + 2 * kx;
    const int b = P.bs[q] & 3;
    if (!b) return;
    const int x = kx * 8, y = j * 4, max = (1 << P.bd) - 1;
    const int qpp = P.qp[q - 1], qpq = P.qp[q];
    const SegRules sr = seg_rules(P, x, y, q - 1, q);
    int beta, tc;
    luma_thresholds(P, b, qpp, qpq, sr.bo, sr.to, &beta, &tc);
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
                for (int k = 1; k < 7; ++k)
                    if (k < 4 ? !sr.np : !sr.nq) r[k] = (S)s[line][k];
            }
        }
    }
    // Chroma: bS 2 on the 16-sample luma grid, two chroma lines.
    if (b == 2 && (x & 15) == 0) {
        for (int c = 1; c < 3; ++c) {
            int tcc = chroma_tc(P, qpp, qpq, c - 1, sr.to);
            S* pl = (S*)P.cur[c];
            chroma_lines(pl + (size_t)(y >> 1) * P.pitch[c] + (x >> 1), 1, P.pitch[c], tcc, max, sr.np, sr.nq);
        }
    }
}

template <typename S>
__device__ void deblock_h_seg(PicRef P, int t) {
    const int w4 = P.w4;
    const int i = t % w4, e = 1 + t / w4;
    const int y = e * 8;
    if (y >= P.h) return;
    const int q = (2 * e) * w4 + i;
    const int b = (P.bs[q] >> 2) & 3;
    if (!b) return;
    const int x = i * 4, max = (1 << P.bd) - 1;
    const int qpp = P.qp[q - w4], qpq = P.qp[q];
    const SegRules sr = seg_rules(P, x, y, q - w4, q);
    int beta, tc;
    luma_thresholds(P, b, qpp, qpq, sr.bo, sr.to, &beta, &tc);
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
                if (k < 4 ? sr.np : sr.nq) continue;
                S* r = Y + (size_t)(y - 4 + k) * pitch + x;
#pragma unroll
                for (int line = 0; line < 4; ++line) r[line] = (S)s[line][k];
            }
        }
    }
    if (b == 2 && (y & 15) == 0) {
        for (int c = 1; c < 3; ++c) {
            int tcc = chroma_tc(P, qpp, qpq, c - 1, sr.to);
            S* pl = (S*)P.cur[c];
            chroma_lines(pl + (size_t)(y >> 1) * P.pitch[c] + (x >> 1), P.pitch[c], 1, tcc, max, sr.np, sr.nq);
        }
    }
}

__global__ void __launch_bounds__(256) hevc_stage2(BatchDev b) {
    ConstI32* pre = prefix(b, 2);
    const int it = xcd_item();
    if (it >= pre[b.n]) return;
    // Uniform by construction; readfirstlane keeps the compiler's
    // uniformity analysis (and scalar loads of the descriptor) on it.
    const int p = __builtin_amdgcn_readfirstlane(find_pic(pre, b.n, it));
    PicRef P = picture(b, p);
    if (!P.deblock) return;
    int t = (it - pre[p]) * 256 + threadIdx.x;
    if (t >= (P.w >> 3) * P.h4) return;
    if (P.bd == 8) deblock_v_seg<uint8_t>(P, t);
    else deblock_v_seg<uint16_t>(P, t);
}

__global__ void __launch_bounds__(256) hevc_stage3(BatchDev b) {
    ConstI32* pre = prefix(b, 3);
    const int it = xcd_item();
    if (it >= pre[b.n]) return;
    // Uniform by construction; readfirstlane keeps the compiler's
    // uniformity analysis (and scalar loads of the descriptor) on it.
    const int p = __builtin_amdgcn_readfirstlane(find_pic(pre, b.n, it));
    PicRef P = picture(b, p);
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
// SAO of one group of G = 16 / sizeof(S) samples of one row (one 16-byte
// vector) inside one CTB with parameters `sp`: rows y-1, y, y+1 and the
// samples left and right of the group are loaded first; the class decides
// which of them are the two neighbours (static register indices).
template <typename S>
__device__ __forceinline__ void sao_group(PicRef P, int c, int x0, int y, const rv_sao& sp, bool on) {
    constexpr int G = 16 / sizeof(S);
    const int pw = P.pw[c], ph = P.ph[c], pitch = P.pitch[c];
    const S* src = (const S*)P.cur[c];
    S* dst = (S*)P.out[c];
    int v[3][G + 2];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const S* row = src + (size_t)clampi(y + r - 1, 0, ph - 1) * pitch;
        union {
            uint4 q;
            S s[G];
        } u;
        u.q = *(const uint4*)(row + x0);  // within the row's 256-byte-aligned pitch
#pragma unroll
        for (int k = 0; k < G; ++k) v[r][k + 1] = u.s[k];
        v[r][0] = row[max(x0 - 1, 0)];
        v[r][G + 1] = row[min(x0 + G, pw - 1)];
    }
    const int max = (1 << P.bd) - 1;
    int o[G];
#pragma unroll
    for (int k = 0; k < G; ++k) o[k] = v[1][k + 1];
    // The offsets as a byte table (offset + 128) read with v_perm_b32: one
    // select for what is otherwise a chain of compares. Offsets beyond a
    // signed byte (13 bits and deeper) take the chain.
    const int o0 = sp.offsets[0], o1 = sp.offsets[1], o2 = sp.offsets[2], o3 = sp.offsets[3];
    const bool tbl = (abs(o0) | abs(o1) | abs(o2) | abs(o3)) <= 127;
    if (on && sp.type_idx == 1) {
        const int shift = P.bd - 5;
        if (tbl) {
            // Bands 0..3 of the table carry their offsets, any other band 0.
            const uint32_t lo = (uint32_t)(o0 + 128) | (uint32_t)(o1 + 128) << 8 | (uint32_t)(o2 + 128) << 16 |
                                (uint32_t)(o3 + 128) << 24;
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const uint32_t band = (uint32_t)((o[k] >> shift) - sp.band_or_class) & 31;
                const int off = (int)__builtin_amdgcn_perm(128u, lo, min(band, 4u) | 0x0c0c0c00u);
                o[k] = clampi(o[k] + off - 128, 0, max);
            }
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k) {
                int band = ((o[k] >> shift) - sp.band_or_class) & 31;
                if (band < 4) o[k] = clampi(o[k] + sp.offsets[band], 0, max);
            }
        }
    } else if (on && sp.type_idx == 2) {
        const int cls = sp.band_or_class & 3;
        const int vy = cls == 0 ? 0 : -1;
        const bool rows_ok = y + vy >= 0 && y - vy < ph && y - vy >= 0 && y + vy < ph;
        // edgeIdx 0..4 (sign sum + 2) picks byte edgeIdx of {hi:lo}.
        const uint32_t lo = (uint32_t)(o0 + 128) | (uint32_t)(o1 + 128) << 8 | 128u << 16 | (uint32_t)(o2 + 128) << 24;
        const uint32_t hi = (uint32_t)(o3 + 128);
        auto apply = [&](int k, int a, int bb) {
            if (tbl) {
                const int s1 = clampi(o[k] - a, -1, 1), s2 = clampi(o[k] - bb, -1, 1);
                const uint32_t sel = (uint32_t)(s1 + s2) + 0x0c0c0c02u;
                o[k] = clampi(o[k] + (int)__builtin_amdgcn_perm(hi, lo, sel) - 128, 0, max);
            } else {
                const int e = sgn(o[k] - a) + sgn(o[k] - bb);  // edgeIdx - 2
                const int off = e == -2 ? o0 : e == -1 ? o1 : e == 0 ? 0 : e == 1 ? o2 : o3;
                o[k] = clampi(o[k] + off, 0, max);
            }
        };
        // Groups clear of the left and right picture edges skip the
        // per-sample column checks.
        const bool inner = x0 >= 1 && x0 + G < pw;
        if (rows_ok) {
            switch (cls) {
                case 0:  // (x - 1, y), (x + 1, y)
                    if (inner) {
#pragma unroll
                        for (int k = 0; k < G; ++k) apply(k, v[1][k], v[1][k + 2]);
                    } else {
#pragma unroll
                        for (int k = 0; k < G; ++k)
                            if (x0 + k - 1 >= 0 && x0 + k + 1 < pw) apply(k, v[1][k], v[1][k + 2]);
                    }
                    break;
                case 1:  // (x, y - 1), (x, y + 1)
                    if (inner) {
#pragma unroll
                        for (int k = 0; k < G; ++k) apply(k, v[0][k + 1], v[2][k + 1]);
                    } else {
#pragma unroll
                        for (int k = 0; k < G; ++k)
                            if (x0 + k < pw) apply(k, v[0][k + 1], v[2][k + 1]);
                    }
                    break;
                case 2:  // (x - 1, y - 1), (x + 1, y + 1)
                    if (inner) {
#pragma unroll
                        for (int k = 0; k < G; ++k) apply(k, v[0][k], v[2][k + 2]);
                    } else {
#pragma unroll
                        for (int k = 0; k < G; ++k)
                            if (x0 + k - 1 >= 0 && x0 + k + 1 < pw) apply(k, v[0][k], v[2][k + 2]);
                    }
                    break;
                default:  // (x + 1, y - 1), (x - 1, y + 1)
                    if (inner) {
#pragma unroll
                        for (int k = 0; k < G; ++k) apply(k, v[0][k + 2], v[2][k]);
                    } else {
#pragma unroll
                        for (int k = 0; k < G; ++k)
                            if (x0 + k - 1 >= 0 && x0 + k + 1 < pw) apply(k, v[0][k + 2], v[2][k]);
                    }
                    break;
            }
        }
    }
    S* drow = dst + (size_t)y * pitch + x0;
    if (x0 + G <= pw) {
        union {
            uint4 q;
            S s[G];
        } u;
#pragma unroll
        for (int k = 0; k < G; ++k) u.s[k] = (S)o[k];
        *(uint4*)drow = u.q;
    } else {
        for (int k = 0; k < G && x0 + k < pw; ++k) drow[k] = (S)o[k];
    }
}

// One sample, for CTBs narrower than a vector group.
template <typename S>
__device__ __forceinline__ void sao_sample(PicRef P, int c, int x, int y, const rv_sao& sp, bool on) {
    const int pw = P.pw[c], ph = P.ph[c], pitch = P.pitch[c], max = (1 << P.bd) - 1;
    const S* src = (const S*)P.cur[c];
    const int vv = src[(size_t)y * pitch + x];
    int r = vv;
    if (on && sp.type_idx == 1) {
        int band = ((vv >> (P.bd - 5)) - sp.band_or_class) & 31;
        if (band < 4) r = clampi(vv + sp.offsets[band], 0, max);
    } else if (on && sp.type_idx == 2) {
        const int cls = sp.band_or_class & 3;
        const int hx = cls == 1 ? 0 : (cls == 3 ? 1 : -1), vy = cls == 0 ? 0 : -1;
        const int xa = x + hx, ya = y + vy, xb = x - hx, yb = y - vy;
        if (xa >= 0 && xa < pw && xb >= 0 && xb < pw && ya >= 0 && ya < ph && yb >= 0 && yb < ph) {
            const int a = src[(size_t)ya * pitch + xa], bb = src[(size_t)yb * pitch + xb];
            const int e = sgn(vv - a) + sgn(vv - bb);
            const int off = e == -2 ? sp.offsets[0] : e == -1 ? sp.offsets[1] : e == 0 ? 0 : e == 1 ? sp.offsets[2] : sp.offsets[3];
            r = clampi(vv + off, 0, max);
        }
    }
    ((S*)P.out[c])[(size_t)y * pitch + x] = (S)r;
}

// One sample of a CTB with per-CTB rules (sao_ctb, src/hevc/sao.rs:
// 252-431): exempt samples keep their deblocked value, and an edge
// neighbour in another CTB counts only if that CTB's bit of nb (l, r, t,
// b, tl, tr, bl, br) is set.
template <typename S>
__device__ void sao_sample_rules(PicRef P, int c, int x, int y, const rv_sao& sp, bool on, uint32_t nb) {
    const int pw = P.pw[c], ph = P.ph[c], pitch = P.pitch[c], max = (1 << P.bd) - 1, sh = c ? 1 : 0;
    const S* src = (const S*)P.cur[c];
    const int vv = src[(size_t)y * pitch + x];
    int r = vv;
    const bool exempt = (P.cuinfo[(size_t)((y << sh) >> 2) * P.w4 + ((x << sh) >> 2)] & kCuSaoExempt) != 0;
    if (on && !exempt && sp.type_idx == 1) {
        int band = ((vv >> (P.bd - 5)) - sp.band_or_class) & 31;
        if (band < 4) r = clampi(vv + sp.offsets[band], 0, max);
    } else if (on && !exempt && sp.type_idx == 2) {
        const int cls = sp.band_or_class & 3;
        const int hx = cls == 1 ? 0 : (cls == 3 ? 1 : -1), vy = cls == 0 ? 0 : -1;
        const int cx0 = (x << sh) >> P.log2_ctb, cy0 = (y << sh) >> P.log2_ctb;
        auto usable = [&](int xn, int yn) {
            if (xn < 0 || yn < 0 || xn >= pw || yn >= ph) return false;
            const int dx = ((xn << sh) >> P.log2_ctb) - cx0, dy = ((yn << sh) >> P.log2_ctb) - cy0;
            if (dx == 0 && dy == 0) return true;
            const int k = dy == 0 ? (dx < 0 ? 0 : 1) : dx == 0 ? (dy < 0 ? 2 : 3) : 4 + (dy > 0) * 2 + (dx > 0);
            return ((nb >> k) & 1) != 0;
        };
        const int xa = x + hx, ya = y + vy, xb = x - hx, yb = y - vy;
        if (usable(xa, ya) && usable(xb, yb)) {
            const int a = src[(size_t)ya * pitch + xa], bb = src[(size_t)yb * pitch + xb];
            const int e = sgn(vv - a) + sgn(vv - bb);
            const int off = e == -2 ? sp.offsets[0] : e == -1 ? sp.offsets[1] : e == 0 ? 0 : e == 1 ? sp.offsets[2] : sp.offsets[3];
            r = clampi(vv + off, 0, max);
        }
    }
    ((S*)P.out[c])[(size_t)y * pitch + x] = (S)r;
}

// One CTB per workgroup (256 lanes), all three components: the parameters
// are uniform across each wave, so the type and class branches are scalar.
// In CTBs whose SAO may not read a neighbour inside the picture, or that
// hold an exempt block, the top and bottom rows go sample by sample (one
// lane each), as do vector groups with an exempt block; the other groups
// take the vector path and then redo their first or last column when it is
// the CTB's (the other samples with a neighbour in another CTB) by the
// rules.
template <typename S>
__device__ __forceinline__ void sao_ctb_rules(PicRef P, int rx, int ry, int tid, uint32_t nb) {
    constexpr int G = 16 / sizeof(S);
    for (int c = 0; c < 3; ++c) {
        const int sh = c ? 1 : 0;
        const int cs = (1 << P.log2_ctb) >> sh;
        const int x0 = rx * cs, y0 = ry * cs;
        const int w = min(cs, P.pw[c] - x0), h = min(cs, P.ph[c] - y0);
        const rv_sao sp = P.sao[3 * (ry * P.wc + rx) + c];
        const bool on = (c == 0 ? P.sao_luma : P.sao_chroma) && sp.type_idx != 0;
        if (cs % G != 0) {
            for (int t = tid; t < w * h; t += 256) sao_sample_rules<S>(P, c, x0 + t % w, y0 + t / w, sp, on, nb);
            continue;
        }
        // The top and bottom rows: one lane per sample.
        const int ring = h > 1 ? 2 * w : w;
        for (int t = tid; t < ring; t += 256)
            sao_sample_rules<S>(P, c, x0 + t % w, t < w ? y0 : y0 + h - 1, sp, on, nb);
        const int gpr = cs / G;
        for (int t = tid; t < cs * gpr; t += 256) {
            const int yy = t / gpr, gx = t - yy * gpr;
            if (yy == 0 || yy >= h - 1 || gx * G >= w) continue;
            const int x = x0 + gx * G, y = y0 + yy, n = min(G, w - gx * G);
            // The 4x4 luma blocks under the group.
            const uint8_t* ci = P.cuinfo + (size_t)((y << sh) >> 2) * P.w4 + ((x << sh) >> 2);
            bool ex = false;
            for (int k = 0; k < ((n << sh) + 3) >> 2; ++k) ex |= (ci[k] & kCuSaoExempt) != 0;
            if (ex) {
                for (int k = 0; k < n; ++k) sao_sample_rules<S>(P, c, x + k, y, sp, on, nb);
                continue;
            }
            sao_group<S>(P, c, x, y, sp, on);
            if (!on || sp.type_idx != 2) continue;
            if (gx == 0) sao_sample_rules<S>(P, c, x, y, sp, on, nb);
            if (gx * G + n == w) sao_sample_rules<S>(P, c, x + n - 1, y, sp, on, nb);
        }
    }
}

template <typename S>
__device__ __forceinline__ void sao_ctb(PicRef P, int rx, int ry, int tid) {
    constexpr int G = 16 / sizeof(S);
    const uint32_t cw = P.ctbw ? P.ctbw[(size_t)ry * P.wc + rx] : 0u;
    if (cw & kCtbSaoPerSample) {
        sao_ctb_rules<S>(P, rx, ry, tid, cw & 0xFF);
        return;
    }
    for (int c = 0; c < 3; ++c) {
        const int sh = c ? 1 : 0;
        const int cs = (1 << P.log2_ctb) >> sh;
        const int x0 = rx * cs, y0 = ry * cs;
        const int w = min(cs, P.pw[c] - x0), h = min(cs, P.ph[c] - y0);
        const rv_sao sp = P.sao[3 * (ry * P.wc + rx) + c];
        const bool on = (c == 0 ? P.sao_luma : P.sao_chroma) && sp.type_idx != 0;
        if (cs % G == 0) {
            const int gpr = cs / G;
            for (int t = tid; t < cs * gpr; t += 256) {
                const int yy = t / gpr, gx = t - yy * gpr;
                if (yy < h && gx * G < w) sao_group<S>(P, c, x0 + gx * G, y0 + yy, sp, on);
            }
        } else {
            for (int t = tid; t < w * h; t += 256) sao_sample<S>(P, c, x0 + t % w, y0 + t / w, sp, on);
        }
    }
}

template <int D>
__global__ void __launch_bounds__(256) hevc_stage4(BatchDev b) {
    ConstI32* pre = prefix(b, 4);
    const int it = xcd_item();
    if (it >= pre[b.n]) return;
    // Uniform by construction; readfirstlane keeps the compiler's
    // uniformity analysis (and scalar loads of the descriptor) on it.
    const int p = __builtin_amdgcn_readfirstlane(find_pic(pre, b.n, it));
    PicRef P = picture(b, p);
    const int i = it - pre[p];
    const int ry = i / P.wc, rx = i - ry * P.wc;
    if (eight_bit<D>(P.bd)) sao_ctb<uint8_t>(P, rx, ry, threadIdx.x);
    else sao_ctb<uint16_t>(P, rx, ry, threadIdx.x);
}

// ---------------------------------------------------------------------
// Per-entry kernels: one call of one HevcDsp entry (src/dsp/hevc.rs:
// 165-229) on one workgroup (transforms, fused MC, deblocking) or a small
// grid (element-wise entries), with the scalar kernels' arithmetic and the
// device functions the batched stages use.
// ---------------------------------------------------------------------

// idct_scalar<N> (src/dsp/hevc.rs:500-526): entries beyond max_x / max_y
// are zero, as idct1 never reads them; lane x transforms column x, then
// lane y row y.
template <int N>
__device__ void entry_idct(const HevcEntry& e, int16_t* t) {
    int16_t* c = (int16_t*)e.dst;
    const int tid = threadIdx.x;
    if (tid < N) {
        int s[N], o[N];
#pragma unroll
        for (int k = 0; k < N; ++k) s[k] = (tid <= e.max_x && k <= e.max_y) ? c[k * N + tid] : 0;
        idct1d<N>(s, o);
#pragma unroll
        for (int y = 0; y < N; ++y) t[y * 33 + tid] = (int16_t)clampi((o[y] + 64) >> 7, -32768, 32767);
    }
    __syncthreads();
    if (tid < N) {
        int s[N], o[N];
#pragma unroll
        for (int k = 0; k < N; ++k) s[k] = t[tid * 33 + k];
        idct1d<N>(s, o);
        const int sh = e.shift, rnd = 1 << (sh - 1);
#pragma unroll
        for (int x = 0; x < N; ++x) c[tid * N + x] = (int16_t)clampi((o[x] + rnd) >> sh, -32768, 32767);
    }
}

// idst4_scalar (src/dsp/hevc.rs:529-556); the bounds are not used.
__device__ void entry_idst(const HevcEntry& e, int16_t* t) {
    int16_t* c = (int16_t*)e.dst;
    const int tid = threadIdx.x;
    if (tid < 4) {
        int s[4], o[4];
        for (int k = 0; k < 4; ++k) s[k] = c[k * 4 + tid];
        idst1d(s, o);
        for (int i = 0; i < 4; ++i) t[i * 33 + tid] = (int16_t)clampi((o[i] + 64) >> 7, -32768, 32767);
    }
    __syncthreads();
    if (tid < 4) {
        int s[4], o[4];
        for (int k = 0; k < 4; ++k) s[k] = t[tid * 33 + k];
        idst1d(s, o);
        const int sh = e.shift, rnd = 1 << (sh - 1);
        for (int i = 0; i < 4; ++i) c[tid * 4 + i] = (int16_t)clampi((o[i] + rnd) >> sh, -32768, 32767);
    }
}

// The element-wise entries: add_residual, copy, the one-stage filters,
// the combiners and SAO (src/dsp/hevc.rs:558-566, 677-818, 997-1087,
// 1181-1224). Element i is (y, x) = (i / w, i % w) of the w x h block.
template <typename S>
__device__ void entry_elem(const HevcEntry& e) {
    const int n = e.w * e.h;
    const int16_t* a16 = (const int16_t*)e.src;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const int y = i / e.w, x = i - y * e.w;
        S* drow = (S*)e.dst + y * e.dst_stride + x;
        switch (e.op) {
            case HE_ADD_RES: *drow = (S)clampi((int)*drow + a16[i], 0, e.max); break;
            case HE_COPY:
                ((int16_t*)e.dst)[i] = (int16_t)((int)((const S*)e.src)[y * e.src_stride + x] << e.shift);
                break;
            case HE_FILT_H:
            case HE_FILT_V:
            case HE_FILT_V2: {
                const int taps = e.luma ? 8 : 4;
                const long along = e.op == HE_FILT_H ? 1 : e.src_stride;
                const long base = y * e.src_stride + x;
                int acc = 0;
                for (int k = 0; k < taps; ++k) {
                    const int f = e.luma ? kF.q[e.frac][k] : kF.e[e.frac][k];
                    const int v = e.op == HE_FILT_V2 ? (int)a16[base + k * along] : (int)((const S*)e.src)[base + k * along];
                    acc += f * v;
                }
                ((int16_t*)e.dst)[i] = (int16_t)(acc >> (e.op == HE_FILT_V2 ? 6 : e.shift));
                break;
            }
            case HE_UNI: {
                const int rnd = e.shift > 0 ? 1 << (e.shift - 1) : 0;
                *drow = (S)clampi((a16[i] + rnd) >> e.shift, 0, e.max);
                break;
            }
            case HE_BI:
                *drow = (S)clampi((a16[i] + e.other[i] + (1 << (e.shift - 1))) >> e.shift, 0, e.max);
                break;
            case HE_WUNI:
                if (e.log2_wd >= 1)
                    *drow = (S)clampi(((a16[i] * e.w0 + (1 << (e.log2_wd - 1))) >> e.log2_wd) + e.o0, 0, e.max);
                else
                    *drow = (S)clampi(a16[i] * e.w0 + e.o0, 0, e.max);
                break;
            case HE_WBI:
                *drow = (S)clampi((a16[i] * e.w0 + e.other[i] * e.w1 + ((e.o0 + e.o1 + 1) << e.log2_wd)) >>
                                      (e.log2_wd + 1),
                                  0, e.max);
                break;
            case HE_SAO_BAND: {
                const int v = ((const S*)e.src)[y * e.src_stride + x];
                *drow = (S)clampi(v + e.table[v >> e.shift], 0, e.max);
                break;
            }
            case HE_SAO_EDGE: {
                const long at = e.off + y * e.dst_stride + x;
                const S* s = (const S*)e.src;
                const int v = s[at], va = s[at + e.na], vb = s[at + e.nb];
                ((S*)e.dst)[at] = (S)clampi(v + e.table[2 + sgn(v - va) + sgn(v - vb)], 0, e.max);
                break;
            }
            default: break;
        }
    }
}

// qpel_uni / epel_uni / qpel_bi / epel_bi: interp_scalar's two-pass
// composition (src/dsp/hevc.rs:823-882) then uni_scalar / bi_scalar at
// 14 - bd / 15 - bd (:887-991). `src` is the filter window's origin.
template <typename S>
__device__ void entry_fused(const HevcEntry& e, int16_t* tmp) {
    const int taps = e.luma ? 8 : 4, reach = e.luma ? 3 : 1;
    const S* src = (const S*)e.src;
    const long ss = e.src_stride;
    const int w = e.w, h = e.h, bd = e.bd;
    const int shift1 = min(bd, 12) - 8, shift3 = 14 - bd;
    auto fxk = [&](int k) { return e.luma ? (int)kF.q[e.fx][k] : (int)kF.e[e.fx][k]; };
    auto fyk = [&](int k) { return e.luma ? (int)kF.q[e.fy][k] : (int)kF.e[e.fy][k]; };
    if (e.fx && e.fy) {
        for (int i = threadIdx.x; i < w * (h + taps - 1); i += 256) {
            const int y = i / w, x = i - y * w;
            int acc = 0;
            for (int k = 0; k < taps; ++k) acc += fxk(k) * src[y * ss + x + k];
            tmp[i] = (int16_t)(acc >> shift1);
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < w * h; i += 256) {
        const int y = i / w, x = i - y * w;
        int p, acc = 0;
        if (!e.fx && !e.fy) {
            p = (int16_t)((int)src[(y + reach) * ss + x + reach] << shift3);
        } else if (!e.fy) {
            for (int k = 0; k < taps; ++k) acc += fxk(k) * src[(y + reach) * ss + x + k];
            p = (int16_t)(acc >> shift1);
        } else if (!e.fx) {
            for (int k = 0; k < taps; ++k) acc += fyk(k) * src[(y + k) * ss + x + reach];
            p = (int16_t)(acc >> shift1);
        } else {
            for (int k = 0; k < taps; ++k) acc += fyk(k) * tmp[(y + k) * w + x];
            p = (int16_t)(acc >> 6);
        }
        S* d = (S*)e.dst + y * e.dst_stride + x;
        const int mx = (1 << bd) - 1;
        if (e.other) {
            const int sh = 15 - bd;
            *d = (S)clampi((e.other[i] + p + (1 << (sh - 1))) >> sh, 0, mx);
        } else {
            const int sh = 14 - bd, rnd = sh > 0 ? 1 << (sh - 1) : 0;
            *d = (S)clampi((p + rnd) >> sh, 0, mx);
        }
    }
}

// deblock_luma_scalar (src/dsp/hevc.rs:1355-1382): lane k < 2 filters
// segment k with luma_segment; an exempt side is not written back.
template <typename S>
__device__ void entry_dbk_luma(const HevcEntry& e) {
    const int seg = threadIdx.x;
    if (seg >= 2 || (e.beta[seg] == 0 && e.tc[seg] == 0)) return;
    const long step = e.vertical ? 1 : e.dst_stride, along = e.vertical ? e.dst_stride : 1;
    S* d = (S*)e.dst + e.off + 4 * seg * along;
    int s[4][8];
#pragma unroll
    for (int line = 0; line < 4; ++line)
#pragma unroll
        for (int k = 0; k < 8; ++k) s[line][k] = d[line * along + (k - 4) * step];
    bool changed = false;
    luma_segment(s, e.beta[seg], e.tc[seg], e.max, changed);
    if (!changed) return;
#pragma unroll
    for (int line = 0; line < 4; ++line)
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < 4 ? !e.no_p[seg] : !e.no_q[seg]) d[line * along + (k - 4) * step] = (S)s[line][k];
}

// deblock_chroma_scalar (:1384-1411): lane k < 4 filters segment k.
template <typename S>
__device__ void entry_dbk_chroma(const HevcEntry& e) {
    const int seg = threadIdx.x;
    if (seg >= 4) return;
    const long step = e.vertical ? 1 : e.dst_stride, along = e.vertical ? e.dst_stride : 1;
    chroma_lines<S>((S*)e.dst + e.off + 2 * seg * along, step, along, e.tc[seg], e.max, e.no_p[seg] != 0,
                    e.no_q[seg] != 0);
}

__global__ void __launch_bounds__(256) hevc_entry(HevcEntry e) {
    __shared__ int16_t lds[64 * 71];
    switch (e.op) {
        case HE_IDCT:
            switch (e.w) {
                case 4: entry_idct<4>(e, lds); break;
                case 8: entry_idct<8>(e, lds); break;
                case 16: entry_idct<16>(e, lds); break;
                default: entry_idct<32>(e, lds); break;
            }
            break;
        case HE_IDST: entry_idst(e, lds); break;
        case HE_FUSED:
            if (e.ss == 1) entry_fused<uint8_t>(e, lds);
            else entry_fused<uint16_t>(e, lds);
            break;
        case HE_DBK_LUMA:
            if (e.ss == 1) entry_dbk_luma<uint8_t>(e);
            else entry_dbk_luma<uint16_t>(e);
            break;
        case HE_DBK_CHROMA:
            if (e.ss == 1) entry_dbk_chroma<uint8_t>(e);
            else entry_dbk_chroma<uint16_t>(e);
            break;
        default:
            if (e.ss == 1) entry_elem<uint8_t>(e);
            else entry_elem<uint16_t>(e);
            break;
    }
}

}  // namespace

hipError_t hevc_entry_launch(const HevcEntry& e, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(hevc_entry, dim3(blocks), dim3(256), 0, st, e);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) rv_copy_probe(uint4* __restrict__ dst, const uint4* __restrict__ src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

hipError_t launch_copy_probe(void* dst, const void* src, int64_t bytes, hipStream_t st) {
    const int64_t n = bytes / 16;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(rv_copy_probe, dim3(grid), dim3(256), 0, st, (uint4*)dst, (const uint4*)src, n);
    return hipGetLastError();
}

// Work items per stage for one picture (the grid unit of each stage).
void hevc_stage_items(HevcPicDev& P) {
    P.items[0] = P.n_side + P.n_intra + P.n_tiles16 + (P.n_tiles + 3) / 4;
    P.items[1] = P.n_txw + ((P.w4 >> 1) * (P.h4 >> 1) + 63) / 64;
    P.items[2] = ((P.w >> 3) * P.h4 + 255) / 256;
    P.items[3] = (P.h >= 16) ? ((((P.h >> 3) - 1) * P.w4) + 255) / 256 : 0;
    P.items[4] = P.wc * P.hc;
}

namespace {

// Kernel variants per stage: [0] serves any picture; the others are
// depth-specialised (eight_bit<D>) with their own occupancy targets.
struct StageVariant {
    const void* fn;
    int depth;    // 0 any, 1 8-bit pictures only, 2 deeper pictures only
    int threads;  // workgroup size
    const char* name;
    const void* strips;  // stage 0: the strip kernel launched with it
};
#define RV_VARIANT(k, d, t, n) \
    StageVariant { reinterpret_cast<const void*>(&k), d, t, n, nullptr }
#define RV_VARIANT0(k, s, d, n) \
    StageVariant { reinterpret_cast<const void*>(&k), d, 64, n, reinterpret_cast<const void*>(&s) }
// Stage 0 pairs a tile kernel (side data, intra copies, the blocks of
// deeper pictures and blocks whose windows cross a plane edge) with a strip
// kernel (the other blocks of 8-bit pictures; deeper pictures have none).
const StageVariant kStage0[] = {
    RV_VARIANT0((hevc_stage0<0, 5>), (hevc_mc_strips<3>), 0, "any/5, strips/3"),
    RV_VARIANT0((hevc_stage0<1, 7>), (hevc_mc_strips<3>), 1, "8-bit/7, strips/3"),
    RV_VARIANT0((hevc_stage0<1, 8, 1>), (hevc_mc_strips<4, 1>), 1, "8-bit/8 per-list, strips/4"),
    RV_VARIANT0((hevc_stage0<1, 8, 1>), (hevc_mc_strips<3>), 1, "8-bit/8 per-list, strips/3"),
    RV_VARIANT0((hevc_stage0<2, 5>), (hevc_mc_strips<3>), 2, "deep/5"),
    RV_VARIANT0((hevc_stage0<2, 6, 1>), (hevc_mc_strips<3>), 2, "deep/6 per-list"),
    RV_VARIANT0((hevc_stage0<0, 6, 1>), (hevc_mc_strips<3>), 0, "any/6 per-list, strips/3"),
    RV_VARIANT0((hevc_stage0<1, 8, 1, true>), (hevc_mc_strips<3>), 1, "8-bit/8 per-list 8x8 windows, strips/3"),
    RV_VARIANT0((hevc_stage0<2, 6, 1, true>), (hevc_mc_strips<3>), 2, "deep/6 per-list 8x8 windows")};
const StageVariant kStage1[] = {RV_VARIANT((hevc_stage1<0, 4>), 0, 64, "any/4"),
                                RV_VARIANT((hevc_stage1<1, 4>), 1, 64, "8-bit/4"),
                                RV_VARIANT((hevc_stage1<2, 4>), 2, 64, "deep/4")};
const StageVariant kStage2[] = {RV_VARIANT(hevc_stage2, 0, 256, "any")};
const StageVariant kStage3[] = {RV_VARIANT(hevc_stage3, 0, 256, "any")};
const StageVariant kStage4[] = {RV_VARIANT(hevc_stage4<0>, 0, 256, "any"), RV_VARIANT(hevc_stage4<1>, 1, 256, "8-bit"),
                                RV_VARIANT(hevc_stage4<2>, 2, 256, "deep")};
#undef RV_VARIANT0
#undef RV_VARIANT
struct StageTable {
    const StageVariant* v;
    int n;
};
template <size_t N> constexpr StageTable table(const StageVariant (&v)[N]) { return StageTable{v, (int)N}; }
const StageTable kStageTables[kStages] = {table(kStage0), table(kStage1), table(kStage2), table(kStage3),
                                          table(kStage4)};

}  // namespace

int hevc_stage_variants(int stage) { return stage >= 0 && stage < kStages ? kStageTables[stage].n : 0; }
int hevc_variant_depth(int stage, int v) {
    return v >= 0 && v < hevc_stage_variants(stage) ? kStageTables[stage].v[v].depth : -1;
}
const char* hevc_variant_name(int stage, int v) {
    return v >= 0 && v < hevc_stage_variants(stage) ? kStageTables[stage].v[v].name : "";
}

hipError_t hevc_launch_stage(int stage, const BatchDev& b, int total, hipStream_t st, int variant) {
    if (variant < 0 || variant >= hevc_stage_variants(stage)) return hipErrorInvalidValue;
    BatchDev arg = b;
    void* args[] = {&arg};
    if (total > 0) {
        const StageVariant& k = kStageTables[stage].v[variant];
        const dim3 grid(8 * ((total + 7) / 8));  // see xcd_item()
        const hipError_t e = hipLaunchKernel(k.fn, grid, dim3(k.threads), args, 0, st);
        if (e != hipSuccess) return e;
    }
    // Stage 0 also runs the strip kernel (blocks with interior windows), of
    // the stage variant's depth class.
    if (stage == 0 && b.stotal > 0) {
        const void* fn = kStageTables[0].v[variant].strips;
        const hipError_t e = hipLaunchKernel(fn, dim3(8 * ((b.stotal + 7) / 8)), dim3(64), args, 0, st);
        if (e != hipSuccess) return e;
    }
    // Stage 1 ends with the intra prediction of device-intra CUs: the
    // preparation over all their TUs, then one launch per dependency level,
    // one TU per workgroup.
    if (stage == 1 && b.ilevels > 0 && b.iall > 0) {
        hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(&hevc_intra_prep), dim3(b.iall), dim3(64), args,
                                       0, st);
        if (e != hipSuccess) return e;
        int64_t at = 0;
        for (int l = 0; l < b.ilevels; ++l) {
            if (b.itotal[l] <= 0) continue;
            const IntraTask* tasks = b.itask + at;
            at += b.itotal[l];
            void* largs[] = {&tasks};
            e = hipLaunchKernel(reinterpret_cast<const void*>(&hevc_intra_level), dim3(b.itotal[l]), dim3(64), largs,
                                0, st);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

}  // namespace rv
