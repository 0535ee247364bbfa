// TEST INFRASTRUCTURE ONLY — the CPU oracle for the HEVC reconstruction
// path. A restatement, in C++, of the reference's scalar kernel table
// (src/dsp/hevc.rs) and of the drivers that call it for one inter picture
// (src/hevc/inter.rs, src/hevc/deblock.rs, src/hevc/sao.rs, the residual
// and side-data lines of src/hevc/ctu.rs). Only tests, smoke() and the CPU
// baseline of bench.py load it. Each function cites the lines it follows.
//
// Pinned by (tests/test_oracle_hevc.py): the reference's own closed-form
// tests (butterfly == matrix product, src/dsp/hevc.rs:1511-1536; the
// add_residual known answers :1495-1507), DC-only closed forms, filter
// invariants (constant planes, symmetric half-sample filter on ramps,
// fraction 0 == copy, two-pass == direct 2-D sum at 8 bits), combine
// identities (bi(a,a) == uni(a), explicit weights 2^d / 0 == default),
// deblocking (bS 0 / tc 0 identity, flat planes stay flat, hand-computed
// step edges), SAO (zero offsets identity, band closed form, edge classes
// on peaks/valleys), and border equivalence (padded plane + gather ==
// coordinate clamping).
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

// Table 8-12 / 8-13 of H.265 (src/hevc/tables_gen.rs:173-191).
const int QPEL[4][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0},
    {-1, 4, -10, 58, 17, -5, 1, 0},
    {-1, 4, -11, 40, 40, -11, 4, -1},
    {0, 1, -5, 17, 58, -10, 4, -1},
};
const int EPEL[8][4] = {
    {0, 0, 0, 0},    {-2, 58, 10, -2}, {-4, 54, 16, -2}, {-6, 46, 28, -4},
    {-4, 36, 36, -4}, {-4, 28, 46, -6}, {-2, 16, 54, -4}, {-2, 10, 58, -2},
};
// tC' and beta' by Q (Table 8-12 of H.265; src/hevc/tables_gen.rs:195-199).
const int TC_TABLE[54] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                          1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4,
                          4, 4, 5, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 22, 24};
const int BETA_TABLE[52] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,
                            0,  0,  0,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15,
                            16, 17, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38,
                            40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64};

// The 32-point inverse transform matrix of 8.6.4.2. The standard lists it
// as a table; every entry is +-c[m] for the 32 distinct values c[m] of its
// first column, with m = ((2n + 1) k) mod 128 folded onto 0..32 by the
// cosine symmetries. Built from that rule here (the reference transcribes
// the table, src/hevc/tables_gen.rs:136-169; tests check both agree on the
// properties the standard gives the table).
struct Transform32 {
    int m[32][32];
    Transform32() {
        static const int c[33] = {64, 90, 90, 90, 89, 88, 87, 85, 83, 82, 80,
                                  78, 75, 73, 70, 67, 64, 61, 57, 54, 50, 46,
                                  43, 38, 36, 31, 25, 22, 18, 13, 9,  4,  0};
        for (int k = 0; k < 32; ++k)
            for (int n = 0; n < 32; ++n) {
                int a = ((2 * n + 1) * k) % 128;
                int v;
                if (a <= 32) v = c[a];
                else if (a <= 64) v = -c[64 - a];
                else if (a <= 96) v = -c[a - 64];
                else v = c[128 - a];
                m[k][n] = v;
            }
    }
};
const Transform32 T32;

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline int sgn(int v) { return (v > 0) - (v < 0); }

// scale_coefficients (src/hevc/residual.rs:1223-1283), 8.6.3: within the
// nonzero bounds, d = Clip3(-32768, 32767, ((c * m * levelScale[qP % 6] <<
// (qP / 6)) + (1 << (bdShift - 1))) >> bdShift), bdShift = BitDepth +
// Log2(nTbS) - 5. m is 16 for flat scaling (list == nullptr), else the
// list's factor: list[y * 4 + x] (4x4), list[y * 8 + x] (8x8), dc at (0, 0)
// and list[(y / r) * 8 + x / r], r = n / 8, for 16x16 / 32x32. The
// reference's flat fast path (:1243-1254) computes the same values in i32.
// Transform skip (flat above 4x4) has no place in the work lists.
void scale_coefficients(int16_t* coeffs, int log2, int qp, int bd, const uint8_t* list, int dc, int max_x,
                        int max_y) {
    static const int LEVEL_SCALE[6] = {40, 45, 51, 57, 64, 72};
    const int n = 1 << log2;
    const int bd_shift = bd + log2 - 5;
    const int64_t round = (int64_t)1 << (bd_shift - 1);
    const int64_t ls = LEVEL_SCALE[qp % 6];
    const int q6 = qp / 6;
    for (int y = 0; y <= max_y; ++y)
        for (int x = 0; x <= max_x; ++x) {
            const int c = coeffs[y * n + x];
            if (c == 0) continue;
            int64_t m = 16;
            if (list) {
                if (n == 4) m = list[y * 4 + x];
                else if (n == 8) m = list[y * 8 + x];
                else if (x == 0 && y == 0) m = dc;
                else m = list[(y / (n / 8)) * 8 + x / (n / 8)];
            }
            const int64_t v = (((int64_t)c * m * ls) << q6) + round;
            const int64_t d = v >> bd_shift;
            coeffs[y * n + x] = (int16_t)(d < -32768 ? -32768 : (d > 32767 ? 32767 : d));
        }
}

// The list and DC of a TU (src/hevc/ctu.rs:2016-2029): sizeId = log2 - 2,
// matrixId = cIdx for intra CUs, 3 + cIdx otherwise; DC 16 below 16x16.
const uint8_t* tu_scaling(const rv_hevc_scaling* sl, int log2, int comp, bool intra, int* dc) {
    *dc = 16;
    if (!sl) return nullptr;
    const int mid = (intra ? 0 : 3) + comp;
    if (log2 >= 4) *dc = sl->dc[log2 - 4][mid];
    return log2 == 2 ? sl->list4[mid] : sl->list8[log2 - 3][mid];
}

template <typename S> inline int rd(const S* p, long i) { return (int)p[i]; }

// ---------------------------------------------------------------------
// Inverse transforms (src/dsp/hevc.rs:474-566)
// ---------------------------------------------------------------------

// idct1: the even/odd partial butterfly (src/dsp/hevc.rs:474-497).
void idct1(const int16_t* src, int stride, int n, int nz, int* out) {
    if (n == 1) {
        out[0] = 64 * (int)src[0];
        return;
    }
    int half = n / 2;
    int e[16];
    idct1(src, stride * 2, half, (nz + 1) / 2, e);
    int step = 32 / n;
    for (int k = 0; k < half; ++k) {
        int o = 0;
        for (int j = 1; j < nz; j += 2) o += T32.m[j * step][k] * (int)src[j * stride];
        out[k] = e[k] + o;
        out[n - 1 - k] = e[k] - o;
    }
}

// idct_scalar (src/dsp/hevc.rs:499-526).
void idct(int16_t* c, int N, int bd_shift, int max_x, int max_y) {
    int round2 = 1 << (bd_shift - 1);
    if (max_x == 0 && max_y == 0) {
        int v = clampi((c[0] * 64 + 64) >> 7, -32768, 32767);
        int16_t r = (int16_t)clampi((v * 64 + round2) >> bd_shift, -32768, 32767);
        for (int i = 0; i < N * N; ++i) c[i] = r;
        return;
    }
    int16_t tmp[32 * 32];
    std::memset(tmp, 0, sizeof(tmp));
    int out[32];
    for (int x = 0; x <= max_x; ++x) {
        idct1(c + x, N, N, max_y + 1, out);
        for (int y = 0; y < N; ++y) tmp[y * N + x] = (int16_t)clampi((out[y] + 64) >> 7, -32768, 32767);
    }
    for (int y = 0; y < N; ++y) {
        idct1(tmp + y * N, 1, N, max_x + 1, out);
        for (int x = 0; x < N; ++x)
            c[y * N + x] = (int16_t)clampi((out[x] + round2) >> bd_shift, -32768, 32767);
    }
}

// idst4_scalar (src/dsp/hevc.rs:528-556).
void idst4(int16_t* c, int bd_shift) {
    static const int M[4][4] = {{29, 55, 74, 84}, {74, 74, 0, -74}, {84, -29, -74, 55}, {55, -84, 74, -29}};
    int round2 = 1 << (bd_shift - 1);
    int16_t tmp[16];
    for (int x = 0; x < 4; ++x)
        for (int i = 0; i < 4; ++i) {
            int s = 0;
            for (int j = 0; j < 4; ++j) s += M[j][i] * c[j * 4 + x];
            tmp[i * 4 + x] = (int16_t)clampi((s + 64) >> 7, -32768, 32767);
        }
    for (int y = 0; y < 4; ++y)
        for (int i = 0; i < 4; ++i) {
            int s = 0;
            for (int j = 0; j < 4; ++j) s += M[j][i] * tmp[y * 4 + j];
            c[y * 4 + i] = (int16_t)clampi((s + round2) >> bd_shift, -32768, 32767);
        }
}

// add_residual_scalar (src/dsp/hevc.rs:558-566).
template <typename S> void add_residual(S* dst, long stride, const int16_t* res, int n, int max) {
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            S* p = dst + y * stride + x;
            *p = (S)clampi((int)*p + res[y * n + x], 0, max);
        }
}

// ---------------------------------------------------------------------
// Interpolation (src/dsp/hevc.rs:677-882)
// ---------------------------------------------------------------------

template <typename S> void copy_k(int16_t* dst, const S* src, long ss, int w, int h, int shift) {
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) dst[y * w + x] = (int16_t)(rd(src, y * ss + x) << shift);
}
template <typename S>
void filt_h(const int* f, int taps, int16_t* dst, const S* src, long ss, int w, int h, int shift) {
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int acc = 0;
            for (int k = 0; k < taps; ++k) acc += f[k] * rd(src, y * ss + x + k);
            dst[y * w + x] = (int16_t)(acc >> shift);
        }
}
template <typename S>
void filt_v(const int* f, int taps, int16_t* dst, const S* src, long ss, int w, int h, int shift) {
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int acc = 0;
            for (int k = 0; k < taps; ++k) acc += f[k] * rd(src, (y + k) * ss + x);
            dst[y * w + x] = (int16_t)(acc >> shift);
        }
}
// qpel_v2 / epel_v2: vertical taps over 14-bit rows, >> 6 (:737-755, :800-818).
void filt_v2(const int* f, int taps, int16_t* dst, const int16_t* src, long ss, int w, int h) {
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int acc = 0;
            for (int k = 0; k < taps; ++k) acc += f[k] * (int)src[(y + k) * ss + x];
            dst[y * w + x] = (int16_t)(acc >> 6);
        }
}

// interp_scalar (src/dsp/hevc.rs:823-882): src at the window origin.
template <typename S>
void interp_k(bool luma, int16_t* pred, int16_t* tmp, const S* src, long ss, int w, int h, int fx, int fy,
              int bd) {
    int reach = luma ? 3 : 1, taps = luma ? 8 : 4;
    int shift1 = std::min(bd, 12) - 8;
    int shift3 = 14 - bd;
    const int* fxt = luma ? QPEL[fx] : EPEL[fx];
    const int* fyt = luma ? QPEL[fy] : EPEL[fy];
    if (fx == 0 && fy == 0) {
        copy_k(pred, src + reach * ss + reach, ss, w, h, shift3);
    } else if (fy == 0) {
        filt_h(fxt, taps, pred, src + reach * ss, ss, w, h, shift1);
    } else if (fx == 0) {
        filt_v(fyt, taps, pred, src + reach, ss, w, h, shift1);
    } else {
        filt_h(fxt, taps, tmp, src, ss, w, h + taps - 1, shift1);
        filt_v2(fyt, taps, pred, tmp, w, w, h);
    }
}

// uni_scalar / bi_scalar / weighted_* (src/dsp/hevc.rs:997-1087).
template <typename S> void uni_k(S* dst, long stride, const int16_t* src, int w, int h, int shift, int max) {
    int round = shift > 0 ? 1 << (shift - 1) : 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) dst[y * stride + x] = (S)clampi((src[y * w + x] + round) >> shift, 0, max);
}
template <typename S>
void bi_k(S* dst, long stride, const int16_t* a, const int16_t* b, int w, int h, int shift, int max) {
    int round = 1 << (shift - 1);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            dst[y * stride + x] = (S)clampi((a[y * w + x] + b[y * w + x] + round) >> shift, 0, max);
}
template <typename S>
void wuni_k(S* dst, long stride, const int16_t* src, int w, int h, int log2_wd, int wt, int o, int max) {
    if (log2_wd >= 1) {
        int round = 1 << (log2_wd - 1);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                dst[y * stride + x] = (S)clampi(((src[y * w + x] * wt + round) >> log2_wd) + o, 0, max);
    } else {
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) dst[y * stride + x] = (S)clampi(src[y * w + x] * wt + o, 0, max);
    }
}
template <typename S>
void wbi_k(S* dst, long stride, const int16_t* a, const int16_t* b, int w, int h, int log2_wd, int w0, int w1,
           int o0, int o1, int max) {
    int round = (o0 + o1 + 1) << log2_wd;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int v = (a[y * w + x] * w0 + b[y * w + x] * w1 + round) >> (log2_wd + 1);
            dst[y * stride + x] = (S)clampi(v, 0, max);
        }
}

// ---------------------------------------------------------------------
// SAO kernels (src/dsp/hevc.rs:1181-1224)
// ---------------------------------------------------------------------
template <typename S>
void sao_band_k(S* dst, long ds, const S* src, long ss, int w, int h, const int16_t* table, int shift, int max) {
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int v = src[y * ss + x];
            dst[y * ds + x] = (S)clampi(v + table[v >> shift], 0, max);
        }
}
template <typename S>
void sao_edge_k(S* dst, const S* src, long origin, long stride, int w, int h, long na, long nb, const int16_t* off,
                int max) {
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            long i = origin + y * stride + x;
            int v = src[i], a = src[i + na], b = src[i + nb];
            int e = 2 + sgn(v - a) + sgn(v - b);
            dst[i] = (S)clampi(v + off[e], 0, max);
        }
}

// ---------------------------------------------------------------------
// Deblocking kernels (src/dsp/hevc.rs:1230-1411)
// ---------------------------------------------------------------------

// luma_edge_scalar (:1234-1324): one 4-line segment.
template <typename S>
void luma_edge(S* d, long pos, long step, long along, int beta, int tc, bool no_p, bool no_q, int max) {
    auto s = [&](int line, int k) -> int { return d[pos + line * along + k * step]; };
    int dp0 = std::abs(s(0, -3) - 2 * s(0, -2) + s(0, -1));
    int dp3 = std::abs(s(3, -3) - 2 * s(3, -2) + s(3, -1));
    int dq0 = std::abs(s(0, 2) - 2 * s(0, 1) + s(0, 0));
    int dq3 = std::abs(s(3, 2) - 2 * s(3, 1) + s(3, 0));
    int dpq0 = dp0 + dq0, dpq3 = dp3 + dq3, dp = dp0 + dp3, dq = dq0 + dq3;
    int dd = dpq0 + dpq3;
    if (dd >= beta) return;
    auto dsam = [&](int line, int dpq) -> bool {
        return dpq < (beta >> 2) &&
               std::abs(s(line, -4) - s(line, -1)) + std::abs(s(line, 0) - s(line, 3)) < (beta >> 3) &&
               std::abs(s(line, -1) - s(line, 0)) < ((5 * tc + 1) >> 1);
    };
    bool strong = dsam(0, 2 * dpq0) && dsam(3, 2 * dpq3);
    bool dep = dp < ((beta + (beta >> 1)) >> 3);
    bool deq = dq < ((beta + (beta >> 1)) >> 3);
    for (int line = 0; line < 4; ++line) {
        long base = pos + line * along;
        auto at = [&](int k) -> S& { return d[base + k * step]; };
        int p0 = at(-1), p1 = at(-2), p2 = at(-3), p3 = at(-4);
        int q0 = at(0), q1 = at(1), q2 = at(2), q3 = at(3);
        if (strong) {
            if (!no_p) {
                at(-1) = (S)clampi((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3, p0 - 2 * tc, p0 + 2 * tc);
                at(-2) = (S)clampi((p2 + p1 + p0 + q0 + 2) >> 2, p1 - 2 * tc, p1 + 2 * tc);
                at(-3) = (S)clampi((2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3, p2 - 2 * tc, p2 + 2 * tc);
            }
            if (!no_q) {
                at(0) = (S)clampi((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3, q0 - 2 * tc, q0 + 2 * tc);
                at(1) = (S)clampi((p0 + q0 + q1 + q2 + 2) >> 2, q1 - 2 * tc, q1 + 2 * tc);
                at(2) = (S)clampi((p0 + q0 + q1 + 3 * q2 + 2 * q3 + 4) >> 3, q2 - 2 * tc, q2 + 2 * tc);
            }
        } else {
            int delta = (9 * (q0 - p0) - 3 * (q1 - p1) + 8) >> 4;
            if (std::abs(delta) < tc * 10) {
                delta = clampi(delta, -tc, tc);
                if (!no_p) at(-1) = (S)clampi(p0 + delta, 0, max);
                if (!no_q) at(0) = (S)clampi(q0 - delta, 0, max);
                if (dep && !no_p) {
                    int dp_ = clampi((((p2 + p0 + 1) >> 1) - p1 + delta) >> 1, -(tc >> 1), tc >> 1);
                    at(-2) = (S)clampi(p1 + dp_, 0, max);
                }
                if (deq && !no_q) {
                    int dq_ = clampi((((q2 + q0 + 1) >> 1) - q1 - delta) >> 1, -(tc >> 1), tc >> 1);
                    at(1) = (S)clampi(q1 + dq_, 0, max);
                }
            }
        }
    }
}

// chroma_edge_scalar (:1326-1353).
template <typename S>
void chroma_edge(S* d, long pos, long step, long along, int n, int tc, bool no_p, bool no_q, int max) {
    for (int line = 0; line < n; ++line) {
        long base = pos + line * along;
        int p0 = d[base - step], p1 = d[base - 2 * step], q0 = d[base], q1 = d[base + step];
        int delta = clampi((((q0 - p0) << 2) + p1 - q1 + 4) >> 3, -tc, tc);
        if (!no_p) d[base - step] = (S)clampi(p0 + delta, 0, max);
        if (!no_q) d[base] = (S)clampi(q0 - delta, 0, max);
    }
}

// deblock_luma_scalar (:1355-1382): two 4-line segments.
template <typename S>
void deblock_luma(S* d, long off, long step, long along, const int* beta, const int* tc, const bool* np,
                  const bool* nq, int max) {
    for (int seg = 0; seg < 2; ++seg) {
        if (beta[seg] == 0 && tc[seg] == 0) continue;
        luma_edge(d, off + 4 * seg * along, step, along, beta[seg], tc[seg], np[seg], nq[seg], max);
    }
}
// deblock_chroma_scalar (:1384-1411): four 2-line segments.
template <typename S>
void deblock_chroma(S* d, long off, long step, long along, const int* tc, const bool* np, const bool* nq,
                    int max) {
    for (int seg = 0; seg < 4; ++seg) {
        if (tc[seg] == 0) continue;
        chroma_edge(d, off + 2 * seg * along, step, along, 2, tc[seg], np[seg], nq[seg], max);
    }
}

// ---------------------------------------------------------------------
// Picture state (src/hevc/frame.rs:12-17, 133-214; src/hevc/pic.rs)
// ---------------------------------------------------------------------
const int LUMA_PAD = 80, CHROMA_PAD = 40;

template <typename S> struct Plane {
    std::vector<S> data;
    int width = 0, height = 0, pad = 0;
    long stride = 0;
    void init(int w, int h, int p) {
        width = w;
        height = h;
        pad = p;
        stride = w + 2 * p;
        data.assign((size_t)stride * (h + 2 * p), 0);
    }
    long origin() const { return (long)pad * stride + pad; }
    long offset(long x, long y) const { return origin() + y * stride + x; }
    S at_clamped(int x, int y) const {
        x = clampi(x, 0, width - 1);
        y = clampi(y, 0, height - 1);
        return data[offset(x, y)];
    }
    void load(const S* src) {
        for (int y = 0; y < height; ++y) std::memcpy(&data[offset(0, y)], src + (long)y * width, width * sizeof(S));
    }
    void store(S* dst) const {
        for (int y = 0; y < height; ++y) std::memcpy(dst + (long)y * width, &data[offset(0, y)], width * sizeof(S));
    }
    // extend_rows / extend_top / extend_bottom (src/hevc/frame.rs:169-214).
    void extend() {
        for (int y = 0; y < height; ++y) {
            long row = origin() + y * stride;
            S l = data[row], r = data[row + width - 1];
            for (int i = 1; i <= pad; ++i) {
                data[row - i] = l;
                data[row + width - 1 + i] = r;
            }
        }
        long first = origin() - pad;
        for (int i = 1; i <= pad; ++i) std::memcpy(&data[first - i * stride], &data[first], stride * sizeof(S));
        long last = origin() + (long)(height - 1) * stride - pad;
        for (int i = 1; i <= pad; ++i) std::memcpy(&data[last + i * stride], &data[last], stride * sizeof(S));
    }
};

template <typename S> struct Frame {
    Plane<S> p[3];
    void init(int w, int h) {
        p[0].init(w, h, LUMA_PAD);
        p[1].init(w / 2, h / 2, CHROMA_PAD);
        p[2].init(w / 2, h / 2, CHROMA_PAD);
    }
};

// MotionInfo (src/hevc/frame.rs:44-58).
struct MotionInfo {
    int mv[2][2];
    int ref_delta[2];
    int ref_idx[2];
    int flags;  // bit 0/1 long-term per list, bit 2 intra
    bool uses(int l) const { return ref_idx[l] >= 0; }
    bool long_term(int l) const { return (flags >> l) & 1; }
};

struct PicInfo {
    int w4, h4, wc, hc, log2_ctb;
    std::vector<uint8_t> pred_mode, cbf_luma, edges, filter_exempt;
    std::vector<int8_t> qp_y;
    std::vector<MotionInfo> motion;
    // Slices and tiles (PicInfo::ctb_slice / ctb_slice_addr / slices,
    // Geometry::ctb_tile / ctb_rs_to_ts, src/hevc/pic.rs:41-96); empty:
    // one slice (`one`) and one tile.
    std::vector<uint16_t> ctb_slice, ctb_tile;
    std::vector<uint32_t> ctb_slice_addr, ctb_ts;
    std::vector<rv_hevc_slice> slices;
    rv_hevc_slice one{};
    bool lf_across_tiles = true;
    int idx4(int x, int y) const { return (y >> 2) * w4 + (x >> 2); }
    int ctb_of(int x, int y) const { return (y >> log2_ctb) * wc + (x >> log2_ctb); }
    bool multi() const { return !ctb_slice.empty(); }
    const rv_hevc_slice& slice_of_ctb(int c) const { return multi() ? slices[ctb_slice[c]] : one; }
};

// PicInfo::fill4 (src/hevc/pic.rs:374-406).
template <typename T> void fill4(std::vector<T>& a, int w4, int x, int y, int w, int h, T v) {
    for (int by = y >> 2; by < (y + h) >> 2; ++by)
        for (int bx = x >> 2; bx < (x + w) >> 2; ++bx) a[by * w4 + bx] = v;
}

// f(lo, hi) over [0, n) split into `threads` contiguous ranges, each on a
// thread of its own (the baseline's row and block parallelism; the
// reference spreads a picture's rows over its pool, src/threading.rs:176-356).
template <typename F> void parallel_for(int n, int threads, F f) {
    if (threads <= 1 || n <= 1) {
        f(0, n);
        return;
    }
    const int t = std::min(threads, n);
    std::vector<std::thread> ts;
    for (int i = 1; i < t; ++i)
        ts.emplace_back([&f, n, t, i] { f((int)((long long)n * i / t), (int)((long long)n * (i + 1) / t)); });
    f(0, (int)((long long)n / t));
    for (auto& th : ts) th.join();
}

// chroma_qp (src/hevc/ctu.rs:3072-3084), 4:2:0 (ChromaArrayType 1).
int chroma_qp420(int qpi) {
    if (qpi < 30) return qpi;
    if (qpi >= 43) return qpi - 6;
    static const int T[14] = {29, 30, 31, 32, 33, 33, 34, 34, 35, 35, 36, 36, 37, 37};
    return T[qpi - 30];
}

// ---------------------------------------------------------------------
// Inter prediction driver (src/hevc/inter.rs:86-195, 553-742)
// ---------------------------------------------------------------------
template <typename S> struct Mc {
    std::vector<int16_t> pred[2][3];
    std::vector<int16_t> tmp;
    std::vector<S> window;
    Mc() {
        for (auto& l : pred)
            for (auto& c : l) c.assign(64 * 64, 0);
        tmp.assign(64 * (64 + 7) + 64 * 64, 0);
        window.assign(71 * 71, 0);
    }
};

// source (src/hevc/inter.rs:92-125): the filter window, from the padded
// plane when it fits, else gathered with clamping.
template <typename S>
const S* source(std::vector<S>& window, const Plane<S>& pl, int xi, int yi, int w, int h, bool luma, long* stride) {
    int reach = luma ? 3 : 1, taps = luma ? 8 : 4;
    int x0 = xi - reach, y0 = yi - reach, ww = w + taps - 1, hh = h + taps - 1;
    int pad = pl.pad;
    bool inside = x0 >= -pad && y0 >= -pad && x0 + ww <= pl.width + pad && y0 + hh <= pl.height + pad;
    if (inside) {
        *stride = pl.stride;
        return &pl.data[pl.offset(x0, y0)];
    }
    for (int yy = 0; yy < hh; ++yy)
        for (int xx = 0; xx < ww; ++xx) window[yy * ww + xx] = pl.at_clamped(x0 + xx, y0 + yy);
    *stride = ww;
    return window.data();
}

// interp (src/hevc/inter.rs:131-195), through the scalar table's two-pass
// entries.
template <typename S>
void interp_drv(Mc<S>& mc, const Plane<S>& pl, int xi, int yi, int xf, int yf, int w, int h, bool luma, int bd,
                int16_t* out) {
    long ss;
    const S* src = source(mc.window, pl, xi, yi, w, h, luma, &ss);
    interp_k(luma, out, mc.tmp.data(), src, ss, w, h, xf, yf, bd);
}

// copy_rows (src/hevc/inter.rs:257-286).
template <typename S>
bool copy_rows(const Plane<S>& src, Plane<S>& dst, int sx, int sy, int dx, int dy, int bw, int bh) {
    int pad = src.pad;
    if (sx < -pad || sy < -pad || sx + bw > src.width + pad || sy + bh > src.height + pad) return false;
    for (int r = 0; r < bh; ++r)
        std::memcpy(&dst.data[dst.offset(dx, dy + r)], &src.data[src.offset(sx, sy + r)], bw * sizeof(S));
    return true;
}

struct Weight {
    bool explicit_;
    int log2_wd;
    int w[2], o[2];
};

// predict_block (src/hevc/inter.rs:555-742) with the scalar table
// (fused_mc = false, src/dsp/hevc.rs:286): direct copies for whole-sample
// default uni-prediction, else two-pass interpolation and the combiners.
template <typename S>
void predict_block(Mc<S>& mc, Frame<S>& cur, int bd, int x, int y, int w, int h, const Frame<S>* rf[2],
                   const int mv[2][2], const Weight wt[3]) {
    int cw = w / 2, ch = h / 2;
    bool both = rf[0] && rf[1];
    bool direct[3] = {false, false, false};
    if (!both) {
        int l = rf[0] ? 0 : 1;
        const Frame<S>& r = *rf[l];
        int mvx = mv[l][0], mvy = mv[l][1];
        auto plain = [&](int c) { return !wt[c].explicit_; };
        if ((mvx & 3) == 0 && (mvy & 3) == 0 && plain(0)) {
            direct[0] = copy_rows(r.p[0], cur.p[0], x + (mvx >> 2), y + (mvy >> 2), x, y, w, h);
        }
        // 4:2:0: chroma vector = mv in eighth chroma samples (:581-594).
        if ((mvx & 7) == 0 && (mvy & 7) == 0 && plain(1) && plain(2)) {
            int xci = x / 2 + (mvx >> 3), yci = y / 2 + (mvy >> 3);
            direct[1] = copy_rows(r.p[1], cur.p[1], xci, yci, x / 2, y / 2, cw, ch);
            direct[2] = copy_rows(r.p[2], cur.p[2], xci, yci, x / 2, y / 2, cw, ch);
        }
    }
    for (int list = 0; list < 2; ++list) {
        if (!rf[list]) continue;
        int mvx = mv[list][0], mvy = mv[list][1];
        for (int c = 0; c < 3; ++c) {
            if (direct[c]) continue;
            bool luma = c == 0;
            if (luma)
                interp_drv(mc, rf[list]->p[0], x + (mvx >> 2), y + (mvy >> 2), mvx & 3, mvy & 3, w, h, true, bd,
                           mc.pred[list][0].data());
            else
                interp_drv(mc, rf[list]->p[c], x / 2 + (mvx >> 3), y / 2 + (mvy >> 3), mvx & 7, mvy & 7, cw, ch,
                           false, bd, mc.pred[list][c].data());
        }
    }
    int max = (1 << bd) - 1;
    for (int c = 0; c < 3; ++c) {
        if (direct[c]) continue;
        Plane<S>& pl = cur.p[c];
        int px = c ? x / 2 : x, py = c ? y / 2 : y, pw = c ? cw : w, ph = c ? ch : h;
        S* dst = &pl.data[pl.offset(px, py)];
        const int16_t* a = mc.pred[0][c].data();
        const int16_t* b = mc.pred[1][c].data();
        if (!wt[c].explicit_) {
            if (!both) uni_k(dst, pl.stride, rf[0] ? a : b, pw, ph, 14 - bd, max);
            else bi_k(dst, pl.stride, a, b, pw, ph, 15 - bd, max);
        } else {
            if (!both) {
                int l = rf[0] ? 0 : 1;
                wuni_k(dst, pl.stride, l == 0 ? a : b, pw, ph, wt[c].log2_wd, wt[c].w[l], wt[c].o[l], max);
            } else {
                wbi_k(dst, pl.stride, a, b, pw, ph, wt[c].log2_wd, wt[c].w[0], wt[c].w[1], wt[c].o[0], wt[c].o[1],
                      max);
            }
        }
    }
}

// ---------------------------------------------------------------------
// Deblocking driver (src/hevc/deblock.rs), one slice and one tile.
// ---------------------------------------------------------------------

// motion_bs (src/hevc/deblock.rs:15-60).
int motion_bs(const MotionInfo& a, const MotionInfo& b) {
    int na = a.uses(0) + a.uses(1), nb = b.uses(0) + b.uses(1);
    if (na != nb) return 1;
    auto far = [](const int* p, const int* q) { return std::abs(p[0] - q[0]) >= 4 || std::abs(p[1] - q[1]) >= 4; };
    if (na == 1) {
        int la = a.uses(0) ? 0 : 1, lb = b.uses(0) ? 0 : 1;
        if (a.ref_delta[la] != b.ref_delta[lb] || a.long_term(la) != b.long_term(lb)) return 1;
        return far(a.mv[la], b.mv[lb]);
    }
    auto key = [](const MotionInfo& m, int l) { return std::make_pair(m.ref_delta[l], m.long_term(l)); };
    auto pa0 = key(a, 0), pa1 = key(a, 1), pb0 = key(b, 0), pb1 = key(b, 1);
    bool same_set = (pa0 == pb0 && pa1 == pb1) || (pa0 == pb1 && pa1 == pb0);
    if (!same_set) return 1;
    if (pa0 != pa1) {
        if (pa0 == pb0) return far(a.mv[0], b.mv[0]) || far(a.mv[1], b.mv[1]);
        return far(a.mv[0], b.mv[1]) || far(a.mv[1], b.mv[0]);
    }
    bool c1 = far(a.mv[0], b.mv[0]) || far(a.mv[1], b.mv[1]);
    bool c2 = far(a.mv[0], b.mv[1]) || far(a.mv[1], b.mv[0]);
    return c1 && c2;
}

// Whether the edge between CTB cq (holding the q block) and CTB cp may be
// filtered (src/hevc/deblock.rs:125-137, 153-165): across a tile boundary
// only with loop_filter_across_tiles, across a slice boundary only with
// the q block's slice's loop_filter_across_slices.
bool edge_crosses_ok(const PicInfo& info, int cq, int cp) {
    if (!info.multi() || cq == cp) return true;
    if (info.ctb_tile[cq] != info.ctb_tile[cp] && !info.lf_across_tiles) return false;
    if (info.ctb_slice_addr[cq] != info.ctb_slice_addr[cp] && !info.slice_of_ctb(cq).loop_filter_across_slices)
        return false;
    return true;
}

// boundary_strengths (src/hevc/deblock.rs:62-180) over the whole picture:
// edges whose q block lies in a slice with deblocking disabled carry none,
// and edges on CTB boundaries cross slices and tiles as edge_crosses_ok.
void boundary_strengths(const PicInfo& info, std::vector<uint8_t>& ver, std::vector<uint8_t>& hor) {
    int w4 = info.w4;
    ver.assign((size_t)w4 * info.h4, 0);
    hor.assign((size_t)w4 * info.h4, 0);
    const int ctb_mask = (1 << info.log2_ctb) - 1;
    for (int by = 0; by < info.h4; ++by) {
        int y = by * 4;
        bool hor_row = y > 0 && y % 8 == 0;
        int bx_start = hor_row ? 0 : 2, bx_step = hor_row ? 1 : 2;
        for (int bx = bx_start; bx < w4; bx += bx_step) {
            int x = bx * 4, q = by * w4 + bx;
            int edges = info.edges[q];
            bool want_v = x > 0 && x % 8 == 0 && (edges & 3) != 0;
            bool want_h = hor_row && (edges & 12) != 0;
            if (!want_v && !want_h) continue;
            const int cq = info.ctb_of(x, y);
            if (info.slice_of_ctb(cq).deblocking_disabled) continue;
            if (want_v && (x & ctb_mask) == 0 && !edge_crosses_ok(info, cq, cq - 1)) want_v = false;
            if (want_h && (y & ctb_mask) == 0 && !edge_crosses_ok(info, cq, cq - info.wc)) want_h = false;
            bool intra_q = info.pred_mode[q] == 1;
            if (want_v) {
                int p = q - 1;
                int b = (intra_q || info.pred_mode[p] == 1)                              ? 2
                        : ((edges & 1) && (info.cbf_luma[p] || info.cbf_luma[q])) ? 1
                                                                                  : motion_bs(info.motion[p], info.motion[q]);
                ver[q] = (uint8_t)b;
            }
            if (want_h) {
                int p = q - w4;
                int b = (intra_q || info.pred_mode[p] == 1)                              ? 2
                        : ((edges & 4) && (info.cbf_luma[p] || info.cbf_luma[q])) ? 1
                                                                                  : motion_bs(info.motion[p], info.motion[q]);
                hor[q] = (uint8_t)b;
            }
        }
    }
}

struct SliceParams {
    int beta_offset, tc_offset, cb_qp_offset, cr_qp_offset;
};

// deblock_rows (src/hevc/deblock.rs:210-515) for rows 0..h4 (the whole
// picture: the standard's order, which row-by-row application equals).
// The beta / tc offsets are those of the q block's slice (SliceCache,
// :186-201); the chroma QP offsets the PPS's (sl).
template <typename S>
void deblock_picture(Frame<S>& f, const PicInfo& info, const SliceParams& sl, int bd, int threads) {
    std::vector<uint8_t> ver, hor;
    boundary_strengths(info, ver, hor);
    int w4 = info.w4, h4 = info.h4;
    int max = (1 << bd) - 1, sh_l = bd - 8, sh_c = bd - 8;
    auto offsets = [&](int q, int* beta_offset, int* tc_offset) {
        if (!info.multi()) {
            *beta_offset = sl.beta_offset;
            *tc_offset = sl.tc_offset;
            return;
        }
        const rv_hevc_slice& s = info.slice_of_ctb(info.ctb_of((q % w4) * 4, (q / w4) * 4));
        *beta_offset = s.beta_offset;
        *tc_offset = s.tc_offset;
    };
    auto luma_params = [&](int b, int p, int q, int* beta, int* tc, bool* np, bool* nq) {
        int qp = (info.qp_y[p] + info.qp_y[q] + 1) >> 1, bo, to;
        offsets(q, &bo, &to);
        *beta = BETA_TABLE[clampi(qp + bo, 0, 51)] * (1 << sh_l);
        *tc = TC_TABLE[clampi(qp + 2 * (b - 1) + to, 0, 53)] * (1 << sh_l);
        *np = (info.filter_exempt[p] & 5) != 0;
        *nq = (info.filter_exempt[q] & 5) != 0;
    };
    auto chroma_tc = [&](int p, int q, int* out) {
        int qp_avg = (info.qp_y[p] + info.qp_y[q] + 1) >> 1, bo, to;
        offsets(q, &bo, &to);
        for (int c = 0; c < 2; ++c) {
            int qpi = qp_avg + (c == 0 ? sl.cb_qp_offset : sl.cr_qp_offset);
            int qpc = qpi < 0 ? qpi : chroma_qp420(qpi);
            out[c] = TC_TABLE[clampi(qpc + 2 + to, 0, 53)] * (1 << sh_c);
        }
    };
    // Each pass over bands of four 4x4 rows (16 luma rows): a vertical edge
    // changes only its own rows, and horizontal edges 8 rows apart change
    // disjoint rows (3 either side) and read none another changes, so bands
    // run in any order and on any thread.
    for (int pass = 0; pass < 2; ++pass) {
      const std::vector<uint8_t>& bs = pass == 0 ? ver : hor;
      parallel_for((h4 + 3) / 4, threads, [&](int band0, int band1) {
        const int lo = band0 * 4, hi = std::min(band1 * 4, h4);
        Plane<S>& Y = f.p[0];
        if (pass == 0) {
            for (int by = lo; by < hi; by += 2) {
                bool two = by + 1 < h4;
                for (int bx = 2; bx < w4; bx += 2) {
                    int b0 = bs[by * w4 + bx], b1 = two ? bs[(by + 1) * w4 + bx] : 0;
                    if ((b0 | b1) == 0) continue;
                    int beta[2] = {0, 0}, tc[2] = {0, 0};
                    bool np[2] = {false, false}, nq[2] = {false, false};
                    int bb[2] = {b0, b1};
                    for (int seg = 0; seg < 2; ++seg)
                        if (bb[seg]) {
                            int q = (by + seg) * w4 + bx;
                            luma_params(bb[seg], q - 1, q, &beta[seg], &tc[seg], &np[seg], &nq[seg]);
                        }
                    deblock_luma(Y.data.data(), Y.offset(bx * 4, by * 4), 1, Y.stride, beta, tc, np, nq, max);
                }
            }
        } else {
            for (int by = std::max(lo, 2); by < hi; by += 2) {
                for (int bx = 0; bx < w4; bx += 2) {
                    int b0 = bs[by * w4 + bx], b1 = bx + 1 < w4 ? bs[by * w4 + bx + 1] : 0;
                    if ((b0 | b1) == 0) continue;
                    int beta[2] = {0, 0}, tc[2] = {0, 0};
                    bool np[2] = {false, false}, nq[2] = {false, false};
                    int bb[2] = {b0, b1};
                    for (int seg = 0; seg < 2; ++seg)
                        if (bb[seg]) {
                            int q = by * w4 + bx + seg;
                            luma_params(bb[seg], q - w4, q, &beta[seg], &tc[seg], &np[seg], &nq[seg]);
                        }
                    deblock_luma(Y.data.data(), Y.offset(bx * 4, by * 4), Y.stride, 1, beta, tc, np, nq, max);
                }
            }
        }
        // Chroma, 4:2:0: bS == 2 edges on the 16-luma grid (:409-513).
        Plane<S>& Cb = f.p[1];
        Plane<S>& Cr = f.p[2];
        if (pass == 0) {
            for (int by = lo; by < hi; by += 4) {
                int cnt = std::min(h4 - by, 4);
                for (int bx = 0; bx < w4; bx += 4) {
                    bool any = false;
                    int tcs[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
                    bool np[4] = {false, false, false, false}, nq[4] = {false, false, false, false};
                    for (int seg = 0; seg < cnt; ++seg) {
                        int q = (by + seg) * w4 + bx;
                        if (bs[q] != 2) continue;
                        any = true;
                        int t[2];
                        chroma_tc(q - 1, q, t);
                        tcs[0][seg] = t[0];
                        tcs[1][seg] = t[1];
                        np[seg] = (info.filter_exempt[q - 1] & 5) != 0;
                        nq[seg] = (info.filter_exempt[q] & 5) != 0;
                    }
                    if (!any) continue;
                    int x = bx * 4, y = by * 4;
                    deblock_chroma(Cb.data.data(), Cb.offset(x / 2, y / 2), 1, Cb.stride, tcs[0], np, nq, max);
                    deblock_chroma(Cr.data.data(), Cr.offset(x / 2, y / 2), 1, Cr.stride, tcs[1], np, nq, max);
                }
            }
        } else {
            for (int by = lo; by < hi; ++by) {
                if ((by * 4) % 16 != 0) continue;
                for (int bx = 0; bx < w4; bx += 4) {
                    int cnt = std::min(w4 - bx, 4);
                    bool any = false;
                    int tcs[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
                    bool np[4] = {false, false, false, false}, nq[4] = {false, false, false, false};
                    for (int seg = 0; seg < cnt; ++seg) {
                        int q = by * w4 + bx + seg;
                        if (bs[q] != 2) continue;
                        any = true;
                        int t[2];
                        chroma_tc(q - w4, q, t);
                        tcs[0][seg] = t[0];
                        tcs[1][seg] = t[1];
                        np[seg] = (info.filter_exempt[q - w4] & 5) != 0;
                        nq[seg] = (info.filter_exempt[q] & 5) != 0;
                    }
                    if (!any) continue;
                    int x = bx * 4, y = by * 4;
                    deblock_chroma(Cb.data.data(), Cb.offset(x / 2, y / 2), Cb.stride, 1, tcs[0], np, nq, max);
                    deblock_chroma(Cr.data.data(), Cr.offset(x / 2, y / 2), Cr.stride, 1, tcs[1], np, nq, max);
                }
            }
        }
      });
    }
}

// ---------------------------------------------------------------------
// SAO driver (src/hevc/sao.rs:110-431): a neighbouring CTB is usable as
// Neighbours::of (:193-238) says, and a CTB holding a block exempt from the
// loop filters (filter_exempt bit 0) takes the per-sample path, which
// leaves the exempt samples alone (:140-151, :280-283).
// ---------------------------------------------------------------------
template <typename S>
void sao_ctb(const Plane<S>& src, Plane<S>& dst, const PicInfo& info, int x0, int y0, int w, int h, int sw,
             int sh, int bd, const rv_sao& p, const bool nb[8], bool exempt_any) {
    // nb: l, r, t, b, tl, tr, bl, br.
    int max = (1 << bd) - 1;
    long stride = src.stride;
    int pic_w = dst.width, pic_h = dst.height;
    auto at = [&](int x, int y) { return src.offset(x, y); };
    auto exempt = [&](int x, int y) {
        return exempt_any && (info.filter_exempt[info.idx4(x * sw, y * sh)] & 1) != 0;
    };
    if (p.type_idx == 1) {
        int shift = bd - 5;
        int16_t table[32] = {0};
        for (int k = 0; k < 4; ++k) table[(k + p.band_or_class) & 31] = p.offsets[k];
        long off = at(x0, y0);
        if (!exempt_any) {
            sao_band_k(&dst.data[off], stride, &src.data[off], stride, w, h, table, shift, max);
            return;
        }
        for (int y = y0; y < y0 + h; ++y)
            for (int x = x0; x < x0 + w; ++x) {
                if (exempt(x, y)) continue;
                long i = at(x, y);
                int v = src.data[i];
                dst.data[i] = (S)clampi(v + table[v >> shift], 0, max);
            }
        return;
    }
    if (p.type_idx != 2) return;
    static const int HP[4][2] = {{-1, 1}, {0, 0}, {-1, 1}, {1, -1}};
    static const int VP[4][2] = {{0, 0}, {-1, 1}, {-1, 1}, {-1, 1}};
    const int* hp = HP[p.band_or_class & 3];
    const int* vp = VP[p.band_or_class & 3];
    int16_t off_tab[5] = {p.offsets[0], p.offsets[1], 0, p.offsets[2], p.offsets[3]};
    bool L = nb[0], R = nb[1], Tt = nb[2], B = nb[3], TL = nb[4], TR = nb[5], BL = nb[6], BR = nb[7];
    int xs = x0, xe = x0 + w, ys = y0, ye = y0 + h;
    bool uses_x = hp[0] != 0, uses_y = vp[0] != 0, diag = uses_x && uses_y;
    if ((uses_x && !L) || (diag && (!TL || !BL))) xs += 1;
    if ((uses_x && !R) || (diag && (!TR || !BR))) xe -= 1;
    if ((uses_y && !Tt) || (diag && (!TL || !TR))) ys += 1;
    if ((uses_y && !B) || (diag && (!BL || !BR))) ye -= 1;
    long na = vp[0] * stride + hp[0], nbb = vp[1] * stride + hp[1];
    bool interior_ok = !exempt_any && xs < xe && ys < ye;
    if (interior_ok) sao_edge_k(dst.data.data(), src.data.data(), at(xs, ys), stride, xe - xs, ye - ys, na, nbb, off_tab, max);
    auto usable = [&](int x, int y, int xn, int yn) -> bool {
        if (xn < 0 || yn < 0 || xn >= pic_w || yn >= pic_h) return false;
        int cx = (xn * sw) >> info.log2_ctb, cy = (yn * sh) >> info.log2_ctb;
        int cx0 = (x * sw) >> info.log2_ctb, cy0 = (y * sh) >> info.log2_ctb;
        int dx = cx - cx0, dy = cy - cy0;
        if (dx == 0 && dy == 0) return true;
        if (dx == -1 && dy == 0) return L;
        if (dx == 1 && dy == 0) return R;
        if (dx == 0 && dy == -1) return Tt;
        if (dx == 0 && dy == 1) return B;
        if (dx == -1 && dy == -1) return TL;
        if (dx == 1 && dy == -1) return TR;
        if (dx == -1 && dy == 1) return BL;
        return BR;
    };
    auto sample = [&](int x, int y) {
        if (exempt(x, y)) return;
        int xa = x + hp[0], ya = y + vp[0], xb = x + hp[1], yb = y + vp[1];
        if (!usable(x, y, xa, ya) || !usable(x, y, xb, yb)) return;
        long i = at(x, y);
        int v = src.data[i], a = src.data[i + na], b = src.data[i + nbb];
        int e = 2 + sgn(v - a) + sgn(v - b);
        dst.data[i] = (S)clampi(v + off_tab[e], 0, max);
    };
    if (interior_ok) {
        for (int y = y0; y < ys; ++y)
            for (int x = x0; x < x0 + w; ++x) sample(x, y);
        for (int y = ys; y < ye; ++y) {
            for (int x = x0; x < xs; ++x) sample(x, y);
            for (int x = xe; x < x0 + w; ++x) sample(x, y);
        }
        for (int y = ye; y < y0 + h; ++y)
            for (int x = x0; x < x0 + w; ++x) sample(x, y);
    } else {
        for (int y = y0; y < y0 + h; ++y)
            for (int x = x0; x < x0 + w; ++x) sample(x, y);
    }
}

// Neighbours::of (src/hevc/sao.rs:193-238): which of the eight CTBs around
// (rx, ry) -- l, r, t, b, tl, tr, bl, br -- this CTB's SAO may read: inside
// the picture; across a slice boundary only if the later slice in tile-scan
// order allows filtering across its slice boundaries; across a tile
// boundary only with loop_filter_across_tiles.
void sao_neighbours(const PicInfo& info, int rx, int ry, bool nb[8]) {
    static const int DX[8] = {-1, 1, 0, 0, -1, 1, -1, 1}, DY[8] = {0, 0, -1, 1, -1, -1, 1, 1};
    const int cur = ry * info.wc + rx;
    for (int k = 0; k < 8; ++k) {
        const int nx = rx + DX[k], ny = ry + DY[k];
        if (nx < 0 || ny < 0 || nx >= info.wc || ny >= info.hc) {
            nb[k] = false;
            continue;
        }
        if (!info.multi()) {
            nb[k] = true;
            continue;
        }
        const int n = ny * info.wc + nx;
        bool ok = true;
        if (info.ctb_slice_addr[n] != info.ctb_slice_addr[cur]) {
            if (info.ctb_ts[n] < info.ctb_ts[cur] && !info.slice_of_ctb(cur).loop_filter_across_slices) ok = false;
            if (info.ctb_ts[cur] < info.ctb_ts[n] && !info.slice_of_ctb(n).loop_filter_across_slices) ok = false;
        }
        if (!info.lf_across_tiles && info.ctb_tile[n] != info.ctb_tile[cur]) ok = false;
        nb[k] = ok;
    }
}

template <typename S>
void sao_picture(Frame<S>& f, const Frame<S>& deb, const PicInfo& info, const rv_hevc_pic* pic, int threads) {
    int ctb = 1 << info.log2_ctb;
    int pw = f.p[0].width, ph = f.p[0].height;
    // CTBs read the deblocked copy and write only their own samples: rows of
    // CTBs run on any thread.
    parallel_for(info.hc, threads, [&](int ry0, int ry1) {
    for (int ry = ry0; ry < ry1; ++ry)
        for (int rx = 0; rx < info.wc; ++rx) {
            const rv_sao* params = &pic->sao[3 * (ry * info.wc + rx)];
            rv_sao eff[3];
            for (int c = 0; c < 3; ++c) {
                eff[c] = params[c];
                bool on = c == 0 ? pic->sao_luma : pic->sao_chroma;
                if (!on) eff[c].type_idx = 0;
            }
            if (eff[0].type_idx == 0 && eff[1].type_idx == 0 && eff[2].type_idx == 0) continue;
            bool nb[8];
            sao_neighbours(info, rx, ry, nb);
            bool exempt_any = false;
            {
                const int x0 = rx * ctb, y0 = ry * ctb;
                const int w = std::min(ctb, pw - x0), h = std::min(ctb, ph - y0);
                for (int by = y0 >> 2; by < (y0 + h) >> 2 && !exempt_any; ++by)
                    for (int bx = x0 >> 2; bx < (x0 + w) >> 2; ++bx)
                        if (info.filter_exempt[by * info.w4 + bx] & 1) {
                            exempt_any = true;
                            break;
                        }
            }
            for (int c = 0; c < 3; ++c) {
                if (eff[c].type_idx == 0) continue;
                int sw = c ? 2 : 1, sh = c ? 2 : 1;
                int x0 = rx * ctb / sw, y0 = ry * ctb / sh;
                int w = std::min(ctb / sw, pw / sw - x0), h = std::min(ctb / sh, ph / sh - y0);
                sao_ctb(deb.p[c], f.p[c], info, x0, y0, w, h, sw, sh, pic->bit_depth, eff[c], nb, exempt_any);
            }
        }
    });
}

// ---------------------------------------------------------------------
// Intra prediction (src/hevc/intra.rs; the caller's availability and
// flags, src/hevc/ctu.rs:1727-1833)
// ---------------------------------------------------------------------

// intraPredAngle, modes 2..34 (Table 8-4), and invAngle, modes 11..25
// (Table 8-5) (src/hevc/intra.rs:20-27).
const int INTRA_ANGLE[33] = {32,  26,  21,  17,  13,  9,   5,   2,   0,   -2, -5, -9, -13, -17, -21, -26, -32,
                             -26, -21, -17, -13, -9,  -5,  -2,  0,   2,   5,  9,  13,  17,  21,  26,  32};
const int INV_ANGLE[15] = {-4096, -1638, -910, -630, -482, -390, -315, -256, -315, -390, -482, -630, -910, -1638, -4096};

// A block's reference samples in substitution (scan) order: r[0..2n) =
// p[-1][2n-1] .. p[-1][0], r[2n] = p[-1][-1], r[2n+1..4n] = p[0][-1] ..
// p[2n-1][-1].
struct IntraRefs {
    int n = 0;
    uint16_t r[129];
    int left(int y) const { return r[2 * n - 1 - y]; }
    int top(int x) const { return r[2 * n + 1 + x]; }
    int corner() const { return r[2 * n]; }
};

// prepare (src/hevc/intra.rs:159-259), 8.4.4.2.2: unavailable samples take
// the value of the previous one in scan order; a leading run takes the first
// available one; with none available, all are 1 << (bitDepth - 1).
void intra_substitute(IntraRefs& R, const bool* avail, int bd) {
    const int m = 4 * R.n + 1;
    int first = -1;
    for (int j = 0; j < m && first < 0; ++j)
        if (avail[j]) first = j;
    if (first < 0) {
        for (int j = 0; j < m; ++j) R.r[j] = (uint16_t)(1 << (bd - 1));
        return;
    }
    if (!avail[0]) R.r[0] = R.r[first];
    for (int j = 1; j < m; ++j)
        if (!avail[j]) R.r[j] = R.r[j - 1];
}

// smooth (src/hevc/intra.rs:261-309), 8.4.4.2.3: the [1 2 1] filter along
// the scan order (its ends kept), or for a flat 32x32 luma block under the
// SPS flag the bilinear strong filter between the corner and the far ends.
IntraRefs intra_smooth(const IntraRefs& R, bool strong, int bd) {
    IntraRefs F = R;
    const int n = R.n, m = 4 * n + 1;
    const int c = R.corner();
    const bool bi = strong && n == 32 && std::abs(c + R.top(63) - 2 * R.top(31)) < (1 << (bd - 5)) &&
                    std::abs(c + R.left(63) - 2 * R.left(31)) < (1 << (bd - 5));
    if (bi) {
        for (int i = 0; i < 63; ++i) {
            F.r[2 * n - 1 - i] = (uint16_t)(((63 - i) * c + (i + 1) * R.left(63) + 32) >> 6);
            F.r[2 * n + 1 + i] = (uint16_t)(((63 - i) * c + (i + 1) * R.top(63) + 32) >> 6);
        }
        return F;
    }
    for (int j = 1; j < m - 1; ++j) F.r[j] = (uint16_t)((R.r[j - 1] + 2 * R.r[j] + R.r[j + 1] + 2) >> 2);
    return F;
}

// predict_prepared (src/hevc/intra.rs:311-426) with the scalar predictors
// (src/dsp/hevc.rs:1093-1166): the filter decision (8.4.4.2.3), planar, DC
// with its edge filter, angular with the ref[] of 8.4.4.2.6 and the
// boundary filters of modes 10 / 26. `luma`: the component is filtered and
// gets the boundary filters (4:2:0). Writes dst[y * n + x].
void intra_pred(int* dst, const IntraRefs& raw, int mode, bool luma, bool strong, int bd) {
    const int n = raw.n;
    int log2n = 0;
    while ((1 << log2n) < n) ++log2n;
    bool smoothed = false;
    if (luma && mode != 1 && n != 4) {
        const int dist = std::min(std::abs(mode - 26), std::abs(mode - 10));
        const int thres = n == 8 ? 7 : n == 16 ? 1 : n == 32 ? 0 : 10;
        smoothed = dist > thres;
    }
    const IntraRefs R = smoothed ? intra_smooth(raw, strong, bd) : raw;
    const int max = (1 << bd) - 1;
    const bool edge = luma && n < 32;
    if (mode == 0) {
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x)
                dst[y * n + x] = ((n - 1 - x) * R.left(y) + (x + 1) * R.top(n) + (n - 1 - y) * R.top(x) +
                                  (y + 1) * R.left(n) + n) >> (log2n + 1);
        return;
    }
    if (mode == 1) {
        int sum = n;
        for (int i = 0; i < n; ++i) sum += R.top(i) + R.left(i);
        const int dc = sum >> (log2n + 1);
        for (int i = 0; i < n * n; ++i) dst[i] = dc;
        if (edge) {
            dst[0] = (R.left(0) + 2 * dc + R.top(0) + 2) >> 2;
            for (int x = 1; x < n; ++x) dst[x] = (R.top(x) + 3 * dc + 2) >> 2;
            for (int y = 1; y < n; ++y) dst[y * n] = (R.left(y) + 3 * dc + 2) >> 2;
        }
        return;
    }
    const int angle = INTRA_ANGLE[mode - 2];
    const bool vertical = mode >= 18;
    // ref[k], k = -n .. 2n, at refb[n + k].
    int refb[3 * 32 + 2] = {};  // one past 3n: read (and unused) when iFact is 0
    auto main_ = [&](int i) { return vertical ? R.top(i) : R.left(i); };
    auto side = [&](int i) { return vertical ? R.left(i) : R.top(i); };
    refb[n] = R.corner();
    for (int i = 0; i < n; ++i) refb[n + 1 + i] = main_(i);
    if (angle < 0) {
        const int last = (n * angle) >> 5;
        if (last < -1) {
            const int inv = INV_ANGLE[mode - 11];
            for (int k = last; k <= -1; ++k) {
                const int idx = -1 + ((k * inv + 128) >> 8);
                refb[n + k] = idx < 0 ? R.corner() : side(idx);
            }
        }
    } else {
        for (int i = n; i < 2 * n; ++i) refb[n + 1 + i] = main_(i);
    }
    for (int y = 0; y < n; ++y) {
        const int i_idx = ((y + 1) * angle) >> 5, i_fact = ((y + 1) * angle) & 31;
        for (int x = 0; x < n; ++x) {
            const int a = refb[n + x + i_idx + 1], b = refb[n + x + i_idx + 2];
            const int v = i_fact ? ((32 - i_fact) * a + i_fact * b + 16) >> 5 : a;
            dst[vertical ? y * n + x : x * n + y] = v;
        }
    }
    if (edge) {
        const int c = R.corner();
        if (mode == 26)
            for (int y = 0; y < n; ++y) dst[y * n] = clampi(R.top(0) + ((R.left(y) - c) >> 1), 0, max);
        else if (mode == 10)
            for (int x = 0; x < n; ++x) dst[x] = clampi(R.left(0) + ((R.top(x) - c) >> 1), 0, max);
    }
}

// MinTbAddrZs (6.5.2) at 4x4 granularity for one tile: the CTB's raster
// address, then the z-order of the 4x4 block inside it.
uint32_t min_tb_addr_zs(int bx, int by, int log2_ctb, int wc) {
    const int sh = log2_ctb - 2;
    uint32_t z = (uint32_t)((by >> sh) * wc + (bx >> sh)) << (2 * sh);
    for (int i = 0; i < sh; ++i) {
        const int m = 1 << i;
        z += (m & bx ? (uint32_t)(m * m) : 0u) + (m & by ? (uint32_t)(2 * m * m) : 0u);
    }
    return z;
}

// The availability of a transform block's reference samples
// (src/hevc/ctu.rs:1732-1821 with PicInfo::available_at, src/hevc/pic.rs:
// 344-367): per 4x4 luma block, inside the picture, earlier in z-scan order
// (one slice and one tile: everything earlier is decoded), and with
// constrained intra prediction an intra block. avail in scan order.
void intra_avail(const rv_hevc_pic* pic, const std::vector<uint8_t>& intra4, int comp, int x, int y, int n,
                 bool* avail) {
    const int sc = comp ? 1 : 0, w4 = pic->width / 4;
    const int wc = (pic->width + (1 << pic->log2_ctb) - 1) >> pic->log2_ctb;
    const uint32_t zc = min_tb_addr_zs((x << sc) >> 2, (y << sc) >> 2, pic->log2_ctb, wc);
    auto check = [&](int xs, int ys) {
        const int xl = xs << sc, yl = ys << sc;
        if (xs < 0 || ys < 0 || xl >= pic->width || yl >= pic->height) return false;
        const int bx = xl >> 2, by = yl >> 2;
        if (min_tb_addr_zs(bx, by, pic->log2_ctb, wc) > zc) return false;
        return !pic->constrained_intra_pred || intra4[(size_t)by * w4 + bx] != 0;
    };
    for (int j = 0; j < 2 * n; ++j) avail[j] = check(x - 1, y + 2 * n - 1 - j);
    avail[2 * n] = check(x - 1, y - 1);
    for (int j = 0; j < 2 * n; ++j) avail[2 * n + 1 + j] = check(x + j, y - 1);
}

// IntraPredModeY / IntraPredModeC of a TU of an RV_INTRA_DEVICE CU: the
// luma mode of the PU holding it (NxN: the quadrant), the chroma mode.
int tu_intra_mode(const rv_hevc_cu& cu, const rv_hevc_tu& tu) {
    if (tu.comp) return cu.chroma_mode;
    if (cu.part_mode != 3) return cu.luma_mode[0];
    const int h = 1 << (cu.log2_size - 1);
    return cu.luma_mode[(tu.y >= cu.y + h ? 2 : 0) + (tu.x >= cu.x + h ? 1 : 0)];
}

// ---------------------------------------------------------------------
// Whole picture
// ---------------------------------------------------------------------

// The per-4x4 side data the deblocking filter reads (PicInfo,
// src/hevc/pic.rs:64-96), from the work lists: QP, prediction mode and
// coding-block edges per CU (src/hevc/ctu.rs:751-800), motion and
// prediction-block edges per PU, transform edges and cbf per luma TU
// (src/hevc/ctu.rs:1493-1510).
int build_info(const rv_hevc_pic* pic, PicInfo& info) {
    int W = pic->width, H = pic->height;
    info.log2_ctb = pic->log2_ctb;
    info.w4 = W / 4;
    info.h4 = H / 4;
    int ctb = 1 << pic->log2_ctb;
    info.wc = (W + ctb - 1) / ctb;
    info.hc = (H + ctb - 1) / ctb;
    size_t n4 = (size_t)info.w4 * info.h4;
    info.pred_mode.assign(n4, 0);
    info.cbf_luma.assign(n4, 0);
    info.edges.assign(n4, 0);
    info.filter_exempt.assign(n4, 0);
    info.qp_y.assign(n4, 0);
    MotionInfo none{};
    none.ref_idx[0] = none.ref_idx[1] = -1;
    info.motion.assign(n4, none);
    info.one.beta_offset = pic->beta_offset;
    info.one.tc_offset = pic->tc_offset;
    info.one.deblocking_disabled = 0;
    info.one.loop_filter_across_slices = 1;
    if (pic->ctb_slice) {
        const size_t nc = (size_t)info.wc * info.hc;
        if (pic->n_slices <= 0 || !pic->slices || !pic->ctb_slice_addr || !pic->ctb_tile || !pic->ctb_ts)
            return RV_E_INVALID;
        info.slices.assign(pic->slices, pic->slices + pic->n_slices);
        info.ctb_slice.assign(pic->ctb_slice, pic->ctb_slice + nc);
        info.ctb_slice_addr.assign(pic->ctb_slice_addr, pic->ctb_slice_addr + nc);
        info.ctb_tile.assign(pic->ctb_tile, pic->ctb_tile + nc);
        info.ctb_ts.assign(pic->ctb_ts, pic->ctb_ts + nc);
        for (size_t c = 0; c < nc; ++c)
            if (info.ctb_slice[c] >= pic->n_slices || info.ctb_ts[c] >= nc) return RV_E_RANGE;
        info.lf_across_tiles = pic->loop_filter_across_tiles != 0;
    }
    for (int i = 0; i < pic->n_cus; ++i) {
        const rv_hevc_cu& cu = pic->cus[i];
        int s = 1 << cu.log2_size;
        fill4(info.qp_y, info.w4, cu.x, cu.y, s, s, (int8_t)cu.qp_y);
        if (cu.exempt) fill4(info.filter_exempt, info.w4, cu.x, cu.y, s, s, (uint8_t)cu.exempt);
        for (int yy = cu.y; yy < cu.y + s; yy += 4) info.edges[info.idx4(cu.x, yy)] |= 1;
        for (int xx = cu.x; xx < cu.x + s; xx += 4) info.edges[info.idx4(xx, cu.y)] |= 4;
        if (cu.intra) {
            fill4(info.pred_mode, info.w4, cu.x, cu.y, s, s, (uint8_t)1);
            MotionInfo mi = none;
            mi.flags = 4;
            fill4(info.motion, info.w4, cu.x, cu.y, s, s, mi);
            // PU edges of the single intra PB.
            for (int yy = cu.y; yy < cu.y + s; yy += 4) info.edges[info.idx4(cu.x, yy)] |= 2;
            for (int xx = cu.x; xx < cu.x + s; xx += 4) info.edges[info.idx4(xx, cu.y)] |= 8;
        }
    }
    for (int i = 0; i < pic->n_pus; ++i) {
        const rv_hevc_pu& pu = pic->pus[i];
        if (!(pu.dir & 3)) return RV_E_RANGE;
        MotionInfo mi = none;
        for (int l = 0; l < 2; ++l) {
            if (!((pu.dir >> l) & 1)) continue;
            int ri = pu.ref_idx[l];
            if (ri < 0 || ri >= RV_MAX_REFS || ri >= pic->num_ref[l]) return RV_E_RANGE;
            mi.ref_idx[l] = ri;
            mi.mv[l][0] = pu.mv[l].x;
            mi.mv[l][1] = pu.mv[l].y;
            mi.ref_delta[l] = pic->poc - pic->ref_poc[l][ri];
            if (pic->ref_long_term[l][ri]) mi.flags |= 1 << l;
        }
        fill4(info.motion, info.w4, pu.x, pu.y, pu.w, pu.h, mi);
        for (int yy = pu.y; yy < std::min(pu.y + pu.h, H); yy += 4) info.edges[info.idx4(pu.x, yy)] |= 2;
        for (int xx = pu.x; xx < std::min(pu.x + pu.w, W); xx += 4) info.edges[info.idx4(xx, pu.y)] |= 8;
    }
    for (int i = 0; i < pic->n_tus; ++i) {
        const rv_hevc_tu& tu = pic->tus[i];
        int n = 1 << tu.log2;
        if (tu.comp == 0) {
            for (int yy = tu.y; yy < std::min(tu.y + n, H); yy += 4) info.edges[info.idx4(tu.x, yy)] |= 1;
            for (int xx = tu.x; xx < std::min(tu.x + n, W); xx += 4) info.edges[info.idx4(xx, tu.y)] |= 4;
            if (tu.flags & RV_TU_CODED) fill4(info.cbf_luma, info.w4, tu.x, tu.y, n, n, (uint8_t)1);
        }
    }
    return RV_OK;
}

template <typename S>
int reconstruct(const rv_hevc_pic* pic, const void* const* slot_planes, int n_slots, void* const out[3],
                int stop_after, int threads) {
    int W = pic->width, H = pic->height, bd = pic->bit_depth;
    if (W <= 0 || H <= 0 || (W & 7) || (H & 7)) return RV_E_RANGE;
    // Reference frames (padded, border extended, as a decoded picture is).
    std::vector<Frame<S>> refs(n_slots);
    std::vector<bool> loaded(n_slots, false);
    auto ref_frame = [&](int slot) -> const Frame<S>* {
        if (slot < 0 || slot >= n_slots || !slot_planes[3 * slot]) return nullptr;
        if (!loaded[slot]) {
            refs[slot].init(W, H);
            for (int c = 0; c < 3; ++c) {
                refs[slot].p[c].load((const S*)slot_planes[3 * slot + c]);
                refs[slot].p[c].extend();
            }
            loaded[slot] = true;
        }
        return &refs[slot];
    };
    Frame<S> cur;
    cur.init(W, H);
    PicInfo info;
    int rc = build_info(pic, info);
    if (rc) return rc;
    int max = (1 << bd) - 1;

    // RV_INTRA_HOST CUs get their host-supplied prediction.
    for (int i = 0; i < pic->n_cus; ++i) {
        const rv_hevc_cu& cu = pic->cus[i];
        if (cu.intra != RV_INTRA_HOST) continue;
        int s = 1 << cu.log2_size;
        const uint16_t* pr = pic->intra_pred + cu.pred_off;
        for (int c = 0; c < 3; ++c) {
            int n = c ? s / 2 : s, px = c ? cu.x / 2 : cu.x, py = c ? cu.y / 2 : cu.y;
            Plane<S>& pl = cur.p[c];
            for (int yy = 0; yy < n; ++yy)
                for (int xx = 0; xx < n; ++xx) pl.data[pl.offset(px + xx, py + yy)] = (S)pr[yy * n + xx];
            pr += n * n;
        }
    }
    // Every reference a PU names, loaded before the PUs run on threads.
    for (int i = 0; i < pic->n_pus; ++i)
        for (int l = 0; l < 2; ++l)
            if (((pic->pus[i].dir >> l) & 1) && !ref_frame(pic->ref_slot[l][pic->pus[i].ref_idx[l]])) return RV_E_RANGE;
    // Prediction units: disjoint outputs, so ranges of PUs run on threads,
    // each with its own filter scratch.
    parallel_for(pic->n_pus, threads, [&](int i0, int i1) {
        Mc<S> mc;
        for (int i = i0; i < i1; ++i) {
            const rv_hevc_pu& pu = pic->pus[i];
            const Frame<S>* rf[2] = {nullptr, nullptr};
            int mv[2][2] = {{0, 0}, {0, 0}};
            for (int l = 0; l < 2; ++l) {
                if (!((pu.dir >> l) & 1)) continue;
                rf[l] = ref_frame(pic->ref_slot[l][pu.ref_idx[l]]);
                mv[l][0] = pu.mv[l].x;
                mv[l][1] = pu.mv[l].y;
            }
            // Weighting (src/hevc/ctu.rs:1254-1292).
            Weight wt[3];
            for (int c = 0; c < 3; ++c) {
                wt[c].explicit_ = pic->weighted != 0;
                wt[c].log2_wd = pic->log2_wd[c];
                for (int l = 0; l < 2; ++l) {
                    wt[c].w[l] = 1;
                    wt[c].o[l] = 0;
                    if (rf[l]) {
                        wt[c].w[l] = pic->wp_w[l][pu.ref_idx[l]][c];
                        wt[c].o[l] = pic->wp_o[l][pu.ref_idx[l]][c];
                    }
                }
            }
            predict_block(mc, cur, bd, pu.x, pu.y, pu.w, pu.h, rf, mv, wt);
        }
    });
    // The CU of each 4x4 luma block, where a CU is predicted on the device.
    std::vector<int32_t> cu_at;
    for (int i = 0; i < pic->n_cus; ++i)
        if (pic->cus[i].intra == RV_INTRA_DEVICE) {
            if (cu_at.empty()) cu_at.assign((size_t)info.w4 * info.h4, -1);
            const rv_hevc_cu& cu = pic->cus[i];
            if (cu.part_mode != 0 && cu.part_mode != 3) return RV_E_RANGE;
            for (int k = 0; k < (cu.part_mode == 3 ? 4 : 1); ++k)
                if (cu.luma_mode[k] > 34) return RV_E_RANGE;
            if (cu.chroma_mode > 34) return RV_E_RANGE;
            fill4(cu_at, info.w4, cu.x, cu.y, 1 << cu.log2_size, 1 << cu.log2_size, (int32_t)i);
        }
    auto device_intra_cu = [&](const rv_hevc_tu& tu) -> int {
        if (cu_at.empty()) return -1;
        const int sc = tu.comp ? 1 : 0;
        return cu_at[info.idx4(tu.x << sc, tu.y << sc)];
    };
    // Residual (src/hevc/ctu.rs:2031-2088): TUs of a component do not
    // overlap, so ranges of TUs run on threads. The TUs of RV_INTRA_DEVICE
    // CUs follow, in decoding order.
    if (stop_after != 1) {
        for (int i = 0; i < pic->n_tus; ++i) {
            const rv_hevc_tu& tu = pic->tus[i];
            if (!(tu.flags & RV_TU_CODED)) continue;
            int n = 1 << tu.log2;
            if (tu.coef_off + (uint64_t)tu.n_nz > pic->n_nz) return RV_E_RANGE;
            for (uint32_t k = 0; k < tu.n_nz; ++k)
                if (pic->nz[tu.coef_off + k].pos >= n * n) return RV_E_RANGE;
            if ((tu.flags & RV_TU_LEVELS) && tu.qp > 51 + 6 * (bd - 8)) return RV_E_RANGE;
        }
        auto residual = [&](const rv_hevc_tu& tu, std::vector<int16_t>& coeffs) {
            int n = 1 << tu.log2;
            // The dense raster the reference's IdctFn takes: the nonzero
            // list scattered into zeros.
            std::fill(coeffs.begin(), coeffs.begin() + n * n, (int16_t)0);
            for (uint32_t k = 0; k < tu.n_nz; ++k) {
                const rv_nz& e = pic->nz[tu.coef_off + k];
                coeffs[e.pos] = e.val;
            }
            if (tu.flags & RV_TU_LEVELS) {
                int dc;
                const uint8_t* list = tu_scaling(pic->scaling, tu.log2, tu.comp, (tu.flags & RV_TU_INTRA) != 0, &dc);
                scale_coefficients(coeffs.data(), tu.log2, tu.qp, bd, list, dc, tu.max_x, tu.max_y);
            }
            int bd_shift = 20 - bd;
            if (tu.flags & RV_TU_DST) idst4(coeffs.data(), bd_shift);
            else idct(coeffs.data(), n, bd_shift, tu.max_x, tu.max_y);
            Plane<S>& pl = cur.p[tu.comp];
            add_residual(&pl.data[pl.offset(tu.x, tu.y)], pl.stride, coeffs.data(), n, max);
        };
        parallel_for(pic->n_tus, threads, [&](int i0, int i1) {
            std::vector<int16_t> coeffs(32 * 32);
            for (int i = i0; i < i1; ++i) {
                const rv_hevc_tu& tu = pic->tus[i];
                if (!(tu.flags & RV_TU_CODED) || device_intra_cu(tu) >= 0) continue;
                residual(tu, coeffs);
            }
        });
        // Intra prediction of each TU, then its residual (src/hevc/ctu.rs:
        // 1727-1833, 2016-2088), in list (decoding) order.
        if (!cu_at.empty()) {
            std::vector<int16_t> coeffs(32 * 32);
            std::vector<int> pred(32 * 32);
            for (int i = 0; i < pic->n_tus; ++i) {
                const rv_hevc_tu& tu = pic->tus[i];
                const int ci = device_intra_cu(tu);
                if (ci < 0) continue;
                const int n = 1 << tu.log2;
                IntraRefs R;
                R.n = n;
                bool avail[129];
                intra_avail(pic, info.pred_mode, tu.comp, tu.x, tu.y, n, avail);
                Plane<S>& pl = cur.p[tu.comp];
                for (int j = 0; j < 4 * n + 1; ++j) {
                    if (!avail[j]) continue;
                    const int xs = j < 2 * n ? tu.x - 1 : j == 2 * n ? tu.x - 1 : tu.x + (j - 2 * n - 1);
                    const int ys = j < 2 * n ? tu.y + (2 * n - 1 - j) : tu.y - 1;
                    R.r[j] = (uint16_t)pl.data[pl.offset(xs, ys)];
                }
                intra_substitute(R, avail, bd);
                intra_pred(pred.data(), R, tu_intra_mode(pic->cus[ci], tu), tu.comp == 0,
                           pic->strong_intra_smoothing != 0, bd);
                for (int y = 0; y < n; ++y)
                    for (int x = 0; x < n; ++x) pl.data[pl.offset(tu.x + x, tu.y + y)] = (S)pred[y * n + x];
                if (tu.flags & RV_TU_CODED) residual(tu, coeffs);
            }
        }
    }
    if (stop_after == 0 || stop_after >= 3) {
        if (pic->deblock) {
            SliceParams sl{pic->beta_offset, pic->tc_offset, pic->cb_qp_offset, pic->cr_qp_offset};
            deblock_picture(cur, info, sl, bd, threads);
        }
    }
    if ((stop_after == 0 || stop_after >= 4) && (pic->sao_luma || pic->sao_chroma) && pic->sao) {
        Frame<S> deb = cur;  // the deblocked copy SAO reads (src/hevc/sao.rs:1-4)
        sao_picture(cur, deb, info, pic, threads);
    }
    for (int c = 0; c < 3; ++c) cur.p[c].store((S*)out[c]);
    return RV_OK;
}

}  // namespace

extern "C" {

void orc_hevc_idct(int16_t* coeffs, int log2, int bd_shift, int max_x, int max_y) {
    idct(coeffs, 1 << log2, bd_shift, max_x, max_y);
}
void orc_hevc_idst4(int16_t* coeffs, int bd_shift) { idst4(coeffs, bd_shift); }
void orc_hevc_intra_tables(int32_t* angle33, int32_t* inv15) {
    for (int i = 0; i < 33; ++i) angle33[i] = INTRA_ANGLE[i];
    for (int i = 0; i < 15; ++i) inv15[i] = INV_ANGLE[i];
}
void orc_hevc_intra_substitute(uint16_t* scan, const uint8_t* avail, int n, int bit_depth) {
    IntraRefs R;
    R.n = n;
    bool a[129] = {};
    for (int j = 0; j < 4 * n + 1; ++j) {
        R.r[j] = scan[j];
        a[j] = avail[j] != 0;
    }
    intra_substitute(R, a, bit_depth);
    for (int j = 0; j < 4 * n + 1; ++j) scan[j] = R.r[j];
}
void orc_hevc_intra_pred(int32_t* dst, const uint16_t* scan, int n, int mode, int luma, int strong, int bit_depth) {
    IntraRefs R;
    R.n = n;
    for (int j = 0; j < 4 * n + 1; ++j) R.r[j] = scan[j];
    intra_pred(dst, R, mode, luma != 0, strong != 0, bit_depth);
}
int orc_hevc_intra_avail(const rv_hevc_pic* pic, int tu_index, uint8_t* avail) {
    PicInfo info;
    const int rc = build_info(pic, info);
    if (rc) return rc;
    if (tu_index < 0 || tu_index >= pic->n_tus) return RV_E_RANGE;
    const rv_hevc_tu& tu = pic->tus[tu_index];
    const int n = 1 << tu.log2;
    bool a[129];
    intra_avail(pic, info.pred_mode, tu.comp, tu.x, tu.y, n, a);
    for (int j = 0; j < 4 * n + 1; ++j) avail[j] = a[j] ? 1 : 0;
    return RV_OK;
}
void orc_hevc_scale_coefficients(int16_t* coeffs, int log2, int qp, int bit_depth, const uint8_t* list, int dc,
                                 int max_x, int max_y) {
    scale_coefficients(coeffs, log2, qp, bit_depth, list, dc, max_x, max_y);
}

void orc_hevc_idct_matrix(const int16_t* c, int n, int bd_shift, int16_t* out) {
    int step = 32 / n;
    std::vector<int> tmp(n * n);
    for (int x = 0; x < n; ++x)
        for (int i = 0; i < n; ++i) {
            long long s = 0;
            for (int j = 0; j < n; ++j) s += (long long)T32.m[j * step][i] * c[j * n + x];
            long long v = (s + 64) >> 7;
            tmp[i * n + x] = (int)std::min(std::max(v, -32768LL), 32767LL);
        }
    long long round = 1LL << (bd_shift - 1);
    for (int y = 0; y < n; ++y)
        for (int i = 0; i < n; ++i) {
            long long s = 0;
            for (int j = 0; j < n; ++j) s += (long long)T32.m[j * step][i] * tmp[y * n + j];
            long long v = (s + round) >> bd_shift;
            out[y * n + i] = (int16_t)std::min(std::max(v, -32768LL), 32767LL);
        }
}

#define ORC_DISPATCH(ss, call8, call16) \
    do {                                \
        if ((ss) == 1) {                \
            typedef uint8_t S;          \
            call8;                      \
        } else {                        \
            typedef uint16_t S;         \
            call16;                     \
        }                               \
    } while (0)
#define ORC_BOTH(ss, call) ORC_DISPATCH(ss, call, call)

void orc_hevc_add_residual(int ss, void* dst, int stride, const int16_t* res, int n, int max) {
    ORC_BOTH(ss, add_residual((S*)dst, stride, res, n, max));
}
void orc_hevc_copy(int ss, int16_t* dst, const void* src, int src_stride, int w, int h, int shift) {
    ORC_BOTH(ss, copy_k(dst, (const S*)src, src_stride, w, h, shift));
}
void orc_hevc_qpel_h(int ss, int16_t* dst, const void* src, int src_stride, int w, int h, int frac, int shift) {
    ORC_BOTH(ss, filt_h(QPEL[frac], 8, dst, (const S*)src, src_stride, w, h, shift));
}
void orc_hevc_qpel_v(int ss, int16_t* dst, const void* src, int src_stride, int w, int h, int frac, int shift) {
    ORC_BOTH(ss, filt_v(QPEL[frac], 8, dst, (const S*)src, src_stride, w, h, shift));
}
void orc_hevc_qpel_v2(int16_t* dst, const int16_t* src, int src_stride, int w, int h, int frac) {
    filt_v2(QPEL[frac], 8, dst, src, src_stride, w, h);
}
void orc_hevc_epel_h(int ss, int16_t* dst, const void* src, int src_stride, int w, int h, int frac, int shift) {
    ORC_BOTH(ss, filt_h(EPEL[frac], 4, dst, (const S*)src, src_stride, w, h, shift));
}
void orc_hevc_epel_v(int ss, int16_t* dst, const void* src, int src_stride, int w, int h, int frac, int shift) {
    ORC_BOTH(ss, filt_v(EPEL[frac], 4, dst, (const S*)src, src_stride, w, h, shift));
}
void orc_hevc_epel_v2(int16_t* dst, const int16_t* src, int src_stride, int w, int h, int frac) {
    filt_v2(EPEL[frac], 4, dst, src, src_stride, w, h);
}
void orc_hevc_uni(int ss, void* dst, int stride, const int16_t* src, int w, int h, int shift, int max) {
    ORC_BOTH(ss, uni_k((S*)dst, stride, src, w, h, shift, max));
}
void orc_hevc_bi(int ss, void* dst, int stride, const int16_t* a, const int16_t* b, int w, int h, int shift,
                 int max) {
    ORC_BOTH(ss, bi_k((S*)dst, stride, a, b, w, h, shift, max));
}
void orc_hevc_weighted_uni(int ss, void* dst, int stride, const int16_t* src, int w, int h, int log2_wd, int wt,
                           int o, int max) {
    ORC_BOTH(ss, wuni_k((S*)dst, stride, src, w, h, log2_wd, wt, o, max));
}
void orc_hevc_weighted_bi(int ss, void* dst, int stride, const int16_t* a, const int16_t* b, int w, int h,
                          int log2_wd, int w0, int w1, int o0, int o1, int max) {
    ORC_BOTH(ss, wbi_k((S*)dst, stride, a, b, w, h, log2_wd, w0, w1, o0, o1, max));
}
// qpel_uni / epel_uni / qpel_bi / epel_bi (src/dsp/hevc.rs:887-991).
void orc_hevc_mc_fused(int ss, int luma, void* dst, int dst_stride, const void* src, int src_stride, int w, int h,
                       int fx, int fy, const int16_t* other, int bit_depth) {
    std::vector<int16_t> tmp(64 * 71), pred(64 * 64);
    int max = (1 << bit_depth) - 1;
    ORC_BOTH(ss, {
        interp_k(luma != 0, pred.data(), tmp.data(), (const S*)src, src_stride, w, h, fx, fy, bit_depth);
        if (other) bi_k((S*)dst, dst_stride, other, pred.data(), w, h, 15 - bit_depth, max);
        else uni_k((S*)dst, dst_stride, pred.data(), w, h, 14 - bit_depth, max);
    });
}
void orc_hevc_sao_band(int ss, void* dst, int dst_stride, const void* src, int src_stride, int w, int h,
                       const int16_t* table32, int shift, int max) {
    ORC_BOTH(ss, sao_band_k((S*)dst, dst_stride, (const S*)src, src_stride, w, h, table32, shift, max));
}
void orc_hevc_sao_edge(int ss, void* dst, const void* src, int origin, int stride, int w, int h, int na, int nb,
                       const int16_t* off5, int max) {
    ORC_BOTH(ss, sao_edge_k((S*)dst, (const S*)src, origin, stride, w, h, na, nb, off5, max));
}
void orc_hevc_deblock_luma(int ss, int vertical, void* data, int off, int stride, const int32_t beta[2],
                           const int32_t tc[2], const uint8_t no_p[2], const uint8_t no_q[2], int max) {
    int b[2] = {beta[0], beta[1]}, t[2] = {tc[0], tc[1]};
    bool np[2] = {no_p[0] != 0, no_p[1] != 0}, nq[2] = {no_q[0] != 0, no_q[1] != 0};
    long step = vertical ? 1 : stride, along = vertical ? stride : 1;
    ORC_BOTH(ss, deblock_luma((S*)data, off, step, along, b, t, np, nq, max));
}
void orc_hevc_deblock_chroma(int ss, int vertical, void* data, int off, int stride, const int32_t tc[4],
                             const uint8_t no_p[4], const uint8_t no_q[4], int max) {
    int t[4] = {tc[0], tc[1], tc[2], tc[3]};
    bool np[4], nq[4];
    for (int i = 0; i < 4; ++i) {
        np[i] = no_p[i] != 0;
        nq[i] = no_q[i] != 0;
    }
    long step = vertical ? 1 : stride, along = vertical ? stride : 1;
    ORC_BOTH(ss, deblock_chroma((S*)data, off, step, along, t, np, nq, max));
}

int orc_hevc_bs_map(const rv_hevc_pic* pic, uint8_t* ver, uint8_t* hor) {
    if (!pic || !ver || !hor) return RV_E_INVALID;
    if (pic->width <= 0 || pic->height <= 0 || (pic->width & 7) || (pic->height & 7)) return RV_E_RANGE;
    PicInfo info;
    int rc = build_info(pic, info);
    if (rc) return rc;
    std::vector<uint8_t> v, h;
    boundary_strengths(info, v, h);
    std::copy(v.begin(), v.end(), ver);
    std::copy(h.begin(), h.end(), hor);
    return RV_OK;
}

int orc_hevc_reconstruct_mt(const rv_hevc_pic* pic, const void* const* slot_planes, int n_slots,
                            void* const out[3], int stop_after, int threads) {
    if (!pic || !slot_planes || !out || threads < 1) return RV_E_INVALID;
    if (pic->bit_depth == 8) return reconstruct<uint8_t>(pic, slot_planes, n_slots, out, stop_after, threads);
    if (pic->bit_depth > 8 && pic->bit_depth <= 12)
        return reconstruct<uint16_t>(pic, slot_planes, n_slots, out, stop_after, threads);
    return RV_E_RANGE;
}

int orc_hevc_reconstruct(const rv_hevc_pic* pic, const void* const* slot_planes, int n_slots, void* const out[3],
                         int stop_after) {
    return orc_hevc_reconstruct_mt(pic, slot_planes, n_slots, out, stop_after, 1);
}

}  // extern "C"
