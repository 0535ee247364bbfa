// TEST INFRASTRUCTURE ONLY — the CPU oracle for the H.264 reconstruction
// path (8-bit, 4:2:0, frame pictures). A C++ restatement of the reference's
// scalar kernel table (src/dsp/h264.rs) and of its drivers for one inter
// picture (src/h264/inter.rs, the weighting / residual lines of
// src/h264/recon.rs, the non-MBAFF part of src/h264/deblock.rs). Only
// tests, smoke() and the CPU baseline of bench.py load it.
//
// Pinned by (tests/test_oracle_h264.py): full-sample position == copy,
// half-sample positions reproduce linear ramps and constants, the
// quarter positions are the rounded means of their neighbours, the
// 2-D centre j equals a direct 6x6 sum, the 4x4 / 8x8 inverse transforms
// against the standard's matrix form (8.5.12.2 / 8.5.13.2) and the DC
// closed form (a DC of 64 adds 1 everywhere, src/h264/transform.rs:263-276),
// implicit weights against their closed form, deblocking identity at
// bS 0 / alpha 0, flat planes, and hand-computed step edges.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

const int PRED_STRIDE = 16;  // src/dsp/h264.rs:22-24

// Tables 8-15, 8-16, 8-17 of H.264 (src/h264/tables_gen.rs:331-392,
// src/h264/tables.rs:83).
const int ALPHA[52] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,   0,   0,   0,   0,   0,   0,   4,   4,
                       5,  6,  7,  8,  9,  10, 12, 13, 15, 17,  20,  22,  25,  28,  32,  36,  40,  45,
                       50, 56, 63, 71, 80, 90, 101, 113, 127, 144, 162, 182, 203, 226, 255, 255};
const int BETA[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  0,  0,  0,  0,  2,  2,  2,  3,  3,  3,  3,  4,  4,  4,
                      6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 18, 18};
const int TC0[52][3] = {
    {0, 0, 0},  {0, 0, 0},  {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},    {0, 0, 0},
    {0, 0, 0},  {0, 0, 0},  {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},   {0, 0, 0},    {0, 0, 0},
    {0, 0, 0},  {0, 0, 1},  {0, 0, 1},   {0, 0, 1},   {0, 0, 1},   {0, 1, 1},   {0, 1, 1},    {1, 1, 1},
    {1, 1, 1},  {1, 1, 1},  {1, 1, 1},   {1, 1, 2},   {1, 1, 2},   {1, 1, 2},   {1, 1, 2},    {1, 2, 3},
    {1, 2, 3},  {2, 2, 3},  {2, 2, 4},   {2, 3, 4},   {2, 3, 4},   {3, 3, 5},   {3, 4, 6},    {3, 4, 6},
    {4, 5, 7},  {4, 5, 8},  {4, 6, 9},   {5, 7, 10},  {6, 8, 11},  {6, 8, 13},  {7, 10, 14},  {8, 11, 16},
    {9, 12, 18}, {10, 13, 20}, {11, 15, 23}, {13, 17, 25}};

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------------
// Luma interpolation (src/dsp/h264.rs:395-493)
// ---------------------------------------------------------------------
inline int tap6(int a, int b, int c, int d, int e, int f) { return a - 5 * b + 20 * c + 20 * d - 5 * e + f; }

struct QpelSrc {
    const uint8_t* s;
    long st;
    int max;
    int G(int x, int y) const { return s[(y + 2) * st + x + 2]; }
    int b1(int x, int yy) const {
        const uint8_t* r = s + yy * st + x;
        return tap6(r[0], r[1], r[2], r[3], r[4], r[5]);
    }
    int h1(int xx, int y) const {
        const uint8_t* c = s + y * st + xx;
        return tap6(c[0], c[st], c[2 * st], c[3 * st], c[4 * st], c[5 * st]);
    }
    int b(int x, int y) const { return clampi((b1(x, y + 2) + 16) >> 5, 0, max); }
    int hh(int x, int y) const { return clampi((h1(x + 2, y) + 16) >> 5, 0, max); }
    int j(int x, int y) const {
        int j1 = tap6(b1(x, y), b1(x, y + 1), b1(x, y + 2), b1(x, y + 3), b1(x, y + 4), b1(x, y + 5));
        return clampi((j1 + 512) >> 10, 0, max);
    }
};

void qpel(int pos, uint8_t* dst, const uint8_t* src, long stride, int w, int h, int max) {
    QpelSrc q{src, stride, max};
    int XF = pos & 3, YF = pos >> 2;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int v;
            switch (YF * 4 + XF) {
                case 0: v = q.G(x, y); break;
                case 1: v = (q.G(x, y) + q.b(x, y) + 1) >> 1; break;
                case 2: v = q.b(x, y); break;
                case 3: v = (q.G(x + 1, y) + q.b(x, y) + 1) >> 1; break;
                case 4: v = (q.G(x, y) + q.hh(x, y) + 1) >> 1; break;
                case 8: v = q.hh(x, y); break;
                case 12: v = (q.G(x, y + 1) + q.hh(x, y) + 1) >> 1; break;
                case 10: v = q.j(x, y); break;
                case 5: v = (q.b(x, y) + q.hh(x, y) + 1) >> 1; break;
                case 7: v = (q.b(x, y) + q.hh(x + 1, y) + 1) >> 1; break;
                case 13: v = (q.hh(x, y) + q.b(x, y + 1) + 1) >> 1; break;
                case 15: v = (q.hh(x + 1, y) + q.b(x, y + 1) + 1) >> 1; break;
                case 6: v = (q.b(x, y) + q.j(x, y) + 1) >> 1; break;
                case 14: v = (q.j(x, y) + q.b(x, y + 1) + 1) >> 1; break;
                case 9: v = (q.hh(x, y) + q.j(x, y) + 1) >> 1; break;
                default: v = (q.j(x, y) + q.hh(x + 1, y) + 1) >> 1; break;  // 11: (3, 2)
            }
            dst[y * PRED_STRIDE + x] = (uint8_t)v;
        }
}

// chroma_scalar (src/dsp/h264.rs:499-520).
void chroma(uint8_t* dst, const uint8_t* src, long stride, int w, int h, int xf, int yf) {
    int a = (8 - xf) * (8 - yf), b = xf * (8 - yf), c = (8 - xf) * yf, d = xf * yf;
    for (int y = 0; y < h; ++y) {
        const uint8_t* r0 = src + y * stride;
        const uint8_t* r1 = src + (y + 1) * stride;
        for (int x = 0; x < w; ++x) {
            int v = a * r0[x] + b * r0[x + 1] + c * r1[x] + d * r1[x + 1];
            dst[y * PRED_STRIDE + x] = (uint8_t)((v + 32) >> 6);
        }
    }
}

// copy / avg / weighted_uni / weighted_bi (src/dsp/h264.rs:522-597).
void copy_k(uint8_t* dst, long stride, const uint8_t* src, int w, int h) {
    for (int y = 0; y < h; ++y) std::memcpy(dst + y * stride, src + y * PRED_STRIDE, w);
}
void avg_k(uint8_t* dst, long stride, const uint8_t* a, const uint8_t* b, int w, int h) {
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            dst[y * stride + x] = (uint8_t)((a[y * PRED_STRIDE + x] + b[y * PRED_STRIDE + x] + 1) >> 1);
}
void wuni_k(uint8_t* dst, long stride, const uint8_t* src, int w, int h, int log_wd, int wt, int o, int max) {
    if (log_wd >= 1) {
        int round = 1 << (log_wd - 1);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                dst[y * stride + x] = (uint8_t)clampi(((src[y * PRED_STRIDE + x] * wt + round) >> log_wd) + o, 0, max);
    } else {
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                dst[y * stride + x] = (uint8_t)clampi(src[y * PRED_STRIDE + x] * wt + o, 0, max);
    }
}
void wbi_k(uint8_t* dst, long stride, const uint8_t* a, const uint8_t* b, int w, int h, int log_wd, int w0, int w1,
           int o0, int o1, int max) {
    int off = (o0 + o1 + 1) >> 1, round = 1 << log_wd;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int v = ((a[y * PRED_STRIDE + x] * w0 + b[y * PRED_STRIDE + x] * w1 + round) >> (log_wd + 1)) + off;
            dst[y * stride + x] = (uint8_t)clampi(v, 0, max);
        }
}

// ---------------------------------------------------------------------
// Deblocking kernels (src/dsp/h264.rs:603-763)
// ---------------------------------------------------------------------

// deblock_line (:603-656). tc0 < 0 on entry means "bS 4".
void deblock_line(int* p, int* q, bool strong_bs, int tc0, int alpha, int beta, bool chroma, int max) {
    int p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
    int q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    if (!(std::abs(p0 - q0) < alpha && std::abs(p1 - p0) < beta && std::abs(q1 - q0) < beta)) return;
    int ap = std::abs(p2 - p0), aq = std::abs(q2 - q0);
    if (!strong_bs) {
        int tc = chroma ? tc0 + 1 : tc0 + (ap < beta) + (aq < beta);
        int delta = clampi((((q0 - p0) << 2) + (p1 - q1) + 4) >> 3, -tc, tc);
        p[0] = clampi(p0 + delta, 0, max);
        q[0] = clampi(q0 - delta, 0, max);
        if (!chroma) {
            if (ap < beta) p[1] = p1 + clampi((p2 + ((p0 + q0 + 1) >> 1) - (p1 << 1)) >> 1, -tc0, tc0);
            if (aq < beta) q[1] = q1 + clampi((q2 + ((p0 + q0 + 1) >> 1) - (q1 << 1)) >> 1, -tc0, tc0);
        }
    } else {
        bool strong = std::abs(p0 - q0) < ((alpha >> 2) + 2);
        if (!chroma && ap < beta && strong) {
            p[0] = (p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3;
            p[1] = (p2 + p1 + p0 + q0 + 2) >> 2;
            p[2] = (2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3;
        } else {
            p[0] = (2 * p1 + p0 + q1 + 2) >> 2;
        }
        if (!chroma && aq < beta && strong) {
            q[0] = (p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3;
            q[1] = (p0 + q0 + q1 + q2 + 2) >> 2;
            q[2] = (2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3;
        } else {
            q[0] = (2 * q1 + q0 + p1 + 2) >> 2;
        }
    }
}

// deblock_luma_lines<16> (:658-700) / deblock_chroma_scalar (:731-763).
// step moves along the edge, across across it.
void deblock_lines(bool chroma_k, uint8_t* d, long off, long step, long across, int alpha, int beta,
                   const int16_t* tc0, int max) {
    int lines = chroma_k ? 8 : 16, per = chroma_k ? 2 : 4;
    for (int i = 0; i < lines; ++i) {
        int t = 0;
        if (tc0) {
            if (tc0[i / per] < 0) continue;
            t = tc0[i / per];
        }
        long base = off + i * step;
        int p[4] = {0, 0, 0, 0}, q[4] = {0, 0, 0, 0};
        int n = chroma_k ? 2 : 4;
        for (int k = 0; k < n; ++k) {
            p[k] = d[base - (k + 1) * across];
            q[k] = d[base + k * across];
        }
        deblock_line(p, q, tc0 == nullptr, t, alpha, beta, chroma_k, max);
        int wn = chroma_k ? 1 : 3;
        for (int k = 0; k < wn; ++k) {
            d[base - (k + 1) * across] = (uint8_t)p[k];
            d[base + k * across] = (uint8_t)q[k];
        }
    }
}

// ---------------------------------------------------------------------
// Inverse transforms, 8-bit: wrapping i16 (src/dsp/h264.rs:766-1066)
// ---------------------------------------------------------------------
inline int16_t wa(int16_t a, int16_t b) { return (int16_t)(uint16_t)((uint16_t)a + (uint16_t)b); }
inline int16_t ws(int16_t a, int16_t b) { return (int16_t)(uint16_t)((uint16_t)a - (uint16_t)b); }

// add_rows_scalar (:803-816).
void add_rows(uint8_t* dst, long stride, const int16_t* v, int n, int max) {
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            int16_t r = (int16_t)(wa(v[y * n + x], 32) >> 6);
            uint8_t* p = dst + y * stride + x;
            *p = (uint8_t)clampi((int)wa((int16_t)*p, r), 0, max);
        }
}

// idct4_add_scalar (:774-801).
void idct4_add(uint8_t* dst, long stride, const int16_t* coeffs, int max) {
    int16_t d[16], tmp[16];
    std::memcpy(d, coeffs, sizeof(d));
    for (int i = 0; i < 4; ++i) {
        int16_t d0 = d[i * 4], d1 = d[i * 4 + 1], d2 = d[i * 4 + 2], d3 = d[i * 4 + 3];
        int16_t e0 = wa(d0, d2), e1 = ws(d0, d2), e2 = ws((int16_t)(d1 >> 1), d3), e3 = wa(d1, (int16_t)(d3 >> 1));
        tmp[i * 4] = wa(e0, e3);
        tmp[i * 4 + 1] = wa(e1, e2);
        tmp[i * 4 + 2] = ws(e1, e2);
        tmp[i * 4 + 3] = ws(e0, e3);
    }
    for (int j = 0; j < 4; ++j) {
        int16_t f0 = tmp[j], f1 = tmp[4 + j], f2 = tmp[8 + j], f3 = tmp[12 + j];
        int16_t g0 = wa(f0, f2), g1 = ws(f0, f2), g2 = ws((int16_t)(f1 >> 1), f3), g3 = wa(f1, (int16_t)(f3 >> 1));
        d[j] = wa(g0, g3);
        d[4 + j] = wa(g1, g2);
        d[8 + j] = ws(g1, g2);
        d[12 + j] = ws(g0, g3);
    }
    add_rows(dst, stride, d, 4, max);
}

// idct8_1d_i16 (:913-958).
void idct8_1d(const int16_t* d, int16_t* o) {
    int16_t a0 = wa(d[0], d[4]), a4 = ws(d[0], d[4]);
    int16_t a2 = ws((int16_t)(d[2] >> 1), d[6]), a6 = wa(d[2], (int16_t)(d[6] >> 1));
    int16_t b0 = wa(a0, a6), b2 = wa(a4, a2), b4 = ws(a4, a2), b6 = ws(a0, a6);
    int16_t a1 = ws(ws(ws(d[5], d[3]), d[7]), (int16_t)(d[7] >> 1));
    int16_t a3 = ws(ws(wa(d[1], d[7]), d[3]), (int16_t)(d[3] >> 1));
    int16_t a5 = wa(wa(ws(d[7], d[1]), d[5]), (int16_t)(d[5] >> 1));
    int16_t a7 = wa(wa(wa(d[3], d[5]), d[1]), (int16_t)(d[1] >> 1));
    int16_t b1 = wa(a1, (int16_t)(a7 >> 2)), b7 = ws(a7, (int16_t)(a1 >> 2));
    int16_t b3 = wa(a3, (int16_t)(a5 >> 2)), b5 = ws((int16_t)(a3 >> 2), a5);
    o[0] = wa(b0, b7);
    o[1] = wa(b2, b5);
    o[2] = wa(b4, b3);
    o[3] = wa(b6, b1);
    o[4] = ws(b6, b1);
    o[5] = ws(b4, b3);
    o[6] = ws(b2, b5);
    o[7] = ws(b0, b7);
}

// idct8_add_scalar (:960-982).
void idct8_add(uint8_t* dst, long stride, const int16_t* coeffs, int max) {
    int16_t tmp[64], out[64];
    for (int i = 0; i < 8; ++i) idct8_1d(coeffs + i * 8, tmp + i * 8);
    for (int j = 0; j < 8; ++j) {
        int16_t col[8], o[8];
        for (int k = 0; k < 8; ++k) col[k] = tmp[k * 8 + j];
        idct8_1d(col, o);
        for (int i = 0; i < 8; ++i) out[i * 8 + j] = o[i];
    }
    add_rows(dst, stride, out, 8, max);
}

// dc_add_scalar (:984-996), 8-bit.
void dc_add(uint8_t* dst, long stride, int dc, int n, int max) {
    int v = (int)(int16_t)(wa((int16_t)dc, 32) >> 6);
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            uint8_t* p = dst + y * stride + x;
            *p = (uint8_t)clampi(*p + v, 0, max);
        }
}

// residual4_scalar / residual8_scalar with narrow4 / narrow8 (:998-1066).
void residual4(uint8_t* dst, long stride, const int32_t* coefs, int dc, bool has_dc, int max) {
    int16_t c[16];
    int ac = 0;
    for (int i = 0; i < 16; ++i) {
        c[i] = (int16_t)clampi(coefs[i], -32768, 32767);
        if (i != 0) ac |= coefs[i];
    }
    if (has_dc) c[0] = (int16_t)dc;
    if (ac != 0) idct4_add(dst, stride, c, max);
    else if (c[0] != 0) dc_add(dst, stride, c[0], 4, max);
}
void residual8(uint8_t* dst, long stride, const int32_t* coefs, int max) {
    int16_t c[64];
    int ac = 0;
    for (int i = 0; i < 64; ++i) {
        c[i] = (int16_t)clampi(coefs[i], -32768, 32767);
        if (i != 0) ac |= coefs[i];
    }
    if (ac != 0) idct8_add(dst, stride, c, max);
    else if (c[0] != 0) dc_add(dst, stride, c[0], 8, max);
}

// ---------------------------------------------------------------------
// Scaling (8.5.9, 8.5.10, 8.5.11, 8.5.12.1, 8.5.13.1)
// ---------------------------------------------------------------------

// normAdjust4x4(m, i, j) of 8.5.9: v[m][0] where i and j are both even,
// v[m][1] where both are odd, v[m][2] otherwise (Dequant::new,
// src/h264/transform.rs:33-41, over src/h264/tables.rs's DEQUANT4_INIT,
// which orders the same values by the number of odd indices).
const int NORM4[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16}, {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
// normAdjust8x8(m, i, j): v[m][0] i%4 == 0 && j%4 == 0; v[m][1] i, j odd;
// v[m][2] i%4 == 2 && j%4 == 2; v[m][3] one index a multiple of 4, the other
// odd; v[m][4] one a multiple of 4, the other 2 mod 4; v[m][5] otherwise
// (transform.rs:42-48, via DEQUANT8_INIT_SCAN).
const int NORM8[6][6] = {{20, 18, 32, 19, 25, 24}, {22, 19, 35, 21, 28, 26}, {26, 23, 42, 24, 33, 31},
                         {28, 25, 45, 26, 35, 33}, {32, 28, 51, 30, 40, 38}, {36, 32, 58, 34, 46, 43}};

int norm4(int m, int i, int j) {
    if (i % 2 == 0 && j % 2 == 0) return NORM4[m][0];
    if (i % 2 == 1 && j % 2 == 1) return NORM4[m][1];
    return NORM4[m][2];
}
int norm8(int m, int i, int j) {
    if (i % 4 == 0 && j % 4 == 0) return NORM8[m][0];
    if (i % 2 == 1 && j % 2 == 1) return NORM8[m][1];
    if (i % 4 == 2 && j % 4 == 2) return NORM8[m][2];
    if ((i % 4 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 4 == 0)) return NORM8[m][3];
    if ((i % 4 == 0 && j % 4 == 2) || (i % 4 == 2 && j % 4 == 0)) return NORM8[m][4];
    return NORM8[m][5];
}

// LevelScale4x4(m, i, j) = weightScale4x4(i, j) * normAdjust4x4(m, i, j) at
// raster pos = i * 4 + j (Dequant::new, transform.rs:24-56); flat lists are
// 16 everywhere.
int level_scale4(const rv_h264_scaling* sl, int list, int m, int pos) {
    return (sl ? sl->list4[list][pos] : 16) * norm4(m, pos / 4, pos % 4);
}
int level_scale8(const rv_h264_scaling* sl, int list, int m, int pos) {
    return (sl ? sl->list8[list][pos] : 16) * norm8(m, pos / 8, pos % 8);
}

// dequant_level (src/h264/mb.rs:1420-1428): (level * table << shift + 32)
// >> 6 in wrapping i32, shift = qP / 6 + 2 for 4x4 blocks and qP / 6 for
// 8x8 (MbDequant::for_mb, :1404-1413) — 8.5.12.1 / 8.5.13.1 folded.
int32_t dequant_level(int32_t level, int32_t table, int shift) {
    const uint32_t t = ((uint32_t)level * (uint32_t)table) << shift;
    return (int32_t)(t + 32u) >> 6;
}

// luma_dc_transform (transform.rs:179-213): the Intra_16x16 DC levels
// (raster 4x4) through the inverse Hadamard, rows then columns, then
// scaled at qP with LevelScale4x4(qP % 6, 0, 0).
void luma_dc_transform(int32_t* dc, int scale00, int qp) {
    int32_t t[16], f[16];
    for (int i = 0; i < 4; ++i) {
        const int32_t c0 = dc[i * 4], c1 = dc[i * 4 + 1], c2 = dc[i * 4 + 2], c3 = dc[i * 4 + 3];
        const int32_t e0 = c0 + c1, e1 = c0 - c1, e2 = c2 - c3, e3 = c2 + c3;
        t[i * 4] = e0 + e3;
        t[i * 4 + 1] = e0 - e3;
        t[i * 4 + 2] = e1 - e2;
        t[i * 4 + 3] = e1 + e2;
    }
    for (int j = 0; j < 4; ++j) {
        const int32_t c0 = t[j], c1 = t[4 + j], c2 = t[8 + j], c3 = t[12 + j];
        const int32_t e0 = c0 + c1, e1 = c0 - c1, e2 = c2 - c3, e3 = c2 + c3;
        f[j] = e0 + e3;
        f[4 + j] = e0 - e3;
        f[8 + j] = e1 - e2;
        f[12 + j] = e1 + e2;
    }
    const int q6 = qp / 6;
    for (int i = 0; i < 16; ++i) {
        const uint32_t p = (uint32_t)f[i] * (uint32_t)scale00;
        dc[i] = qp >= 36 ? (int32_t)(p << (q6 - 6)) : (int32_t)(p + (1u << (5 - q6))) >> (6 - q6);
    }
}

// chroma_dc_transform_420 (transform.rs:216-228): the 2x2 Hadamard, then
// ((f * LevelScale4x4(qP % 6, 0, 0)) << (qP / 6)) >> 5.
void chroma_dc_transform_420(int32_t* dc, int scale00, int qp) {
    const int32_t c0 = dc[0], c1 = dc[1], c2 = dc[2], c3 = dc[3];
    const int32_t f[4] = {c0 + c1 + c2 + c3, c0 - c1 + c2 - c3, c0 + c1 - c2 - c3, c0 - c1 - c2 + c3};
    for (int i = 0; i < 4; ++i) dc[i] = (int32_t)(((uint32_t)f[i] * (uint32_t)scale00) << (qp / 6)) >> 5;
}

// The coefficients and DC of a block of levels as the reference's parser
// and recon hand them to residual4 / residual8 (src/h264/cavlc.rs:300-310,
// src/h264/recon.rs:700-716, 1200-1226): AC levels through dequant_level
// with the macroblock's tables; with has_dc, the group's DC levels through
// the DC transform and this block's entry taken. Returns false for a QP
// outside 0..51 or a DC group outside coefs.
bool scale_block(const rv_h264_pic* pic, const rv_h264_blk& bk, int32_t* out, int* dc_out) {
    const int mbw = pic->width / 16;
    const bool chroma = bk.comp != 0;
    const int mbx = chroma ? bk.x / 8 : bk.x / 16, mby = chroma ? bk.y / 8 : bk.y / 16;
    const rv_h264_mb& m = pic->mbs[mby * mbw + mbx];
    const int qp = chroma ? m.qpc[bk.comp - 1] : m.qp;
    if (qp < 0 || qp > 51) return false;
    const bool inter = !m.intra;
    const int n = bk.size * bk.size;
    const int32_t* lv = pic->coefs + bk.coef_off;
    for (int i = 0; i < n; ++i)
        out[i] = bk.size == 8 ? dequant_level(lv[i], level_scale8(pic->scaling, inter ? 1 : 0, qp % 6, i), qp / 6)
                              : dequant_level(lv[i], level_scale4(pic->scaling, (inter ? 3 : 0) + bk.comp, qp % 6, i),
                                              qp / 6 + 2);
    *dc_out = 0;
    if (!bk.has_dc) return true;
    const int group = chroma ? 4 : 16;
    if (bk.dc < 0 || (uint64_t)bk.dc + group > pic->n_coefs) return false;
    int32_t g[16];
    std::memcpy(g, pic->coefs + bk.dc, group * sizeof(int32_t));
    if (chroma) {
        chroma_dc_transform_420(g, level_scale4(pic->scaling, (inter ? 3 : 0) + bk.comp, qp % 6, 0), qp);
        *dc_out = g[((bk.y % 8) / 4) * 2 + (bk.x % 8) / 4];
    } else {
        luma_dc_transform(g, level_scale4(pic->scaling, 0, qp % 6, 0), qp);
        *dc_out = g[((bk.y % 16) / 4) * 4 + (bk.x % 16) / 4];
    }
    return true;
}

// implicit_pair (src/h264/recon.rs:214-233).
void implicit_pair(int cur_poc, int poc0, int poc1, bool long0, bool long1, int* w0, int* w1) {
    int tb = clampi(cur_poc - poc0, -128, 127);
    int td = clampi(poc1 - poc0, -128, 127);
    if (td == 0 || long0 || long1) {
        *w0 = 32;
        *w1 = 32;
        return;
    }
    int tx = (16384 + std::abs(td / 2)) / td;
    int dsf = clampi((tb * tx + 32) >> 6, -1024, 1023);
    int ww1 = dsf >> 2;
    if (ww1 < -64 || ww1 > 128) {
        *w0 = 32;
        *w1 = 32;
        return;
    }
    *w0 = 64 - ww1;
    *w1 = ww1;
}

// ---------------------------------------------------------------------
// Frames (src/h264/frame.rs:13-15: luma border 32, chroma 16)
// ---------------------------------------------------------------------
struct Plane {
    std::vector<uint8_t> data;
    int width = 0, height = 0, pad = 0;
    long stride = 0;
    void init(int w, int h, int p) {
        width = w;
        height = h;
        pad = p;
        stride = w + 2 * p;
        data.assign((size_t)stride * (h + 2 * p), 0);
    }
    long offset(long x, long y) const { return (long)pad * stride + pad + y * stride + x; }
    void load(const uint8_t* s) {
        for (int y = 0; y < height; ++y) std::memcpy(&data[offset(0, y)], s + (long)y * width, width);
    }
    void store(uint8_t* d) const {
        for (int y = 0; y < height; ++y) std::memcpy(d + (long)y * width, &data[offset(0, y)], width);
    }
    void extend() {
        for (int y = 0; y < height; ++y) {
            long row = offset(0, y);
            for (int i = 1; i <= pad; ++i) {
                data[row - i] = data[row];
                data[row + width - 1 + i] = data[row + width - 1];
            }
        }
        long first = offset(-pad, 0), last = offset(-pad, height - 1);
        for (int i = 1; i <= pad; ++i) {
            std::memcpy(&data[first - i * stride], &data[first], stride);
            std::memcpy(&data[last + i * stride], &data[last], stride);
        }
    }
};
struct Frame {
    Plane p[3];
    void init(int w, int h) {
        p[0].init(w, h, 32);
        p[1].init(w / 2, h / 2, 16);
        p[2].init(w / 2, h / 2, 16);
    }
};

// interp (src/h264/inter.rs:107-140), frame parity: the window from the
// padded plane when it fits, else a clamped gather with stride 32.
template <typename K>
void interp(const Plane& src, int x0, int y0, int ww, int hh, std::vector<uint8_t>& window, K kernel) {
    int pad = src.pad;
    if (x0 >= -pad && y0 >= -pad && x0 + ww <= src.width + pad && y0 + hh <= src.height + pad) {
        kernel(&src.data[src.offset(x0, y0)], src.stride);
    } else {
        for (int y = 0; y < hh; ++y) {
            int yy = clampi(y0 + y, 0, src.height - 1);
            for (int x = 0; x < ww; ++x) {
                int xx = clampi(x0 + x, 0, src.width - 1);
                window[y * 32 + x] = src.data[src.offset(xx, yy)];
            }
        }
        kernel(window.data(), 32);
    }
}

struct BlockMotion {
    int mv[2];
    int ref_idx;
    int ref_id;  // the referenced picture (its DPB slot); -1 none
};

// motion_bs (src/h264/deblock.rs:97-146), frame macroblocks (mvy_limit 4).
int motion_bs(const BlockMotion* m0, const BlockMotion* m1, int pa, int qa) {
    const BlockMotion &p0 = m0[pa], &p1 = m1[pa], &q0 = m0[qa], &q1 = m1[qa];
    int pn = (p0.ref_idx >= 0) + (p1.ref_idx >= 0), qn = (q0.ref_idx >= 0) + (q1.ref_idx >= 0);
    if (pn != qn) return 1;
    auto far = [](const int* a, const int* b) { return std::abs(a[0] - b[0]) >= 4 || std::abs(a[1] - b[1]) >= 4; };
    auto same = [](const BlockMotion& a, const BlockMotion& b) { return a.ref_id == b.ref_id; };
    if (pn == 1) {
        const BlockMotion& pp = p0.ref_idx >= 0 ? p0 : p1;
        const BlockMotion& qq = q0.ref_idx >= 0 ? q0 : q1;
        if (!same(pp, qq)) return 1;
        return far(pp.mv, qq.mv);
    }
    bool straight = same(p0, q0) && same(p1, q1);
    bool crossed = same(p0, q1) && same(p1, q0);
    if (!straight && !crossed) return 1;
    if (same(p0, p1)) {
        bool pa_ = far(p0.mv, q0.mv) || far(p1.mv, q1.mv);
        bool pb_ = far(p0.mv, q1.mv) || far(p1.mv, q0.mv);
        return pa_ && pb_;
    }
    if (straight) return far(p0.mv, q0.mv) || far(p1.mv, q1.mv);
    return far(p0.mv, q1.mv) || far(p1.mv, q0.mv);
}

struct Thr {
    int alpha, beta;
    int16_t lut[4];
};

const uint32_t BS2_SPREAD[16] = {0x00000000, 0x00000002, 0x00000200, 0x00000202, 0x00020000, 0x00020002,
                                 0x00020200, 0x00020202, 0x02000000, 0x02000002, 0x02000200, 0x02000202,
                                 0x02020000, 0x02020002, 0x02020200, 0x02020202};
inline uint32_t gather_col(uint32_t x, int s) {
    x = (x >> s) & 0x1111;
    return (x | (x >> 3) | (x >> 6) | (x >> 9)) & 0xF;
}

// filter_edge (src/h264/deblock.rs:192-213): byte 0 decides bS 4.
void filter_edge(bool chroma_k, bool vertical, uint8_t* data, long off, long stride, uint32_t bs, const Thr& t,
                 int max) {
    long step = vertical ? stride : 1, across = vertical ? 1 : stride;
    if ((bs & 0xFF) == 4) {
        deblock_lines(chroma_k, data, off, step, across, t.alpha, t.beta, nullptr, max);
    } else {
        int16_t tc[4];
        for (int k = 0; k < 4; ++k) tc[k] = t.lut[(bs >> (8 * k)) & 3];
        deblock_lines(chroma_k, data, off, step, across, t.alpha, t.beta, tc, max);
    }
}

// Per-4x4 motion per list (BlockMotion, src/h264/frame.rs:42-74), raster
// per macroblock (index mb * 16 + by * 4 + bx), and the partition of each
// 4x4 block (raster over the picture), from the partition list.
struct MbSide {
    int mbw = 0, mbh = 0, w4 = 0;
    std::vector<BlockMotion> motion[2];
    std::vector<int> part_of;
};

int build_side(const rv_h264_pic* pic, MbSide& sd) {
    int W = pic->width, H = pic->height;
    sd.mbw = W / 16;
    sd.mbh = H / 16;
    sd.w4 = W / 4;
    BlockMotion none{{0, 0}, -1, -1};
    for (auto& m : sd.motion) m.assign((size_t)sd.mbw * sd.mbh * 16, none);
    sd.part_of.assign((size_t)sd.w4 * (H / 4), -1);
    for (int i = 0; i < pic->n_parts; ++i) {
        const rv_h264_part& pt = pic->parts[i];
        int x = pt.x, y = pt.y, w = pt.w, h = pt.h;
        if (!(pt.dir & 3)) return RV_E_RANGE;
        for (int l = 0; l < 2; ++l) {
            if (!((pt.dir >> l) & 1)) continue;
            int ri = pt.ref_idx[l];
            if (ri < 0 || ri >= RV_MAX_REFS || ri >= pic->num_ref[l]) return RV_E_RANGE;
            for (int by = y / 4; by < (y + h) / 4; ++by)
                for (int bx = x / 4; bx < (x + w) / 4; ++bx) {
                    int mb = (by / 4) * sd.mbw + bx / 4, blk = (by % 4) * 4 + (bx % 4);
                    BlockMotion& bm = sd.motion[l][(size_t)mb * 16 + blk];
                    bm.mv[0] = pt.mv[l].x;
                    bm.mv[1] = pt.mv[l].y;
                    bm.ref_idx = ri;
                    bm.ref_id = pic->ref_slot[l][ri];
                }
        }
        for (int by = y / 4; by < (y + h) / 4; ++by)
            for (int bx = x / 4; bx < (x + w) / 4; ++bx) sd.part_of[(size_t)by * sd.w4 + bx] = i;
    }
    return RV_OK;
}

// part_edges (MbInfo, src/h264/mb.rs:350-365) from the partitions: bit
// e * 4 + k is set where the 4x4 blocks either side of internal edge e
// (1..3) at line k belong to different partitions.
void part_edges(const MbSide& sd, int mbx, int mby, uint32_t* pe_v, uint32_t* pe_h) {
    *pe_v = *pe_h = 0;
    const int bx = mbx * 4, by = mby * 4, w4 = sd.w4;
    for (int e = 1; e < 4; ++e)
        for (int k = 0; k < 4; ++k) {
            if (sd.part_of[(size_t)(by + k) * w4 + bx + e - 1] != sd.part_of[(size_t)(by + k) * w4 + bx + e])
                *pe_v |= 1u << (e * 4 + k);
            if (sd.part_of[(size_t)(by + e - 1) * w4 + bx + k] != sd.part_of[(size_t)(by + e) * w4 + bx + k])
                *pe_h |= 1u << (e * 4 + k);
        }
}

// The packed strengths of macroblock `addr` (deblock_mb_rows' bS
// derivation, src/h264/deblock.rs:617-756): bs_v[e] / bs_h[e] for edge e
// (0 = the macroblock edge), one byte per 4-sample segment. Frame
// picture, one slice.
void mb_strengths(const rv_h264_pic* pic, const MbSide& sd, int addr, uint32_t bs_v[4], uint32_t bs_h[4]) {
    const int mbw = sd.mbw, mbx = addr % mbw, mby = addr / mbw;
    const rv_h264_mb& m = pic->mbs[addr];
    const int left = mbx > 0 ? addr - 1 : addr, above = mby > 0 ? addr - mbw : addr;
    const bool filter_left = mbx > 0, filter_top = mby > 0;
    for (int e = 0; e < 4; ++e) bs_v[e] = bs_h[e] = 0;
    const bool internal_odd = !m.transform_8x8;
    uint32_t pe_v, pe_h;
    part_edges(sd, mbx, mby, &pe_v, &pe_h);
    if (m.intra) {
        if (filter_left) bs_v[0] = 0x04040404;
        if (filter_top) bs_h[0] = 0x04040404;
        for (int e = 1; e < 4; ++e) bs_v[e] = bs_h[e] = 0x03030303;
        return;
    }
    const BlockMotion* m0 = sd.motion[0].data();
    const BlockMotion* m1 = sd.motion[1].data();
    auto edge_motion = [&](uint32_t rest, uint32_t chg, auto bs_of) -> uint32_t {
        if (rest == 0) return 0;
        uint32_t packed = 0, held = 0;
        bool known = false;
        for (int k = 0; k < 4; ++k) {
            if ((chg >> k) & 1) known = false;
            if (!((rest >> k) & 1)) continue;
            if (!known) {
                held = (uint32_t)bs_of(k);
                known = true;
            }
            packed |= held << (k * 8);
        }
        return packed;
    };
    uint32_t nz = m.nz_mask;
    int step = internal_odd ? 1 : 2;
    // internal_edge (src/h264/deblock.rs:255-282).
    auto internal = [&](uint32_t coef, uint32_t part, int e, bool vertical) -> uint32_t {
        uint32_t c = vertical ? gather_col(coef, e - 1) : (coef >> ((e - 1) * 4)) & 0xF;
        uint32_t bs = BS2_SPREAD[c];
        uint32_t rest = ((part >> (e * 4)) & 0xF) & ~c;
        for (int k = 0; k < 4; ++k) {
            if (!((rest >> k) & 1)) continue;
            int pb = vertical ? k * 4 + e - 1 : (e - 1) * 4 + k;
            int qb = vertical ? k * 4 + e : e * 4 + k;
            bs |= (uint32_t)motion_bs(m0, m1, addr * 16 + pb, addr * 16 + qb) << (k * 8);
        }
        return bs;
    };
    if (nz != 0 || pe_v != 0) {
        uint32_t coef_v = nz | (nz >> 1);
        for (int e = step; e < 4; e += step) bs_v[e] = internal(coef_v, pe_v, e, true);
    }
    if (nz != 0 || pe_h != 0) {
        uint32_t coef_h = nz | (nz >> 4);
        for (int e = step; e < 4; e += step) bs_h[e] = internal(coef_h, pe_h, e, false);
    }
    if (filter_left) {
        const rv_h264_mb& ml = pic->mbs[left];
        if (ml.intra) {
            bs_v[0] = 0x04040404;
        } else {
            uint32_t lpv, lph;
            part_edges(sd, mbx - 1, mby, &lpv, &lph);
            uint32_t c = gather_col(ml.nz_mask, 3) | gather_col(nz, 0);
            uint32_t rest = ~c & 0xF;
            uint32_t chg = gather_col(lph, 3) | gather_col(pe_h, 0);
            bs_v[0] = BS2_SPREAD[c] | edge_motion(rest, chg, [&](int k) {
                          return motion_bs(m0, m1, left * 16 + k * 4 + 3, addr * 16 + k * 4);
                      });
        }
    }
    if (filter_top) {
        const rv_h264_mb& ma = pic->mbs[above];
        if (ma.intra) {
            bs_h[0] = 0x04040404;
        } else {
            uint32_t apv, aph;
            part_edges(sd, mbx, mby - 1, &apv, &aph);
            uint32_t c = ((ma.nz_mask >> 12) | nz) & 0xF;
            uint32_t rest = ~c & 0xF;
            uint32_t chg = gather_col(apv, 3) | gather_col(pe_v, 0);
            bs_h[0] = BS2_SPREAD[c] | edge_motion(rest, chg, [&](int k) {
                          return motion_bs(m0, m1, above * 16 + 12 + k, addr * 16 + k);
                      });
        }
    }
}


// ---- Intra prediction (clause 8.3), as src/h264/intra.rs -----------------
// Which neighbours of a block may be read (IntraAvail, src/h264/intra.rs:
// 10-21).
struct Av {
    bool top, left, tl, tr;
};

// Reference samples of an N x N block (N = 4, 8): t[0..2N) above, the
// top-right half repeated from t[N - 1] where unavailable; l[0..N) to the
// left; c the corner; zero where unavailable (Edge4::gather / Edge8::gather,
// src/h264/intra.rs:39-75, 333-362). d points at the block's first sample.
struct Nbr {
    int t[16], l[8], c;
};
Nbr gather_nbr(const uint8_t* d, long stride, int n, Av av) {
    Nbr e;
    std::memset(&e, 0, sizeof(e));
    if (av.top) {
        for (int x = 0; x < n; ++x) e.t[x] = d[x - stride];
        for (int x = n; x < 2 * n; ++x) e.t[x] = av.tr ? d[x - stride] : e.t[n - 1];
    }
    if (av.left)
        for (int y = 0; y < n; ++y) e.l[y] = d[y * stride - 1];
    if (av.tl) e.c = d[-stride - 1];
    return e;
}

// The Intra_8x8 reference filter (8.3.2.2.1, src/h264/intra.rs:363-404):
// [1 2 1] / 4 along the edge, the ends mirrored or taken from the corner.
Nbr filter8(const Nbr& e, Av av) {
    Nbr f;
    std::memset(&f, 0, sizeof(f));
    if (av.top) {
        f.t[0] = av.tl ? (e.c + 2 * e.t[0] + e.t[1] + 2) >> 2 : (3 * e.t[0] + e.t[1] + 2) >> 2;
        for (int x = 1; x < 15; ++x) f.t[x] = (e.t[x - 1] + 2 * e.t[x] + e.t[x + 1] + 2) >> 2;
        f.t[15] = (e.t[14] + 3 * e.t[15] + 2) >> 2;
    }
    if (av.tl) {
        if (av.top && av.left) f.c = (e.t[0] + 2 * e.c + e.l[0] + 2) >> 2;
        else if (av.top) f.c = (3 * e.c + e.t[0] + 2) >> 2;
        else if (av.left) f.c = (3 * e.c + e.l[0] + 2) >> 2;
        else f.c = e.c;
    }
    if (av.left) {
        f.l[0] = av.tl ? (e.c + 2 * e.l[0] + e.l[1] + 2) >> 2 : (3 * e.l[0] + e.l[1] + 2) >> 2;
        for (int y = 1; y < 7; ++y) f.l[y] = (e.l[y - 1] + 2 * e.l[y] + e.l[y + 1] + 2) >> 2;
        f.l[7] = (e.l[6] + 3 * e.l[7] + 2) >> 2;
    }
    return f;
}

// Whether an Intra_4x4 / Intra_8x8 mode has the neighbours it reads
// (mode_nxn_available, src/h264/intra.rs:78-88).
bool nxn_mode_ok(int mode, Av av) {
    switch (mode) {
        case 0: case 3: case 7: return av.top;
        case 1: case 8: return av.left;
        case 2: return true;
        case 4: case 5: case 6: return av.top && av.left && av.tl;
        default: return false;
    }
}

// The nine directional modes of an N x N block (N = 4: 8.3.1.2.1-9, N = 8:
// 8.3.2.2.2-10; pred_4x4 / pred_8x8, src/h264/intra.rs:109-325, 427-620),
// written once for both sizes: the 4x4 and 8x8 formulas differ only in N.
// out is raster with stride N; false when the mode is out of range or lacks
// its neighbours.
bool pred_nxn(const Nbr& e, int n, int mode, Av av, int bd, uint8_t* out) {
    if (!nxn_mode_ok(mode, av)) return false;
    const int lg = n == 4 ? 2 : 3;
    // p[x, -1] for x >= -1 on the top row (x = -1 the corner), p[-1, y]
    // for y >= 0 down the left column.
    auto T = [&](int x) { return x < 0 ? e.c : e.t[x]; };
    auto L = [&](int y) { return y < 0 ? e.c : e.l[y]; };
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            int v = 0;
            switch (mode) {
                case 0: v = e.t[x]; break;
                case 1: v = e.l[y]; break;
                case 2: {
                    int st = 0, sl = 0;
                    for (int k = 0; k < n; ++k) st += e.t[k], sl += e.l[k];
                    v = av.top && av.left ? (st + sl + n) >> (lg + 1)
                        : av.top          ? (st + n / 2) >> lg
                        : av.left         ? (sl + n / 2) >> lg
                                          : 1 << (bd - 1);
                    break;
                }
                case 3:  // diagonal down left
                    v = (x == n - 1 && y == n - 1) ? (e.t[2 * n - 2] + 3 * e.t[2 * n - 1] + 2) >> 2
                                                   : (e.t[x + y] + 2 * e.t[x + y + 1] + e.t[x + y + 2] + 2) >> 2;
                    break;
                case 4:  // diagonal down right
                    if (x > y) v = (T(x - y - 2) + 2 * T(x - y - 1) + T(x - y) + 2) >> 2;
                    else if (x < y) v = (L(y - x - 2) + 2 * L(y - x - 1) + L(y - x) + 2) >> 2;
                    else v = (T(0) + 2 * e.c + L(0) + 2) >> 2;
                    break;
                case 5: {  // vertical right
                    const int z = 2 * x - y;
                    if (z >= 0 && !(z & 1)) v = (T(x - (y >> 1) - 1) + T(x - (y >> 1)) + 1) >> 1;
                    else if (z >= 0) v = (T(x - (y >> 1) - 2) + 2 * T(x - (y >> 1) - 1) + T(x - (y >> 1)) + 2) >> 2;
                    else if (z == -1) v = (L(0) + 2 * e.c + T(0) + 2) >> 2;
                    else v = (L(y - 2 * x - 1) + 2 * L(y - 2 * x - 2) + L(y - 2 * x - 3) + 2) >> 2;
                    break;
                }
                case 6: {  // horizontal down
                    const int z = 2 * y - x;
                    if (z >= 0 && !(z & 1)) v = (L(y - (x >> 1) - 1) + L(y - (x >> 1)) + 1) >> 1;
                    else if (z >= 0) v = (L(y - (x >> 1) - 2) + 2 * L(y - (x >> 1) - 1) + L(y - (x >> 1)) + 2) >> 2;
                    else if (z == -1) v = (L(0) + 2 * e.c + T(0) + 2) >> 2;
                    else v = (T(x - 2 * y - 1) + 2 * T(x - 2 * y - 2) + T(x - 2 * y - 3) + 2) >> 2;
                    break;
                }
                case 7:  // vertical left
                    v = (y & 1) ? (e.t[x + (y >> 1)] + 2 * e.t[x + (y >> 1) + 1] + e.t[x + (y >> 1) + 2] + 2) >> 2
                                : (e.t[x + (y >> 1)] + e.t[x + (y >> 1) + 1] + 1) >> 1;
                    break;
                default: {  // 8: horizontal up
                    const int z = x + 2 * y;
                    if (z > 2 * n - 3) v = e.l[n - 1];
                    else if (z == 2 * n - 3) v = (e.l[n - 2] + 3 * e.l[n - 1] + 2) >> 2;
                    else if (!(z & 1)) v = (e.l[y + (x >> 1)] + e.l[y + (x >> 1) + 1] + 1) >> 1;
                    else v = (e.l[y + (x >> 1)] + 2 * e.l[y + (x >> 1) + 1] + e.l[y + (x >> 1) + 2] + 2) >> 2;
                    break;
                }
            }
            out[y * n + x] = (uint8_t)v;
        }
    return true;
}

// Intra_16x16 (8.3.3; predict_16x16 / predict_planar_block,
// src/h264/intra.rs:622-646, 792-923) at d: 0 vertical, 1 horizontal,
// 2 DC, 3 plane.
bool pred16(uint8_t* d, long stride, int mode, Av av, int bd) {
    const int max = (1 << bd) - 1;
    auto top = [&](int x) { return (int)d[x - stride]; };   // x = -1: the corner
    auto left = [&](int y) { return (int)d[y * stride - 1]; };  // y = -1: the corner
    uint8_t out[256];
    if (mode == 0) {
        if (!av.top) return false;
        for (int i = 0; i < 256; ++i) out[i] = (uint8_t)top(i & 15);
    } else if (mode == 1) {
        if (!av.left) return false;
        for (int i = 0; i < 256; ++i) out[i] = (uint8_t)left(i >> 4);
    } else if (mode == 2) {
        int st = 0, sl = 0;
        for (int k = 0; k < 16; ++k) st += av.top ? top(k) : 0, sl += av.left ? left(k) : 0;
        const int v = av.top && av.left ? (st + sl + 16) >> 5 : av.top ? (st + 8) >> 4 : av.left ? (sl + 8) >> 4
                                                                                                  : 1 << (bd - 1);
        std::memset(out, v, 256);
    } else if (mode == 3) {
        if (!(av.top && av.left && av.tl)) return false;
        int hh = 0, vv = 0;
        for (int i = 0; i < 8; ++i) {
            hh += (i + 1) * (top(8 + i) - top(6 - i));
            vv += (i + 1) * (left(8 + i) - left(6 - i));
        }
        const int b = (5 * hh + 32) >> 6, c = (5 * vv + 32) >> 6, a = 16 * (left(15) + top(15));
        for (int y = 0; y < 16; ++y)
            for (int x = 0; x < 16; ++x) out[y * 16 + x] = (uint8_t)clampi((a + b * (x - 7) + c * (y - 7) + 16) >> 5, 0, max);
    } else {
        return false;
    }
    for (int y = 0; y < 16; ++y) std::memcpy(d + y * stride, out + y * 16, 16);
    return true;
}

// 4:2:0 chroma prediction (8.3.4; predict_chroma, src/h264/intra.rs:648-790)
// of the 8 x 8 block at d: 0 DC per 4x4 block (the top-left and bottom-right
// blocks average both edges, the top-right prefers the top, the bottom-left
// the left), 1 horizontal, 2 vertical, 3 plane.
bool pred_chroma(uint8_t* d, long stride, int mode, Av av, int bd) {
    const int max = (1 << bd) - 1;
    auto top = [&](int x) { return (int)d[x - stride]; };
    auto left = [&](int y) { return (int)d[y * stride - 1]; };
    uint8_t out[64];
    if (mode == 0) {
        for (int by = 0; by < 2; ++by)
            for (int bx = 0; bx < 2; ++bx) {
                int st = 0, sl = 0;
                for (int k = 0; k < 4; ++k) {
                    st += av.top ? top(bx * 4 + k) : 0;
                    sl += av.left ? left(by * 4 + k) : 0;
                }
                int v;
                const int none = 1 << (bd - 1);
                if (bx == by) v = av.top && av.left ? (st + sl + 4) >> 3 : av.top ? (st + 2) >> 2 : av.left ? (sl + 2) >> 2 : none;
                else if (bx) v = av.top ? (st + 2) >> 2 : av.left ? (sl + 2) >> 2 : none;
                else v = av.left ? (sl + 2) >> 2 : av.top ? (st + 2) >> 2 : none;
                for (int y = 0; y < 4; ++y)
                    for (int x = 0; x < 4; ++x) out[(by * 4 + y) * 8 + bx * 4 + x] = (uint8_t)v;
            }
    } else if (mode == 1) {
        if (!av.left) return false;
        for (int i = 0; i < 64; ++i) out[i] = (uint8_t)left(i >> 3);
    } else if (mode == 2) {
        if (!av.top) return false;
        for (int i = 0; i < 64; ++i) out[i] = (uint8_t)top(i & 7);
    } else if (mode == 3) {
        if (!(av.top && av.left && av.tl)) return false;
        int hh = 0, vv = 0;
        for (int i = 0; i < 4; ++i) {
            hh += (i + 1) * (top(4 + i) - top(2 - i));
            vv += (i + 1) * (left(4 + i) - left(2 - i));
        }
        const int b = (34 * hh + 32) >> 6, c = (34 * vv + 32) >> 6, a = 16 * (left(7) + top(7));
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) out[y * 8 + x] = (uint8_t)clampi((a + b * (x - 3) + c * (y - 3) + 16) >> 5, 0, max);
    } else {
        return false;
    }
    for (int y = 0; y < 8; ++y) std::memcpy(d + y * stride, out + y * 8, 8);
    return true;
}

// Decoding index of the 4x4 block at raster (bx, by) of a macroblock: 8x8
// quadrants in z order, 4x4 blocks in z order within (BLK4X4_FROM_RASTER).
inline int blk4_order(int bx, int by) { return ((by >> 1) * 2 + (bx >> 1)) * 4 + (by & 1) * 2 + (bx & 1); }

// Availability of a macroblock's neighbours for intra prediction in a
// one-slice frame: left (A), above (B), above-right (C), above-left (D),
// each present and, under constrained_intra_pred, intra (intra_ok,
// src/h264/recon.rs:301-307).
struct MbAv {
    bool a, b, c, d;
};
MbAv mb_avail(const rv_h264_pic* pic, int addr) {
    const int mbw = pic->width / 16, mbx = addr % mbw, mby = addr / mbw;
    auto ok = [&](int n) { return !pic->constrained_intra_pred || pic->mbs[n].intra != 0; };
    MbAv m;
    m.a = mbx > 0 && ok(addr - 1);
    m.b = mby > 0 && ok(addr - mbw);
    m.c = mby > 0 && mbx < mbw - 1 && ok(addr - mbw + 1);
    m.d = mbx > 0 && mby > 0 && ok(addr - mbw - 1);
    return m;
}
// intra_avail_4x4 / intra_avail_8x8 (src/h264/recon.rs:936-1011) for the
// block at 4x4 raster (bx, by) of size 4 or 8.
Av block_avail(const MbAv& m, int bx, int by, int size) {
    Av av;
    av.left = bx > 0 || m.a;
    av.top = by > 0 || m.b;
    av.tl = (bx > 0 && by > 0) || (bx == 0 ? (by == 0 ? m.d : m.a) : m.b);
    const int w4 = size / 4;
    if (by == 0) av.tr = bx + w4 < 4 ? m.b : m.c;
    else if (size == 8) av.tr = bx == 0;
    else av.tr = bx < 3 && blk4_order(bx + 1, by - 1) < blk4_order(bx, by);
    return av;
}

// Residual of one listed block into the picture (residual4 / residual8,
// scaling levels first).
bool add_block(const rv_h264_pic* pic, const rv_h264_blk& bk, uint8_t* dst, long stride) {
    const int n = bk.size;
    if (bk.coef_off + (uint64_t)n * n > pic->n_coefs) return false;
    if (bk.flags & RV_BLK_LEVELS) {
        int32_t scaled[64];
        int dc;
        if (!scale_block(pic, bk, scaled, &dc)) return false;
        if (n == 8) residual8(dst, stride, scaled, 255);
        else residual4(dst, stride, scaled, dc, bk.has_dc != 0, 255);
    } else if (n == 8) {
        residual8(dst, stride, pic->coefs + bk.coef_off, 255);
    } else {
        residual4(dst, stride, pic->coefs + bk.coef_off, bk.dc, bk.has_dc != 0, 255);
    }
    return true;
}
}  // namespace

extern "C" {

void orc_h264_qpel(int pos, uint8_t* dst, const uint8_t* src, int stride, int w, int h, int max) {
    qpel(pos, dst, src, stride, w, h, max);
}
void orc_h264_chroma(uint8_t* dst, const uint8_t* src, int stride, int w, int h, int xf, int yf) {
    chroma(dst, src, stride, w, h, xf, yf);
}
void orc_h264_copy(uint8_t* dst, int stride, const uint8_t* src, int w, int h) { copy_k(dst, stride, src, w, h); }
void orc_h264_idct4_add(uint8_t* dst, int stride, const int16_t* coeffs, int max) {
    idct4_add(dst, stride, coeffs, max);
}
void orc_h264_idct8_add(uint8_t* dst, int stride, const int16_t* coeffs, int max) {
    idct8_add(dst, stride, coeffs, max);
}
void orc_h264_dc_add(uint8_t* dst, int stride, int dc, int n, int max) { dc_add(dst, stride, dc, n, max); }
void orc_h264_avg(uint8_t* dst, int stride, const uint8_t* a, const uint8_t* b, int w, int h) {
    avg_k(dst, stride, a, b, w, h);
}
void orc_h264_weighted_uni(uint8_t* dst, int stride, const uint8_t* src, int w, int h, int log_wd, int wt, int o,
                           int max) {
    wuni_k(dst, stride, src, w, h, log_wd, wt, o, max);
}
void orc_h264_weighted_bi(uint8_t* dst, int stride, const uint8_t* a, const uint8_t* b, int w, int h, int log_wd,
                          int w0, int w1, int o0, int o1, int max) {
    wbi_k(dst, stride, a, b, w, h, log_wd, w0, w1, o0, o1, max);
}
void orc_h264_deblock(int chroma_k, int vertical, uint8_t* data, int off, int stride, int alpha, int beta,
                      const int16_t* tc0, int max) {
    long step = vertical ? stride : 1, across = vertical ? 1 : stride;
    deblock_lines(chroma_k != 0, data, off, step, across, alpha, beta, tc0, max);
}
void orc_h264_residual4(uint8_t* dst, int stride, const int32_t* coefs, int dc, int has_dc, int max) {
    residual4(dst, stride, coefs, dc, has_dc != 0, max);
}
int orc_h264_level_scale4(const rv_h264_scaling* sl, int list, int m, int pos) {
    return level_scale4(sl, list, m, pos);
}
int orc_h264_level_scale8(const rv_h264_scaling* sl, int list, int m, int pos) {
    return level_scale8(sl, list, m, pos);
}
int32_t orc_h264_dequant_level(int32_t level, int32_t table, int shift) { return dequant_level(level, table, shift); }
void orc_h264_luma_dc(int32_t* dc, int scale00, int qp) { luma_dc_transform(dc, scale00, qp); }
void orc_h264_chroma_dc(int32_t* dc, int scale00, int qp) { chroma_dc_transform_420(dc, scale00, qp); }
void orc_h264_residual8(uint8_t* dst, int stride, const int32_t* coefs, int max) {
    residual8(dst, stride, coefs, max);
}
void orc_h264_implicit_pair(int cur_poc, int poc0, int poc1, int long0, int long1, int32_t out[2]) {
    int w0, w1;
    implicit_pair(cur_poc, poc0, poc1, long0 != 0, long1 != 0, &w0, &w1);
    out[0] = w0;
    out[1] = w1;
}

int orc_h264_bs_map(const rv_h264_pic* pic, uint32_t* out) {
    if (!pic || !out) return RV_E_INVALID;
    int W = pic->width, H = pic->height;
    if (W <= 0 || H <= 0 || (W & 15) || (H & 15) || pic->n_mbs != (W / 16) * (H / 16)) return RV_E_RANGE;
    MbSide sd;
    int rc = build_side(pic, sd);
    if (rc) return rc;
    for (int a = 0; a < pic->n_mbs; ++a) mb_strengths(pic, sd, a, out + 8 * a, out + 8 * a + 4);
    return RV_OK;
}

// Intra predictors on a caller's plane (d = the block's first sample;
// every neighbour the availability flags admit must be readable). kind 0:
// N x N directional (n = 4 or 8, with the 8x8 reference filter), 1:
// Intra_16x16, 2: 4:2:0 chroma. av: bit 0 top, 1 left, 2 top-left,
// 3 top-right. Returns 0, or RV_E_INVALID for a mode without its
// neighbours.
int orc_h264_intra_pred(int kind, int n, int mode, int av, uint8_t* d, int stride) {
    const Av a{(av & 1) != 0, (av & 2) != 0, (av & 4) != 0, (av & 8) != 0};
    if (kind == 0) {
        if (n != 4 && n != 8) return RV_E_RANGE;
        Nbr e = gather_nbr(d, stride, n, a);
        if (n == 8) e = filter8(e, a);
        uint8_t out[64];
        if (!pred_nxn(e, n, mode, a, 8, out)) return RV_E_INVALID;
        for (int y = 0; y < n; ++y) std::memcpy(d + (long)y * stride, out + y * n, n);
        return RV_OK;
    }
    if (kind == 1) return pred16(d, stride, mode, a, 8) ? RV_OK : RV_E_INVALID;
    if (kind == 2) return pred_chroma(d, stride, mode, a, 8) ? RV_OK : RV_E_INVALID;
    return RV_E_RANGE;
}

// The Intra_8x8 reference filter alone: in = t[16], l[8], c; out likewise.
void orc_h264_filter8(const int32_t in[25], int av, int32_t out[25]) {
    const Av a{(av & 1) != 0, (av & 2) != 0, (av & 4) != 0, (av & 8) != 0};
    Nbr e;
    for (int i = 0; i < 16; ++i) e.t[i] = in[i];
    for (int i = 0; i < 8; ++i) e.l[i] = in[16 + i];
    e.c = in[24];
    const Nbr f = filter8(e, a);
    for (int i = 0; i < 16; ++i) out[i] = f.t[i];
    for (int i = 0; i < 8; ++i) out[16 + i] = f.l[i];
    out[24] = f.c;
}

// Availability (bit 0 top, 1 left, 2 top-left, 3 top-right) of the block
// at 4x4 raster (bx, by), size 4 or 8 (16: the whole macroblock), of
// macroblock addr.
int orc_h264_block_avail(const rv_h264_pic* pic, int addr, int bx, int by, int size) {
    const MbAv m = mb_avail(pic, addr);
    const Av a = size == 16 ? Av{m.b, m.a, m.d, false} : block_avail(m, bx, by, size);
    return (a.top ? 1 : 0) | (a.left ? 2 : 0) | (a.tl ? 4 : 0) | (a.tr ? 8 : 0);
}

int orc_h264_reconstruct(const rv_h264_pic* pic, const void* const* slot_planes, int n_slots, void* const out[3],
                         int stop_after) {
    if (!pic || !slot_planes || !out) return RV_E_INVALID;
    if (pic->bit_depth != 8) return RV_E_RANGE;
    int W = pic->width, H = pic->height;
    if (W <= 0 || H <= 0 || (W & 15) || (H & 15)) return RV_E_RANGE;
    int mbw = W / 16, mbh = H / 16;
    if (pic->n_mbs != mbw * mbh) return RV_E_RANGE;
    int max = 255;
    std::vector<Frame> refs(n_slots);
    std::vector<bool> loaded(n_slots, false);
    auto ref_frame = [&](int slot) -> const Frame* {
        if (slot < 0 || slot >= n_slots || !slot_planes[3 * slot]) return nullptr;
        if (!loaded[slot]) {
            refs[slot].init(W, H);
            for (int c = 0; c < 3; ++c) {
                refs[slot].p[c].load((const uint8_t*)slot_planes[3 * slot + c]);
                refs[slot].p[c].extend();
            }
            loaded[slot] = true;
        }
        return &refs[slot];
    };
    Frame cur;
    cur.init(W, H);
    MbSide sd;
    int rc = build_side(pic, sd);
    if (rc) return rc;
    // Intra macroblocks: host-supplied prediction.
    for (int a = 0; a < pic->n_mbs; ++a) {
        const rv_h264_mb& m = pic->mbs[a];
        if (m.intra != RV_INTRA_HOST) continue;
        int mx = (a % mbw) * 16, my = (a / mbw) * 16;
        const uint16_t* pr = pic->intra_pred + m.pred_off;
        for (int c = 0; c < 3; ++c) {
            int n = c ? 8 : 16, px = c ? mx / 2 : mx, py = c ? my / 2 : my;
            Plane& pl = cur.p[c];
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x) pl.data[pl.offset(px + x, py + y)] = (uint8_t)pr[y * n + x];
            pr += n * n;
        }
    }
    // Inter partitions (predict_partition, src/h264/inter.rs:172-332).
    std::vector<uint8_t> pred[2][3];
    for (auto& l : pred)
        for (auto& c : l) c.assign(16 * PRED_STRIDE, 0);
    std::vector<uint8_t> window(32 * 32);
    for (int i = 0; i < pic->n_parts; ++i) {
        const rv_h264_part& pt = pic->parts[i];
        int x = pt.x, y = pt.y, w = pt.w, h = pt.h, cw = w / 2, ch = h / 2;
        const Frame* rf[2] = {nullptr, nullptr};
        for (int l = 0; l < 2; ++l) {
            if (!((pt.dir >> l) & 1)) continue;
            int ri = pt.ref_idx[l];
            if (ri < 0 || ri >= RV_MAX_REFS || ri >= pic->num_ref[l]) return RV_E_RANGE;
            rf[l] = ref_frame(pic->ref_slot[l][ri]);
            if (!rf[l]) return RV_E_RANGE;
            int mvx = pt.mv[l].x, mvy = pt.mv[l].y;
            int xi = x + (mvx >> 2), yi = y + (mvy >> 2);
            int pos = (mvy & 3) * 4 + (mvx & 3);
            interp(rf[l]->p[0], xi - 2, yi - 2, w + 5, h + 5, window,
                   [&](const uint8_t* s, long st) { qpel(pos, pred[l][0].data(), s, st, w, h, max); });
            int xci = x / 2 + (mvx >> 3), yci = y / 2 + (mvy >> 3), xf = mvx & 7, yf = mvy & 7;
            for (int c = 1; c < 3; ++c)
                interp(rf[l]->p[c], xci, yci, cw + 1, ch + 1, window,
                       [&](const uint8_t* s, long st) { chroma(pred[l][c].data(), s, st, cw, ch, xf, yf); });
        }
        // Weighting (src/h264/recon.rs:142-170).
        bool both = rf[0] && rf[1];
        int single = rf[0] ? 0 : 1;
        bool weighted = false;
        int log_wd[3] = {0, 0, 0}, wt[3][2] = {{1, 1}, {1, 1}, {1, 1}}, o[3][2] = {{0, 0}, {0, 0}, {0, 0}};
        if (pic->weighting == RV_H264_EXPLICIT) {
            weighted = true;
            for (int c = 0; c < 3; ++c) {
                log_wd[c] = pic->log_wd[c];
                for (int l = 0; l < 2; ++l)
                    if (rf[l]) {
                        wt[c][l] = pic->wp_w[l][pt.ref_idx[l]][c];
                        o[c][l] = pic->wp_o[l][pt.ref_idx[l]][c];
                    }
            }
        } else if (pic->weighting == RV_H264_IMPLICIT && both) {
            weighted = true;
            int w0, w1;
            implicit_pair(pic->poc, pic->ref_poc[0][pt.ref_idx[0]], pic->ref_poc[1][pt.ref_idx[1]],
                          pic->ref_long_term[0][pt.ref_idx[0]] != 0, pic->ref_long_term[1][pt.ref_idx[1]] != 0, &w0,
                          &w1);
            for (int c = 0; c < 3; ++c) {
                log_wd[c] = 5;
                wt[c][0] = w0;
                wt[c][1] = w1;
            }
        }
        for (int c = 0; c < 3; ++c) {
            Plane& pl = cur.p[c];
            int px = c ? x / 2 : x, py = c ? y / 2 : y, pw = c ? cw : w, ph = c ? ch : h;
            uint8_t* dst = &pl.data[pl.offset(px, py)];
            const uint8_t* a = pred[0][c].data();
            const uint8_t* b = pred[1][c].data();
            if (!weighted) {
                if (!both) copy_k(dst, pl.stride, single == 0 ? a : b, pw, ph);
                else avg_k(dst, pl.stride, a, b, pw, ph);
            } else if (!both) {
                wuni_k(dst, pl.stride, single == 0 ? a : b, pw, ph, log_wd[c], wt[c][single], o[c][single], max);
            } else {
                wbi_k(dst, pl.stride, a, b, pw, ph, log_wd[c], wt[c][0], wt[c][1], o[c][0], o[c][1], max);
            }
        }
    }
    // Residual (src/h264/recon.rs:870-926, 1200-1226); the blocks of
    // macroblocks predicted here are kept for their macroblock.
    std::vector<std::vector<int>> own((size_t)pic->n_mbs);
    if (stop_after != 1) {
        for (int i = 0; i < pic->n_blks; ++i) {
            const rv_h264_blk& bk = pic->blks[i];
            if (bk.comp > 2 || (bk.size != 4 && bk.size != 8)) return RV_E_RANGE;
            const int mb = bk.comp ? (bk.y / 8) * mbw + bk.x / 8 : (bk.y / 16) * mbw + bk.x / 16;
            if (mb < 0 || mb >= pic->n_mbs) return RV_E_RANGE;
            if (pic->mbs[mb].intra == RV_INTRA_DEVICE) {
                own[(size_t)mb].push_back(i);
                continue;
            }
            Plane& pl = cur.p[bk.comp];
            if (!add_block(pic, bk, &pl.data[pl.offset(bk.x, bk.y)], pl.stride)) return RV_E_RANGE;
        }
        // Macroblocks predicted from their neighbours, in decoding order
        // (src/h264/recon.rs:611-828, 1021-1128): luma block by block in
        // decoding order, each block's residual added before the next is
        // predicted (Intra_4x4, Intra_8x8) or the whole prediction first
        // (Intra_16x16); then each chroma component and its residual.
        for (int a = 0; a < pic->n_mbs; ++a) {
            const rv_h264_mb& m = pic->mbs[a];
            if (m.intra != RV_INTRA_DEVICE) continue;
            const int mx = (a % mbw) * 16, my = (a / mbw) * 16;
            const MbAv mav = mb_avail(pic, a);
            // This macroblock's block at (comp, x, y), or -1.
            auto find = [&](int comp, int x, int y) {
                for (int i : own[(size_t)a])
                    if (pic->blks[i].comp == comp && pic->blks[i].x == x && pic->blks[i].y == y) return i;
                return -1;
            };
            Plane& Y = cur.p[0];
            auto residual_at = [&](int comp, int x, int y) {
                const int i = find(comp, x, y);
                if (i < 0) return true;
                Plane& pl = cur.p[comp];
                return add_block(pic, pic->blks[i], &pl.data[pl.offset(x, y)], pl.stride);
            };
            if (m.intra_kind == RV_H264_I16X16) {
                uint8_t* d = &Y.data[Y.offset(mx, my)];
                if (!pred16(d, Y.stride, m.modes[0], Av{mav.b, mav.a, mav.d, false}, 8)) return RV_E_INVALID;
                for (int r = 0; r < 16; ++r)
                    if (!residual_at(0, mx + (r & 3) * 4, my + (r >> 2) * 4)) return RV_E_RANGE;
            } else if (m.intra_kind == RV_H264_I4X4 || m.intra_kind == RV_H264_I8X8) {
                const int n = m.intra_kind == RV_H264_I8X8 ? 8 : 4, nb = 16 / (n * n / 16);
                for (int k = 0; k < nb; ++k) {
                    // Decoding order: 8x8 quadrants in z order, and 4x4
                    // blocks in z order within each.
                    int bx, by;
                    if (n == 8) {
                        bx = (k & 1) * 2;
                        by = (k >> 1) * 2;
                    } else {
                        bx = ((k >> 2) & 1) * 2 + (k & 1);
                        by = (k >> 3) * 2 + ((k >> 1) & 1);
                    }
                    const Av av = block_avail(mav, bx, by, n);
                    uint8_t* d = &Y.data[Y.offset(mx + bx * 4, my + by * 4)];
                    Nbr e = gather_nbr(d, Y.stride, n, av);
                    if (n == 8) e = filter8(e, av);
                    uint8_t out[64];
                    if (!pred_nxn(e, n, m.modes[by * 4 + bx], av, 8, out)) return RV_E_INVALID;
                    for (int y = 0; y < n; ++y) std::memcpy(d + y * Y.stride, out + y * n, n);
                    if (!residual_at(0, mx + bx * 4, my + by * 4)) return RV_E_RANGE;
                }
            } else {
                return RV_E_INVALID;
            }
            for (int comp = 1; comp < 3; ++comp) {
                Plane& pl = cur.p[comp];
                if (!pred_chroma(&pl.data[pl.offset(mx / 2, my / 2)], pl.stride, m.chroma_mode,
                                 Av{mav.b, mav.a, mav.d, false}, 8))
                    return RV_E_INVALID;
                for (int b = 0; b < 4; ++b)
                    if (!residual_at(comp, mx / 2 + (b & 1) * 4, my / 2 + (b >> 1) * 4)) return RV_E_RANGE;
            }
        }
    }
    // Deblocking (deblock_mb_rows, src/h264/deblock.rs:567-815), frame
    // picture, one slice, disable_deblocking_filter_idc 0.
    if ((stop_after == 0 || stop_after >= 3) && pic->deblock) {
        Thr thr[52];
        for (int qa = 0; qa < 52; ++qa) {
            int ia = clampi(qa + pic->offset_a, 0, 51), ib = clampi(qa + pic->offset_b, 0, 51);
            thr[qa].alpha = ALPHA[ia];
            thr[qa].beta = BETA[ib];
            thr[qa].lut[0] = -1;
            for (int k = 0; k < 3; ++k) thr[qa].lut[k + 1] = (int16_t)TC0[ia][k];
        }
        auto get = [&](int qp_p, int qp_q) -> const Thr& { return thr[clampi((qp_p + qp_q + 1) >> 1, 0, 51)]; };
        for (int mby = 0; mby < mbh; ++mby)
            for (int mbx = 0; mbx < mbw; ++mbx) {
                int addr = mby * mbw + mbx;
                const rv_h264_mb& m = pic->mbs[addr];
                int left = mbx > 0 ? addr - 1 : addr, above = mby > 0 ? addr - mbw : addr;
                bool internal_odd = !m.transform_8x8;
                uint32_t bs_v[4], bs_h[4];
                mb_strengths(pic, sd, addr, bs_v, bs_h);
                if ((bs_v[0] | bs_v[1] | bs_v[2] | bs_v[3] | bs_h[0] | bs_h[1] | bs_h[2] | bs_h[3]) == 0) continue;
                const rv_h264_mb& ml = pic->mbs[left];
                const rv_h264_mb& ma = pic->mbs[above];
                // filter_luma_style (src/h264/deblock.rs:352-451).
                {
                    int step = internal_odd ? 1 : 2;
                    uint32_t iv = internal_odd ? (bs_v[1] | bs_v[2] | bs_v[3]) : bs_v[2];
                    uint32_t ih = internal_odd ? (bs_h[1] | bs_h[2] | bs_h[3]) : bs_h[2];
                    if (bs_v[0] | bs_h[0] | iv | ih) {
                        Plane& Y = cur.p[0];
                        long base = Y.offset(mbx * 16, mby * 16);
                        const Thr& ti = get(m.qp, m.qp);
                        if (bs_v[0]) filter_edge(false, true, Y.data.data(), base, Y.stride, bs_v[0], get(ml.qp, m.qp), max);
                        for (int e = step; e < 4; e += step)
                            if (bs_v[e]) filter_edge(false, true, Y.data.data(), base + e * 4, Y.stride, bs_v[e], ti, max);
                        if (bs_h[0]) filter_edge(false, false, Y.data.data(), base, Y.stride, bs_h[0], get(ma.qp, m.qp), max);
                        for (int e = step; e < 4; e += step)
                            if (bs_h[e])
                                filter_edge(false, false, Y.data.data(), base + e * 4 * Y.stride, Y.stride, bs_h[e], ti,
                                            max);
                    }
                }
                // filter_chroma_style, 4:2:0 (src/h264/deblock.rs:453-560).
                if (bs_v[0] | bs_h[0] | bs_v[2] | bs_h[2]) {
                    Plane& Cb = cur.p[1];
                    Plane& Cr = cur.p[2];
                    long base = Cb.offset(mbx * 8, mby * 8);
                    long st = Cb.stride;
                    const Thr& tib = get(m.qpc[0], m.qpc[0]);
                    const Thr& tir = get(m.qpc[1], m.qpc[1]);
                    if (bs_v[0]) {
                        filter_edge(true, true, Cb.data.data(), base, st, bs_v[0], get(ml.qpc[0], m.qpc[0]), max);
                        filter_edge(true, true, Cr.data.data(), base, st, bs_v[0], get(ml.qpc[1], m.qpc[1]), max);
                    }
                    if (bs_v[2]) {
                        filter_edge(true, true, Cb.data.data(), base + 4, st, bs_v[2], tib, max);
                        filter_edge(true, true, Cr.data.data(), base + 4, st, bs_v[2], tir, max);
                    }
                    if (bs_h[0]) {
                        filter_edge(true, false, Cb.data.data(), base, st, bs_h[0], get(ma.qpc[0], m.qpc[0]), max);
                        filter_edge(true, false, Cr.data.data(), base, st, bs_h[0], get(ma.qpc[1], m.qpc[1]), max);
                    }
                    if (bs_h[2]) {
                        filter_edge(true, false, Cb.data.data(), base + 4 * st, st, bs_h[2], tib, max);
                        filter_edge(true, false, Cr.data.data(), base + 4 * st, st, bs_h[2], tir, max);
                    }
                }
            }
    }
    for (int c = 0; c < 3; ++c) cur.p[c].store((uint8_t*)out[c]);
    return RV_OK;
}

}  // extern "C"
