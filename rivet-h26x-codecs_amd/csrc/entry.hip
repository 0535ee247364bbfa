// The per-entry API: one extern "C" function per entry of the reference's
// kernel tables, HevcDsp<S> (src/dsp/hevc.rs:165-229) and H264Dsp<u8>
// (src/dsp/h264.rs:123-170), with the entry's arguments in order. A Rust
// slice becomes (pointer, length in elements).
//
// Each call:
//  1. computes, from the arguments alone, the span of every slice the
//     kernel touches, and refuses with RV_E_RANGE any span that leaves its
//     slice (the reference panics on that index);
//  2. copies the spans into a device staging buffer;
//  3. runs the entry's kernel (hevc_entry / h264_entry) on the library
//     stream;
//  4. copies the written spans back and waits.
// Calls hold the library lock for their whole duration, so one staging
// buffer serves them all; it grows to the largest call and is kept.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/rivet_hip.h"
#include "rv_device.h"
#include "rv_kernels.h"
#include "rv_runtime.h"

namespace {

using rv::lib_fail;

void* g_stage = nullptr;
size_t g_stage_cap = 0;

constexpr int64_t kMaxDim = 1 << 16;       // block width / height / stride bound
constexpr int64_t kMaxSpan = int64_t(1) << 34;

// One slice the kernel touches: elements [lo, hi) of `host`.
struct Span {
    const void* host;
    size_t elem;
    int64_t lo, hi;
    bool out;
    char* dev;  // device copy of element lo
};

int range(const char* what) { return lib_fail(RV_E_RANGE, std::string(what)); }

// Elements lo..hi (inclusive) must lie inside a slice of `len` elements.
bool inside(int64_t lo, int64_t hi, size_t len) { return lo >= 0 && hi >= lo && (uint64_t)hi < (uint64_t)len; }

// Last index of a w x h block at `base` with row stride `stride`.
int64_t block_last(int64_t base, int64_t stride, int64_t w, int64_t h) { return base + (h - 1) * stride + (w - 1); }

bool dims_ok(size_t w, size_t h) {
    return w <= (size_t)kMaxDim && h <= (size_t)kMaxDim && (uint64_t)w * h <= (uint64_t(1) << 24);
}
bool stride_ok(size_t s) { return s <= (size_t)kMaxSpan; }

// Copy the spans in, launch, copy the written ones back, wait.
template <typename Launch>
int run(Span* spans, int n, Launch launch) {
    try {
        std::lock_guard<std::mutex> lk(rv::lib_mutex());
        if (!rv::lib_ready()) return lib_fail(RV_E_NODEVICE, "rv_init not called");
        hipStream_t st = rv::lib_stream();
        size_t total = 0;
        for (int i = 0; i < n; ++i) total += ((size_t)(spans[i].hi - spans[i].lo) * spans[i].elem + 255) / 256 * 256;
        if (total > g_stage_cap) {
            if (g_stage) {
                hipError_t e = hipStreamSynchronize(st);
                if (e != hipSuccess) return lib_fail(RV_E_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
                (void)hipFree(g_stage);
                g_stage = nullptr;
                g_stage_cap = 0;
            }
            size_t cap = std::max<size_t>(total, 1 << 20);
            if (hipMalloc(&g_stage, cap) != hipSuccess) {
                g_stage = nullptr;
                return lib_fail(RV_E_NOMEM, "per-entry staging allocation failed");
            }
            g_stage_cap = cap;
        }
        char* p = (char*)g_stage;
        for (int i = 0; i < n; ++i) {
            Span& s = spans[i];
            s.dev = p;
            const size_t bytes = (size_t)(s.hi - s.lo) * s.elem;
            p += (bytes + 255) / 256 * 256;
            if (bytes) {
                hipError_t e = hipMemcpyAsync(s.dev, (const char*)s.host + s.lo * s.elem, bytes, hipMemcpyHostToDevice, st);
                if (e != hipSuccess) return lib_fail(RV_E_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
            }
        }
        hipError_t e = launch(st);
        if (e != hipSuccess) return lib_fail(RV_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
        for (int i = 0; i < n; ++i) {
            const Span& s = spans[i];
            const size_t bytes = (size_t)(s.hi - s.lo) * s.elem;
            if (!s.out || !bytes) continue;
            e = hipMemcpyAsync((char*)s.host + s.lo * s.elem, s.dev, bytes, hipMemcpyDeviceToHost, st);
            if (e != hipSuccess) return lib_fail(RV_E_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
        }
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return lib_fail(RV_E_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
        return RV_OK;
    } catch (...) {
        return lib_fail(RV_E_HIP, "exception in a per-entry call");
    }
}

int elem_blocks(size_t w, size_t h) { return (int)std::min<size_t>((w * h + 255) / 256, 64); }

template <typename S>
constexpr int smax() {
    return sizeof(S) == 1 ? 255 : 65535;
}

// ---------------------------------------------------------------- HEVC

int idct_common(int op, int log2, int16_t* coeffs, size_t len, int bd_shift, size_t max_x, size_t max_y) {
    if (!coeffs) return lib_fail(RV_E_INVALID, "null coeffs");
    if (log2 < 2 || log2 > 5) return range("log2 outside 2..5");
    const int n = 1 << log2;
    if ((size_t)n * n > len) return range("coeffs shorter than n * n");
    if (bd_shift < 1 || bd_shift > 24) return range("bd_shift outside 1..24");
    if (op == rv::HE_IDCT && (max_x >= (size_t)n || max_y >= (size_t)n)) return range("max_x / max_y outside the block");
    rv::HevcEntry e{};
    e.op = op;
    e.w = n;
    e.shift = bd_shift;
    e.max_x = (int)std::min<size_t>(max_x, 31);
    e.max_y = (int)std::min<size_t>(max_y, 31);
    Span sp[1] = {{coeffs, 2, 0, (int64_t)n * n, true, nullptr}};
    return run(sp, 1, [&](hipStream_t st) {
        e.dst = sp[0].dev;
        return rv::hevc_entry_launch(e, 1, st);
    });
}

// copy / h / v filters of samples S, and v2 of i16 (src_is_i16).
template <typename S>
int filt_common(int op, bool luma, int16_t* dst, size_t dst_len, const void* src, size_t src_len, size_t ss, size_t w,
                size_t h, size_t frac, int shift, bool src_is_i16) {
    if (!dst || !src) return lib_fail(RV_E_INVALID, "null buffer");
    if (!dims_ok(w, h) || !stride_ok(ss)) return range("block or stride too large");
    if (!w || !h) return RV_OK;
    const int taps = luma ? 8 : 4;
    if (op != rv::HE_COPY && frac >= (size_t)(luma ? 4 : 8)) return range("frac outside the filter table");
    if (shift < 0 || shift > 15) return range("shift outside 0..15");
    int64_t last;
    if (op == rv::HE_COPY) last = block_last(0, ss, w, h);
    else if (op == rv::HE_FILT_H) last = block_last(0, ss, w + taps - 1, h);
    else last = block_last(0, ss, w, h + taps - 1);
    if (!inside(0, last, src_len)) return range("source window outside src");
    if ((uint64_t)w * h > dst_len) return range("dst shorter than w * h");
    rv::HevcEntry e{};
    e.op = op;
    e.ss = sizeof(S);
    e.luma = luma;
    e.src_stride = (int64_t)ss;
    e.w = (int)w;
    e.h = (int)h;
    e.frac = (int)frac;
    e.shift = shift;
    Span sp[2] = {{src, src_is_i16 ? 2 : sizeof(S), 0, last + 1, false, nullptr},
                  {dst, 2, 0, (int64_t)(w * h), true, nullptr}};
    return run(sp, 2, [&](hipStream_t st) {
        e.src = sp[0].dev;
        e.dst = sp[1].dev;
        return rv::hevc_entry_launch(e, elem_blocks(w, h), st);
    });
}

template <typename S>
int add_residual(S* dst, size_t dst_len, size_t stride, const int16_t* res, size_t res_len, size_t n, int max) {
    if (!dst || !res) return lib_fail(RV_E_INVALID, "null buffer");
    if (!dims_ok(n, n) || !stride_ok(stride)) return range("block or stride too large");
    if (!n) return RV_OK;
    if (max < 0 || max > smax<S>()) return range("max outside the sample range");
    const int64_t last = block_last(0, stride, n, n);
    if (!inside(0, last, dst_len)) return range("block outside dst");
    if ((uint64_t)n * n > res_len) return range("res shorter than n * n");
    rv::HevcEntry e{};
    e.op = rv::HE_ADD_RES;
    e.ss = sizeof(S);
    e.dst_stride = (int64_t)stride;
    e.w = e.h = (int)n;
    e.max = max;
    Span sp[2] = {{dst, sizeof(S), 0, last + 1, true, nullptr}, {res, 2, 0, (int64_t)(n * n), false, nullptr}};
    return run(sp, 2, [&](hipStream_t st) {
        e.dst = sp[0].dev;
        e.src = sp[1].dev;
        return rv::hevc_entry_launch(e, elem_blocks(n, n), st);
    });
}

// uni / bi / weighted_uni / weighted_bi (b == null: one input).
template <typename S>
int combine_common(int op, S* dst, size_t dst_len, size_t stride, const int16_t* a, size_t a_len, const int16_t* b,
                   size_t b_len, size_t w, size_t h, int shift, int log2_wd, int w0, int w1, int o0, int o1, int max) {
    const bool two = op == rv::HE_BI || op == rv::HE_WBI;
    if (!dst || !a || (two && !b)) return lib_fail(RV_E_INVALID, "null buffer");
    if (!dims_ok(w, h) || !stride_ok(stride)) return range("block or stride too large");
    if (!w || !h) return RV_OK;
    if (max < 0 || max > smax<S>()) return range("max outside the sample range");
    if (op == rv::HE_UNI && (shift < 0 || shift > 15)) return range("shift outside 0..15");
    if (op == rv::HE_BI && (shift < 1 || shift > 16)) return range("shift outside 1..16");
    if (op == rv::HE_WUNI || op == rv::HE_WBI) {
        // Ranges in which the reference's i32 arithmetic cannot overflow.
        if (log2_wd < 0 || log2_wd > 14) return range("log2_wd outside 0..14");
        if (std::abs(w0) > 4096 || std::abs(w1) > 4096) return range("weight outside -4096..4096");
        if (std::abs(o0) > 32768 || std::abs(o1) > 32768) return range("offset outside -32768..32768");
    }
    const int64_t last = block_last(0, stride, w, h);
    if (!inside(0, last, dst_len)) return range("block outside dst");
    if ((uint64_t)w * h > a_len || (two && (uint64_t)w * h > b_len)) return range("input shorter than w * h");
    rv::HevcEntry e{};
    e.op = op;
    e.ss = sizeof(S);
    e.dst_stride = (int64_t)stride;
    e.w = (int)w;
    e.h = (int)h;
    e.shift = shift;
    e.max = max;
    e.log2_wd = log2_wd;
    e.w0 = w0;
    e.w1 = w1;
    e.o0 = o0;
    e.o1 = o1;
    Span sp[3] = {{dst, sizeof(S), 0, last + 1, true, nullptr},
                  {a, 2, 0, (int64_t)(w * h), false, nullptr},
                  {two ? b : a, 2, 0, two ? (int64_t)(w * h) : 0, false, nullptr}};
    return run(sp, 3, [&](hipStream_t st) {
        e.dst = sp[0].dev;
        e.src = sp[1].dev;
        e.other = two ? (const int16_t*)sp[2].dev : nullptr;
        return rv::hevc_entry_launch(e, elem_blocks(w, h), st);
    });
}

template <typename S>
int fused(bool luma, S* dst, size_t dst_len, size_t ds, const S* src, size_t src_len, size_t ss, size_t w, size_t h,
          size_t fx, size_t fy, const int16_t* other, size_t other_len, unsigned bd, bool bi) {
    if (!dst || !src || (bi && !other)) return lib_fail(RV_E_INVALID, "null buffer");
    if (!stride_ok(ds) || !stride_ok(ss)) return range("stride too large");
    if (w > 64 || h > 64) return range("block larger than 64 x 64 (MC_TMP_LEN)");
    if (!w || !h) return RV_OK;
    const size_t nf = luma ? 4 : 8;
    if (fx >= nf || fy >= nf) return range("fraction outside the filter table");
    if (sizeof(S) == 1 ? bd != 8 : (bd < 8 || bd > 12)) return range("bit_depth outside the sample type's range");
    const int taps = luma ? 8 : 4;
    const int64_t src_last = block_last(0, ss, w + taps - 1, h + taps - 1);
    const int64_t dst_last = block_last(0, ds, w, h);
    if (!inside(0, src_last, src_len)) return range("filter window outside src");
    if (!inside(0, dst_last, dst_len)) return range("block outside dst");
    if (bi && (uint64_t)w * h > other_len) return range("other shorter than w * h");
    rv::HevcEntry e{};
    e.op = rv::HE_FUSED;
    e.ss = sizeof(S);
    e.luma = luma;
    e.dst_stride = (int64_t)ds;
    e.src_stride = (int64_t)ss;
    e.w = (int)w;
    e.h = (int)h;
    e.fx = (int)fx;
    e.fy = (int)fy;
    e.bd = (int)bd;
    Span sp[3] = {{dst, sizeof(S), 0, dst_last + 1, true, nullptr},
                  {src, sizeof(S), 0, src_last + 1, false, nullptr},
                  {bi ? (const void*)other : (const void*)src, 2, 0, bi ? (int64_t)(w * h) : 0, false, nullptr}};
    return run(sp, 3, [&](hipStream_t st) {
        e.dst = sp[0].dev;
        e.src = sp[1].dev;
        e.other = bi ? (const int16_t*)sp[2].dev : nullptr;
        return rv::hevc_entry_launch(e, 1, st);
    });
}

template <typename S>
int sao_band(S* dst, size_t dst_len, size_t ds, const S* src, size_t src_len, size_t ss, size_t w, size_t h,
             const int16_t* table, int shift, int max) {
    if (!dst || !src || !table) return lib_fail(RV_E_INVALID, "null buffer");
    if (!dims_ok(w, h) || !stride_ok(ds) || !stride_ok(ss)) return range("block or stride too large");
    if (!w || !h) return RV_OK;
    if (shift < 0 || shift > 15) return range("shift outside 0..15");
    if (max < 0 || max > smax<S>()) return range("max outside the sample range");
    const int64_t dl = block_last(0, ds, w, h), sl = block_last(0, ss, w, h);
    if (!inside(0, dl, dst_len)) return range("block outside dst");
    if (!inside(0, sl, src_len)) return range("block outside src");
    // table[v >> shift] must be inside the 32-entry table for every sample.
    for (size_t y = 0; y < h; ++y)
        for (size_t x = 0; x < w; ++x)
            if ((src[y * ss + x] >> shift) >= 32) return range("a sample's band is outside the 32-entry table");
    rv::HevcEntry e{};
    e.op = rv::HE_SAO_BAND;
    e.ss = sizeof(S);
    e.dst_stride = (int64_t)ds;
    e.src_stride = (int64_t)ss;
    e.w = (int)w;
    e.h = (int)h;
    e.shift = shift;
    e.max = max;
    std::memcpy(e.table, table, 32 * sizeof(int16_t));
    Span sp[2] = {{dst, sizeof(S), 0, dl + 1, true, nullptr}, {src, sizeof(S), 0, sl + 1, false, nullptr}};
    return run(sp, 2, [&](hipStream_t st) {
        e.dst = sp[0].dev;
        e.src = sp[1].dev;
        return rv::hevc_entry_launch(e, elem_blocks(w, h), st);
    });
}

template <typename S>
int sao_edge(S* dst, size_t dst_len, const S* src, size_t src_len, size_t origin, size_t stride, size_t w, size_t h,
             ptrdiff_t na, ptrdiff_t nb, const int16_t* off, int max) {
    if (!dst || !src || !off) return lib_fail(RV_E_INVALID, "null buffer");
    if (!dims_ok(w, h) || !stride_ok(stride) || origin > (size_t)kMaxSpan) return range("block or stride too large");
    if (std::abs((int64_t)na) > kMaxSpan || std::abs((int64_t)nb) > kMaxSpan) return range("neighbour offset too large");
    if (!w || !h) return RV_OK;
    if (max < 0 || max > smax<S>()) return range("max outside the sample range");
    const int64_t first = (int64_t)origin, last = block_last((int64_t)origin, stride, w, h);
    const int64_t lo = first + std::min<int64_t>({0, (int64_t)na, (int64_t)nb});
    const int64_t hi = last + std::max<int64_t>({0, (int64_t)na, (int64_t)nb});
    if (!inside(first, last, dst_len)) return range("region outside dst");
    if (!inside(lo, hi, src_len)) return range("region or its neighbours outside src");
    // Both spans start at element `lo` (the same rebase for the shared origin).
    rv::HevcEntry e{};
    e.op = rv::HE_SAO_EDGE;
    e.ss = sizeof(S);
    e.dst_stride = (int64_t)stride;
    e.off = first - lo;
    e.na = na;
    e.nb = nb;
    e.w = (int)w;
    e.h = (int)h;
    e.max = max;
    std::memcpy(e.table, off, 5 * sizeof(int16_t));
    Span sp[2] = {{dst, sizeof(S), lo, last + 1, true, nullptr}, {src, sizeof(S), lo, hi + 1, false, nullptr}};
    return run(sp, 2, [&](hipStream_t st) {
        e.dst = sp[0].dev;
        e.src = sp[1].dev;
        return rv::hevc_entry_launch(e, elem_blocks(w, h), st);
    });
}

// Deblocking: segments of `lines` lines, reaching `reach` samples into p
// and q. Only active segments are touched, as in the reference.
template <typename S>
int hevc_deblock(bool luma, bool vertical, S* data, size_t len, size_t off, size_t stride, const int32_t* beta,
                 const int32_t* tc, const uint8_t* no_p, const uint8_t* no_q, int max) {
    if (!data || !tc || !no_p || !no_q || (luma && !beta)) return lib_fail(RV_E_INVALID, "null argument");
    if (!stride_ok(stride) || off > (size_t)kMaxSpan) return range("offset or stride too large");
    if (max < 0 || max > smax<S>()) return range("max outside the sample range");
    const int nseg = luma ? 2 : 4, lines = luma ? 4 : 2, reach = luma ? 4 : 2;
    const int64_t step = vertical ? 1 : (int64_t)stride, along = vertical ? (int64_t)stride : 1;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int s = 0; s < nseg; ++s) {
        const bool active = luma ? !(beta[s] == 0 && tc[s] == 0) : tc[s] != 0;
        if (!active) continue;
        if (tc[s] < 0 || tc[s] > (1 << 16) || (luma && (beta[s] < 0 || beta[s] > (1 << 18))))
            return range("beta / tc outside the kernels' range");
        const int64_t first = (int64_t)off + (int64_t)s * lines * along, lastl = first + (lines - 1) * along;
        lo = std::min(lo, std::min(first, lastl) - reach * step);
        hi = std::max(hi, std::max(first, lastl) + (reach - 1) * step);
    }
    if (lo == INT64_MAX) return RV_OK;  // every segment left alone
    if (!inside(lo, hi, len)) return range("edge samples outside data");
    rv::HevcEntry e{};
    e.op = luma ? rv::HE_DBK_LUMA : rv::HE_DBK_CHROMA;
    e.ss = sizeof(S);
    e.vertical = vertical;
    e.dst_stride = (int64_t)stride;
    e.off = (int64_t)off - lo;
    e.max = max;
    for (int s = 0; s < nseg; ++s) {
        e.tc[s] = tc[s];
        if (luma) e.beta[s] = beta[s];
        e.no_p[s] = no_p[s] ? 1 : 0;
        e.no_q[s] = no_q[s] ? 1 : 0;
    }
    Span sp[1] = {{data, sizeof(S), lo, hi + 1, true, nullptr}};
    return run(sp, 1, [&](hipStream_t st) {
        e.dst = sp[0].dev;
        return rv::hevc_entry_launch(e, 1, st);
    });
}

// ---------------------------------------------------------------- H.264

constexpr int64_t PS = RV_H264_PRED_STRIDE;

int h264_run(rv::H264Entry& e, Span* sp, int n, int dst_i, int src_i, int b_i, int coef_i) {
    return run(sp, n, [&](hipStream_t st) {
        if (dst_i >= 0) e.dst = (uint8_t*)sp[dst_i].dev;
        if (src_i >= 0) e.src = (const uint8_t*)sp[src_i].dev;
        if (b_i >= 0) e.b = (const uint8_t*)sp[b_i].dev;
        if (coef_i >= 0) e.coefs = sp[coef_i].dev;
        return rv::h264_entry_launch(e, 1, st);
    });
}

bool pred_dims(size_t w, size_t h) { return w >= 1 && h >= 1 && w <= 16 && h <= 16; }

int h264_weights_ok(int log_wd, int w0, int w1, int o0, int o1, int max) {
    if (max < 1 || max > 255) return range("max outside 1..255");
    if (log_wd < 0 || log_wd > 14) return range("log_wd outside 0..14");
    if (std::abs(w0) > 4096 || std::abs(w1) > 4096) return range("weight outside -4096..4096");
    if (std::abs(o0) > 32768 || std::abs(o1) > 32768) return range("offset outside -32768..32768");
    return RV_OK;
}

// Blocks written into a plane: a w x h block at the start of dst.
int plane_block(size_t dst_len, size_t stride, size_t w, size_t h, int64_t* last) {
    if (!stride_ok(stride)) return range("stride too large");
    *last = block_last(0, stride, w, h);
    if (!inside(0, *last, dst_len)) return range("block outside dst");
    return RV_OK;
}

int h264_deblock(bool chroma, bool vertical, bool intra, uint8_t* data, size_t len, size_t off, size_t stride,
                 int alpha, int beta, const int16_t* tc0, int max) {
    if (!data || (!intra && !tc0)) return lib_fail(RV_E_INVALID, "null argument");
    if (!stride_ok(stride) || off > (size_t)kMaxSpan) return range("offset or stride too large");
    if (max < 1 || max > 255) return range("max outside 1..255");
    if (alpha < 0 || alpha > (1 << 16) || beta < 0 || beta > (1 << 16)) return range("alpha / beta outside 0..65536");
    const int lines = chroma ? 8 : 16, per = chroma ? 2 : 4, reach = chroma ? 2 : 4;
    const int64_t step = vertical ? (int64_t)stride : 1, across = vertical ? 1 : (int64_t)stride;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int i = 0; i < lines; ++i) {
        if (!intra && tc0[i / per] < 0) continue;
        if (!intra && tc0[i / per] > (1 << 14)) return range("tc0 outside the kernels' range");
        const int64_t base = (int64_t)off + i * step;
        lo = std::min(lo, base - reach * across);
        hi = std::max(hi, base + (reach - 1) * across);
    }
    if (lo == INT64_MAX) return RV_OK;
    if (!inside(lo, hi, len)) return range("edge samples outside data");
    rv::H264Entry e{};
    e.op = rv::HA_DBK;
    e.chroma = chroma;
    e.vertical = vertical;
    e.intra = intra;
    e.stride = (int64_t)stride;
    e.off = (int64_t)off - lo;
    e.alpha = alpha;
    e.beta = beta;
    e.max = max;
    if (!intra) std::memcpy(e.tc0, tc0, sizeof(e.tc0));
    Span sp[1] = {{data, 1, lo, hi + 1, true, nullptr}};
    return h264_run(e, sp, 1, 0, -1, -1, -1);
}

// idct4/8_add, dc_add, residual4/8 into an n x n block of dst.
int h264_block(int op, int n, uint8_t* dst, size_t dst_len, size_t stride, const void* coefs, size_t coef_elem, int dc,
               int has_dc, int max) {
    if (!dst || (op != rv::HA_DC && !coefs)) return lib_fail(RV_E_INVALID, "null argument");
    if (max < 1 || max > 255) return range("max outside 1..255");
    int64_t last;
    if (int rc = plane_block(dst_len, stride, n, n, &last)) return rc;
    rv::H264Entry e{};
    e.op = op;
    e.stride = (int64_t)stride;
    e.n = n;
    e.dc = dc;
    e.has_dc = has_dc;
    e.max = max;
    Span sp[2] = {{dst, 1, 0, last + 1, true, nullptr},
                  {coefs ? coefs : (const void*)dst, coef_elem, 0, coefs ? (int64_t)n * n : 0, false, nullptr}};
    return h264_run(e, sp, 2, 0, -1, -1, coefs ? 1 : -1);
}

}  // namespace

extern "C" {

int rv_hevc_idct(int log2, int16_t* coeffs, size_t coeffs_len, int bd_shift, size_t max_x, size_t max_y) {
    return idct_common(rv::HE_IDCT, log2, coeffs, coeffs_len, bd_shift, max_x, max_y);
}
int rv_hevc_idst4(int16_t* coeffs, size_t coeffs_len, int bd_shift, size_t max_x, size_t max_y) {
    return idct_common(rv::HE_IDST, 2, coeffs, coeffs_len, bd_shift, max_x, max_y);
}
int rv_hevc_qpel_v2(int16_t* dst, size_t dst_len, const int16_t* src, size_t src_len, size_t ss, size_t w, size_t h,
                    size_t frac) {
    return filt_common<int16_t>(rv::HE_FILT_V2, true, dst, dst_len, src, src_len, ss, w, h, frac, 6, true);
}
int rv_hevc_epel_v2(int16_t* dst, size_t dst_len, const int16_t* src, size_t src_len, size_t ss, size_t w, size_t h,
                    size_t frac) {
    return filt_common<int16_t>(rv::HE_FILT_V2, false, dst, dst_len, src, src_len, ss, w, h, frac, 6, true);
}

#define RV_DEFINE_HEVC_SAMPLE_ENTRIES(S, sfx)                                                                        \
    int rv_hevc_##sfx##_add_residual(S* dst, size_t dst_len, size_t stride, const int16_t* res, size_t res_len,     \
                                     size_t n, int max) {                                                           \
        return add_residual<S>(dst, dst_len, stride, res, res_len, n, max);                                         \
    }                                                                                                               \
    int rv_hevc_##sfx##_qpel_copy(int16_t* dst, size_t dst_len, const S* src, size_t src_len, size_t ss, size_t w,  \
                                  size_t h, int shift) {                                                            \
        return filt_common<S>(rv::HE_COPY, true, dst, dst_len, src, src_len, ss, w, h, 0, shift, false);            \
    }                                                                                                               \
    int rv_hevc_##sfx##_qpel_h(int16_t* dst, size_t dst_len, const S* src, size_t src_len, size_t ss, size_t w,     \
                               size_t h, size_t frac, int shift) {                                                  \
        return filt_common<S>(rv::HE_FILT_H, true, dst, dst_len, src, src_len, ss, w, h, frac, shift, false);       \
    }                                                                                                               \
    int rv_hevc_##sfx##_qpel_v(int16_t* dst, size_t dst_len, const S* src, size_t src_len, size_t ss, size_t w,     \
                               size_t h, size_t frac, int shift) {                                                  \
        return filt_common<S>(rv::HE_FILT_V, true, dst, dst_len, src, src_len, ss, w, h, frac, shift, false);       \
    }                                                                                                               \
    int rv_hevc_##sfx##_epel_copy(int16_t* dst, size_t dst_len, const S* src, size_t src_len, size_t ss, size_t w,  \
                                  size_t h, int shift) {                                                            \
        return filt_common<S>(rv::HE_COPY, false, dst, dst_len, src, src_len, ss, w, h, 0, shift, false);           \
    }                                                                                                               \
    int rv_hevc_##sfx##_epel_h(int16_t* dst, size_t dst_len, const S* src, size_t src_len, size_t ss, size_t w,     \
                               size_t h, size_t frac, int shift) {                                                  \
        return filt_common<S>(rv::HE_FILT_H, false, dst, dst_len, src, src_len, ss, w, h, frac, shift, false);      \
    }                                                                                                               \
    int rv_hevc_##sfx##_epel_v(int16_t* dst, size_t dst_len, const S* src, size_t src_len, size_t ss, size_t w,     \
                               size_t h, size_t frac, int shift) {                                                  \
        return filt_common<S>(rv::HE_FILT_V, false, dst, dst_len, src, src_len, ss, w, h, frac, shift, false);      \
    }                                                                                                               \
    int rv_hevc_##sfx##_uni(S* dst, size_t dst_len, size_t stride, const int16_t* src, size_t src_len, size_t w,    \
                            size_t h, int shift, int max) {                                                         \
        return combine_common<S>(rv::HE_UNI, dst, dst_len, stride, src, src_len, nullptr, 0, w, h, shift, 0, 0, 0, \
                                 0, 0, max);                                                                        \
    }                                                                                                               \
    int rv_hevc_##sfx##_bi(S* dst, size_t dst_len, size_t stride, const int16_t* a, size_t a_len, const int16_t* b, \
                           size_t b_len, size_t w, size_t h, int shift, int max) {                                  \
        return combine_common<S>(rv::HE_BI, dst, dst_len, stride, a, a_len, b, b_len, w, h, shift, 0, 0, 0, 0, 0,  \
                                 max);                                                                              \
    }                                                                                                               \
    int rv_hevc_##sfx##_weighted_uni(S* dst, size_t dst_len, size_t stride, const int16_t* src, size_t src_len,     \
                                     size_t w, size_t h, int log2_wd, int wt, int o, int max) {                     \
        return combine_common<S>(rv::HE_WUNI, dst, dst_len, stride, src, src_len, nullptr, 0, w, h, 0, log2_wd, wt, \
                                 0, o, 0, max);                                                                     \
    }                                                                                                               \
    int rv_hevc_##sfx##_weighted_bi(S* dst, size_t dst_len, size_t stride, const int16_t* a, size_t a_len,          \
                                    const int16_t* b, size_t b_len, size_t w, size_t h, int log2_wd, int w0,        \
                                    int w1, int o0, int o1, int max) {                                              \
        return combine_common<S>(rv::HE_WBI, dst, dst_len, stride, a, a_len, b, b_len, w, h, 0, log2_wd, w0, w1,   \
                                 o0, o1, max);                                                                      \
    }                                                                                                               \
    int rv_hevc_##sfx##_qpel_uni(S* dst, size_t dst_len, size_t ds, const S* src, size_t src_len, size_t ss,        \
                                 size_t w, size_t h, size_t fx, size_t fy, unsigned bd) {                           \
        return fused<S>(true, dst, dst_len, ds, src, src_len, ss, w, h, fx, fy, nullptr, 0, bd, false);             \
    }                                                                                                               \
    int rv_hevc_##sfx##_epel_uni(S* dst, size_t dst_len, size_t ds, const S* src, size_t src_len, size_t ss,        \
                                 size_t w, size_t h, size_t fx, size_t fy, unsigned bd) {                           \
        return fused<S>(false, dst, dst_len, ds, src, src_len, ss, w, h, fx, fy, nullptr, 0, bd, false);            \
    }                                                                                                               \
    int rv_hevc_##sfx##_qpel_bi(S* dst, size_t dst_len, size_t ds, const S* src, size_t src_len, size_t ss,         \
                                size_t w, size_t h, size_t fx, size_t fy, const int16_t* other, size_t other_len,   \
                                unsigned bd) {                                                                      \
        return fused<S>(true, dst, dst_len, ds, src, src_len, ss, w, h, fx, fy, other, other_len, bd, true);        \
    }                                                                                                               \
    int rv_hevc_##sfx##_epel_bi(S* dst, size_t dst_len, size_t ds, const S* src, size_t src_len, size_t ss,         \
                                size_t w, size_t h, size_t fx, size_t fy, const int16_t* other, size_t other_len,   \
                                unsigned bd) {                                                                      \
        return fused<S>(false, dst, dst_len, ds, src, src_len, ss, w, h, fx, fy, other, other_len, bd, true);       \
    }                                                                                                               \
    int rv_hevc_##sfx##_sao_band(S* dst, size_t dst_len, size_t ds, const S* src, size_t src_len, size_t ss,        \
                                 size_t w, size_t h, const int16_t table[32], int shift, int max) {                 \
        return sao_band<S>(dst, dst_len, ds, src, src_len, ss, w, h, table, shift, max);                            \
    }                                                                                                               \
    int rv_hevc_##sfx##_sao_edge(S* dst, size_t dst_len, const S* src, size_t src_len, size_t origin,               \
                                 size_t stride, size_t w, size_t h, ptrdiff_t na, ptrdiff_t nb,                     \
                                 const int16_t off[5], int max) {                                                   \
        return sao_edge<S>(dst, dst_len, src, src_len, origin, stride, w, h, na, nb, off, max);                     \
    }                                                                                                               \
    int rv_hevc_##sfx##_deblock_luma_v(S* data, size_t data_len, size_t off, size_t stride, const int32_t beta[2], \
                                       const int32_t tc[2], const uint8_t no_p[2], const uint8_t no_q[2],           \
                                       int max) {                                                                   \
        return hevc_deblock<S>(true, true, data, data_len, off, stride, beta, tc, no_p, no_q, max);                 \
    }                                                                                                               \
    int rv_hevc_##sfx##_deblock_luma_h(S* data, size_t data_len, size_t off, size_t stride, const int32_t beta[2], \
                                       const int32_t tc[2], const uint8_t no_p[2], const uint8_t no_q[2],           \
                                       int max) {                                                                   \
        return hevc_deblock<S>(true, false, data, data_len, off, stride, beta, tc, no_p, no_q, max);                \
    }                                                                                                               \
    int rv_hevc_##sfx##_deblock_chroma_v(S* data, size_t data_len, size_t off, size_t stride, const int32_t tc[4],  \
                                         const uint8_t no_p[4], const uint8_t no_q[4], int max) {                   \
        return hevc_deblock<S>(false, true, data, data_len, off, stride, nullptr, tc, no_p, no_q, max);             \
    }                                                                                                               \
    int rv_hevc_##sfx##_deblock_chroma_h(S* data, size_t data_len, size_t off, size_t stride, const int32_t tc[4],  \
                                         const uint8_t no_p[4], const uint8_t no_q[4], int max) {                   \
        return hevc_deblock<S>(false, false, data, data_len, off, stride, nullptr, tc, no_p, no_q, max);            \
    }

RV_DEFINE_HEVC_SAMPLE_ENTRIES(uint8_t, u8)
RV_DEFINE_HEVC_SAMPLE_ENTRIES(uint16_t, u16)

int rv_h264_u8_qpel(int pos, uint8_t* dst, size_t dst_len, const uint8_t* src, size_t src_len, size_t ss, size_t w,
                    size_t h, int max) {
    if (!dst || !src) return lib_fail(RV_E_INVALID, "null buffer");
    if (pos < 0 || pos > 15) return range("pos outside 0..15");
    if (!pred_dims(w, h)) return range("block outside 1..16 x 1..16");
    if (max < 1 || max > 255) return range("max outside 1..255");
    if (!stride_ok(ss)) return range("stride too large");
    const int64_t sl = block_last(0, ss, w + 5, h + 5), dl = block_last(0, PS, w, h);
    if (!inside(0, sl, src_len)) return range("six-tap window outside src");
    if (!inside(0, dl, dst_len)) return range("block outside dst");
    rv::H264Entry e{};
    e.op = rv::HA_QPEL;
    e.pos = pos;
    e.src_stride = (int64_t)ss;
    e.w = (int)w;
    e.h = (int)h;
    e.max = max;
    Span sp[2] = {{dst, 1, 0, dl + 1, true, nullptr}, {src, 1, 0, sl + 1, false, nullptr}};
    return h264_run(e, sp, 2, 0, 1, -1, -1);
}

int rv_h264_u8_chroma(uint8_t* dst, size_t dst_len, const uint8_t* src, size_t src_len, size_t stride, size_t w,
                      size_t h, int xf, int yf) {
    if (!dst || !src) return lib_fail(RV_E_INVALID, "null buffer");
    if (xf < 0 || xf > 7 || yf < 0 || yf > 7) return range("fraction outside 0..7");
    if (!pred_dims(w, h)) return range("block outside 1..16 x 1..16");
    if (!stride_ok(stride)) return range("stride too large");
    const int64_t sl = block_last(0, stride, w + 1, h + 1), dl = block_last(0, PS, w, h);
    if (!inside(0, sl, src_len)) return range("bilinear window outside src");
    if (!inside(0, dl, dst_len)) return range("block outside dst");
    rv::H264Entry e{};
    e.op = rv::HA_CHROMA;
    e.src_stride = (int64_t)stride;
    e.w = (int)w;
    e.h = (int)h;
    e.xf = xf;
    e.yf = yf;
    Span sp[2] = {{dst, 1, 0, dl + 1, true, nullptr}, {src, 1, 0, sl + 1, false, nullptr}};
    return h264_run(e, sp, 2, 0, 1, -1, -1);
}

namespace {
// copy / avg / weighted_uni / weighted_bi: PRED_STRIDE inputs into a plane.
int h264_combine(int op, uint8_t* dst, size_t dst_len, size_t stride, const uint8_t* a, size_t a_len, const uint8_t* b,
                 size_t b_len, size_t w, size_t h, int log_wd, int w0, int w1, int o0, int o1, int max) {
    const bool two = op == rv::HA_AVG || op == rv::HA_WBI;
    if (!dst || !a || (two && !b)) return lib_fail(RV_E_INVALID, "null buffer");
    if (!pred_dims(w, h)) return range("block outside 1..16 x 1..16");
    if (op == rv::HA_WUNI || op == rv::HA_WBI)
        if (int rc = h264_weights_ok(log_wd, w0, w1, o0, o1, max)) return rc;
    int64_t dl;
    if (int rc = plane_block(dst_len, stride, w, h, &dl)) return rc;
    const int64_t il = block_last(0, PS, w, h);
    if (!inside(0, il, a_len) || (two && !inside(0, il, b_len))) return range("prediction block outside its slice");
    rv::H264Entry e{};
    e.op = op;
    e.stride = (int64_t)stride;
    e.w = (int)w;
    e.h = (int)h;
    e.log_wd = log_wd;
    e.w0 = w0;
    e.w1 = w1;
    e.o0 = o0;
    e.o1 = o1;
    e.max = max;
    Span sp[3] = {{dst, 1, 0, dl + 1, true, nullptr},
                  {a, 1, 0, il + 1, false, nullptr},
                  {two ? b : a, 1, 0, two ? il + 1 : 0, false, nullptr}};
    return h264_run(e, sp, 3, 0, 1, two ? 2 : -1, -1);
}
}  // namespace

int rv_h264_u8_copy(uint8_t* dst, size_t dst_len, size_t stride, const uint8_t* src, size_t src_len, size_t w,
                    size_t h) {
    return h264_combine(rv::HA_COPY, dst, dst_len, stride, src, src_len, nullptr, 0, w, h, 0, 0, 0, 0, 0, 255);
}
int rv_h264_u8_avg(uint8_t* dst, size_t dst_len, size_t stride, const uint8_t* a, size_t a_len, const uint8_t* b,
                   size_t b_len, size_t w, size_t h) {
    return h264_combine(rv::HA_AVG, dst, dst_len, stride, a, a_len, b, b_len, w, h, 0, 0, 0, 0, 0, 255);
}
int rv_h264_u8_weighted_uni(uint8_t* dst, size_t dst_len, size_t stride, const uint8_t* src, size_t src_len, size_t w,
                            size_t h, int log_wd, int wt, int o, int max) {
    return h264_combine(rv::HA_WUNI, dst, dst_len, stride, src, src_len, nullptr, 0, w, h, log_wd, wt, 0, o, 0, max);
}
int rv_h264_u8_weighted_bi(uint8_t* dst, size_t dst_len, size_t stride, const uint8_t* a, size_t a_len,
                           const uint8_t* b, size_t b_len, size_t w, size_t h, int log_wd, int w0, int w1, int o0,
                           int o1, int max) {
    return h264_combine(rv::HA_WBI, dst, dst_len, stride, a, a_len, b, b_len, w, h, log_wd, w0, w1, o0, o1, max);
}

int rv_h264_u8_deblock_luma_v(uint8_t* data, size_t len, size_t off, size_t stride, int alpha, int beta,
                              const int16_t tc0[4], int max) {
    return h264_deblock(false, true, false, data, len, off, stride, alpha, beta, tc0, max);
}
int rv_h264_u8_deblock_luma_h(uint8_t* data, size_t len, size_t off, size_t stride, int alpha, int beta,
                              const int16_t tc0[4], int max) {
    return h264_deblock(false, false, false, data, len, off, stride, alpha, beta, tc0, max);
}
int rv_h264_u8_deblock_luma_v_intra(uint8_t* data, size_t len, size_t off, size_t stride, int alpha, int beta,
                                    int max) {
    return h264_deblock(false, true, true, data, len, off, stride, alpha, beta, nullptr, max);
}
int rv_h264_u8_deblock_luma_h_intra(uint8_t* data, size_t len, size_t off, size_t stride, int alpha, int beta,
                                    int max) {
    return h264_deblock(false, false, true, data, len, off, stride, alpha, beta, nullptr, max);
}
int rv_h264_u8_deblock_chroma_v(uint8_t* data, size_t len, size_t off, size_t stride, int alpha, int beta,
                                const int16_t tc0[4], int max) {
    return h264_deblock(true, true, false, data, len, off, stride, alpha, beta, tc0, max);
}
int rv_h264_u8_deblock_chroma_h(uint8_t* data, size_t len, size_t off, size_t stride, int alpha, int beta,
                                const int16_t tc0[4], int max) {
    return h264_deblock(true, false, false, data, len, off, stride, alpha, beta, tc0, max);
}
int rv_h264_u8_deblock_chroma_v_intra(uint8_t* data, size_t len, size_t off, size_t stride, int alpha, int beta,
                                      int max) {
    return h264_deblock(true, true, true, data, len, off, stride, alpha, beta, nullptr, max);
}
int rv_h264_u8_deblock_chroma_h_intra(uint8_t* data, size_t len, size_t off, size_t stride, int alpha, int beta,
                                      int max) {
    return h264_deblock(true, false, true, data, len, off, stride, alpha, beta, nullptr, max);
}

int rv_h264_u8_idct4_add(uint8_t* dst, size_t dst_len, size_t stride, const int16_t coeffs[16], int max) {
    return h264_block(rv::HA_IDCT4, 4, dst, dst_len, stride, coeffs, 2, 0, 0, max);
}
int rv_h264_u8_idct8_add(uint8_t* dst, size_t dst_len, size_t stride, const int16_t coeffs[64], int max) {
    return h264_block(rv::HA_IDCT8, 8, dst, dst_len, stride, coeffs, 2, 0, 0, max);
}
int rv_h264_u8_idct4_dc_add(uint8_t* dst, size_t dst_len, size_t stride, int dc, int max) {
    return h264_block(rv::HA_DC, 4, dst, dst_len, stride, nullptr, 1, dc, 0, max);
}
int rv_h264_u8_idct8_dc_add(uint8_t* dst, size_t dst_len, size_t stride, int dc, int max) {
    return h264_block(rv::HA_DC, 8, dst, dst_len, stride, nullptr, 1, dc, 0, max);
}
int rv_h264_u8_residual4(uint8_t* dst, size_t dst_len, size_t stride, const int32_t coefs[16], int dc, int max) {
    const bool has_dc = dc != RV_H264_NO_DC;
    return h264_block(rv::HA_RES4, 4, dst, dst_len, stride, coefs, 4, has_dc ? dc : 0, has_dc, max);
}
int rv_h264_u8_residual8(uint8_t* dst, size_t dst_len, size_t stride, const int32_t coefs[64], int max) {
    return h264_block(rv::HA_RES8, 8, dst, dst_len, stride, coefs, 4, 0, 0, max);
}

}  // extern "C"
