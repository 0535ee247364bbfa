// Host runtime of the HIP tier: device selection, sessions (DPB planes
// resident in device memory), job preparation (work lists validated and
// copied to the device), batches (per-stage item tables, hipGraph replay)
// and completion markers. Every extern "C" entry catches everything and
// returns an RV_E_* code; messages go to a thread-local buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rivet_hip.h"
#include "rv_device.h"
#include "rv_kernels.h"
#include "rv_runtime.h"

namespace {

thread_local std::string g_err;
std::mutex g_mu;
int g_device = -1;
hipStream_t g_stream = nullptr;
hipStream_t g_io = nullptr;  // plane uploads / downloads
hipStream_t g_up = nullptr;  // work-list and batch-table uploads
int g_stage_mask = 0;
// Device error counter: kernels whose bounded waits give up add to it
// (H264PicDev::err). The host reports each increase once, as RV_E_DEVICE,
// from the next synchronising call that reads it.
uint32_t* g_dev_err = nullptr;
uint32_t g_err_seen = 0;
std::mutex g_err_mu;
constexpr int kMarks = 4096;
hipEvent_t g_marks[kMarks];
int g_next_mark = 0;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
int hip_fail(hipError_t e, const char* what) {
    return fail(RV_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define RV_HIP(call)                                  \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

bool ready() { return g_device >= 0 && g_stream; }

// RV_E_DEVICE when the device error counter rose since the last check.
// Called after a wait, so the launches it covers have completed.
int check_device_errors() {
    if (!g_dev_err) return RV_OK;
    uint32_t n = 0;
    hipError_t e = hipMemcpy(&n, g_dev_err, sizeof(n), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "device error counter");
    std::lock_guard<std::mutex> lk(g_err_mu);
    if (n == g_err_seen) return RV_OK;
    const uint32_t k = n - g_err_seen;
    g_err_seen = n;
    return fail(RV_E_DEVICE, std::to_string(k) + " H.264 deblocking exchange wait(s) gave up (a band's strips "
                             "never arrived); the pictures launched since the last check are not valid");
}

}  // namespace

// Device memory of rv_picture_hash_noalloc: the hash kernels' results
// (256 bytes) and one CRC partial per row and component, for pictures of at
// most RV_HASH_NOALLOC_MAX_ROWS rows. Each device holds its own copy from
// module load to process end. Host code never passes this name to a
// kernel (on the host it is a stand-in, not a device address): rv_init
// looks up the device address once, and the entry point uses that.
static __device__ __align__(256) uint8_t scratch[256 + 3 * RV_HASH_NOALLOC_MAX_ROWS * sizeof(uint32_t)];

// Small tables written by a kernel whose data travel in its arguments, so
// that writing them is ordered on the library stream and the host never
// waits for a copy: a host-to-device copy of a few hundred bytes runs as a
// blit kernel that can start only once the stage kernels ahead of it have
// drained, and a host that waited for it could not queue the next batch.
constexpr int kFillWords = 960;
struct WordChunk {
    uint32_t* dst;
    int32_t n;
    uint32_t w[kFillWords];
};
__global__ void __launch_bounds__(256) rv_fill_words(WordChunk c) {
    for (int i = threadIdx.x; i < c.n; i += 256) c.dst[i] = c.w[i];
}

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_dev[kMaxDevices];  // device address of `scratch`, per device ordinal
std::mutex g_noalloc_mu;           // one allocation-free hash at a time uses the array
}  // namespace

namespace rv {
std::mutex& lib_mutex() { return g_mu; }
hipStream_t lib_stream() { return g_stream; }
bool lib_ready() { return ready(); }
int lib_fail(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace rv

namespace {

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Runs f when the scope ends, on every return path.
template <typename F> struct OnExit {
    F f;
    ~OnExit() { f(); }
};
template <typename F> OnExit<F> on_exit(F f) { return OnExit<F>{f}; }

// Bytes after each DPB plane: motion compensation reads reference rows
// with aligned dword loads that may extend up to 4 bytes past a row's last
// sample (past the plane's end on its last row).
constexpr size_t kPlaneSlack = 256;

struct PlaneGeom {
    int pw[3], ph[3], pitch[3];
    size_t bytes[3];
    int ss;
};

// ---- Device buffers recycled by completion ----------------------------
// Job arenas, job scratch and batch tables come from one pool. A released
// buffer goes back to the pool once an event recorded on the kernel stream
// at release has completed, so a free never waits for the GPU and, once
// the pool has warmed up, a prepare allocates nothing.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};
// An event recorded on the kernel stream after a launch (or other stream
// work on pooled buffers), shared by the sessions and released buffers it
// covers; its event is recycled when the last of them lets go.
std::mutex g_launch_mu;
std::vector<hipEvent_t> g_launch_free;  // under g_launch_mu
struct LaunchEv {
    hipEvent_t e = nullptr;
    ~LaunchEv() {
        if (!e) return;
        std::lock_guard<std::mutex> lk(g_launch_mu);
        g_launch_free.push_back(e);
    }
};

std::mutex g_pool_mu;
std::vector<DevBuf> g_pool;
// Released buffers, reusable once the event of the last kernel-stream work
// enqueued before their release has completed. Every device use of a
// pooled buffer is kernel-stream work followed by such an event (launches,
// batch-table fills), so one event per launch covers every buffer released
// after it: releases cost no stream packet of their own.
std::vector<std::pair<std::shared_ptr<LaunchEv>, DevBuf>> g_pool_pending;
std::shared_ptr<LaunchEv> g_last_work;  // under g_pool_mu

// Completed releases back to the pool; the caller holds g_pool_mu.
void pool_reap() {
    for (size_t i = 0; i < g_pool_pending.size();) {
        if (hipEventQuery(g_pool_pending[i].first->e) == hipSuccess) {
            g_pool.push_back(g_pool_pending[i].second);
            g_pool_pending[i] = g_pool_pending.back();
            g_pool_pending.pop_back();
        } else {
            ++i;
        }
    }
}

// The smallest free buffer of at least `bytes`, else a new one with a
// quarter of headroom.
bool pool_get(size_t bytes, DevBuf* out) {
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        pool_reap();
        int best = -1;
        for (int i = 0; i < (int)g_pool.size(); ++i)
            if (g_pool[i].cap >= bytes && (best < 0 || g_pool[i].cap < g_pool[best].cap)) best = i;
        if (best >= 0) {
            *out = g_pool[best];
            g_pool[best] = g_pool.back();
            g_pool.pop_back();
            return true;
        }
    }
    const size_t cap = bytes + bytes / 4 + 4096;
    void* p = nullptr;
    if (hipMalloc(&p, cap) != hipSuccess) return false;
    *out = DevBuf{p, cap};
    return true;
}

// Release after everything launched so far on the kernel stream. Callers
// hold g_mu, so the event never lands in another thread's graph capture.
void pool_put(const DevBuf& b) {
    if (!b.p) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (g_last_work) g_pool_pending.emplace_back(g_last_work, b);
    else g_pool.push_back(b);  // no kernel-stream work has touched it yet
}

// Records the event that covers the kernel-stream work enqueued so far
// (the caller holds g_mu, so nothing else is enqueued meanwhile).
std::shared_ptr<LaunchEv> note_stream_work(hipError_t* err) {
    auto ev = std::make_shared<LaunchEv>();
    {
        std::lock_guard<std::mutex> lk(g_launch_mu);
        if (!g_launch_free.empty()) {
            ev->e = g_launch_free.back();
            g_launch_free.pop_back();
        }
    }
    hipError_t e = ev->e ? hipSuccess : hipEventCreateWithFlags(&ev->e, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(ev->e, g_stream);
    *err = e;
    if (e != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_last_work = ev;
    return ev;
}

// Pinned host staging of the calling thread (work lists and batch tables
// are assembled here and copied with one DMA transfer).
struct Staging {
    char* p = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    char* get(size_t n) {
        if (n > cap) {
            if (p) (void)hipHostFree(p);
            p = nullptr;
            cap = 0;
            const size_t c = n + n / 4 + 4096;
            if (hipHostMalloc((void**)&p, c, hipHostMallocDefault) != hipSuccess) return nullptr;
            cap = c;
        }
        return p;
    }
    ~Staging() {
        if (p) (void)hipHostFree(p);
        if (ev) (void)hipEventDestroy(ev);
    }
};
thread_local Staging t_stage;

// Copy `bytes` of the staging buffer to `dst` on the upload stream, zero
// `zero_bytes` at `zero`, and wait for both.
hipError_t stage_upload(void* dst, size_t bytes, void* zero, size_t zero_bytes) {
    hipError_t e = hipMemcpyAsync(dst, t_stage.p, bytes, hipMemcpyHostToDevice, g_up);
    if (e == hipSuccess && zero_bytes) e = hipMemsetAsync(zero, 0, zero_bytes, g_up);
    if (e == hipSuccess && !t_stage.ev) e = hipEventCreateWithFlags(&t_stage.ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(t_stage.ev, g_up);
    if (e == hipSuccess) e = hipEventSynchronize(t_stage.ev);
    return e;
}


PlaneGeom plane_geom(int w, int h, int bd) {
    PlaneGeom g;
    g.ss = bd > 8 ? 2 : 1;
    for (int c = 0; c < 3; ++c) {
        g.pw[c] = c ? w / 2 : w;
        g.ph[c] = c ? h / 2 : h;
        size_t row = align_up((size_t)g.pw[c] * g.ss, 256);
        g.pitch[c] = (int)(row / g.ss);
        g.bytes[c] = row * g.ph[c];
    }
    return g;
}

}  // namespace

struct rv_session {
    int codec, w, h, bd, slots;
    PlaneGeom g;
    std::vector<void*> planes;  // slots * 3
    std::vector<DevBuf> bufs;   // the planes' pool buffers
    // Completion of the latest launch that included a job of this session
    // (null before the first): plane copies wait for this, not for the
    // other streams' work. Set and read under g_mu.
    std::shared_ptr<LaunchEv> last;
};

struct rv_job {
    rv_session* s = nullptr;
    int codec = 0;
    int out_slot = 0;
    DevBuf arena;    // work lists + descriptor
    DevBuf scratch;  // reconstruction planes + side maps
    rv::HevcPicDev hd{};
    rv::H264PicDev h4{};
    void* desc_dev = nullptr;  // device copy of the descriptor
    // Single-picture tables for the plain path.
    rv::BatchDev single{};
    // Device intra: HEVC TUs or H.264 macroblocks per level (host copy),
    // and their total.
    std::vector<int32_t> itotal;
    int32_t n_itus = 0;
    int64_t bytes[4] = {0, 0, 0, 0};
};

struct rv_batch {
    std::vector<rv_job*> jobs;
    std::vector<rv_job*> hevc, h264;
    int depth_cls = 0;               // HEVC pictures: 1 all 8-bit, 2 all deeper, 0 mixed
    int var[rv::kStages] = {0, 0, 0, 0, 0};  // HEVC kernel variant per stage
    bool var_set = false;            // chosen by rv_batch_set_variant / rv_batch_tune
    rv::BatchDev bh{}, b4{};
    std::vector<int32_t> pre_h, pre_4;  // host copies (totals)
    std::vector<int32_t> itotal;        // device-intra TUs per level (HEVC)
    std::vector<int32_t> itotal4;       // device-intra macroblocks per level (H.264)
    DevBuf tables;
    DevBuf itask;                       // device-intra task table
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
};

namespace {

template <typename T> T* carve(char*& p, size_t n) {
    T* r = (T*)p;
    p += align_up(n * sizeof(T), 256);
    return r;
}
template <typename T> size_t carve_size(size_t n) { return align_up(n * sizeof(T), 256); }

// Record the end of a launch for the sessions of `jobs`; the caller holds
// g_mu and has just enqueued the launch on g_stream.
int note_launch(const std::vector<rv_job*>& jobs) {
    hipError_t e = hipSuccess;
    auto ev = note_stream_work(&e);
    if (!ev) return hip_fail(e, "launch event");
    for (rv_job* j : jobs) j->s->last = ev;
    return RV_OK;
}

// Wait on the host for the latest launch that covered `s`.
hipError_t wait_session(rv_session* s) {
    std::shared_ptr<LaunchEv> ev;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        ev = s->last;
    }
    return ev ? hipEventSynchronize(ev->e) : hipSuccess;
}

// Tuned HEVC kernel variants per depth class (1 all 8-bit, 2 all deeper,
// 0 mixed), shared by every batch of that class; under g_mu.
int g_tuned_var[3][rv::kStages];
bool g_tuned[3] = {false, false, false};

bool tuning_enabled() {
    const char* e = std::getenv("RIVET_HIP_TUNE");
    return !(e && e[0] == '0');
}

// `var`: the HEVC variant of each stage (null: variant 0).
int launch_codec(const rv::BatchDev& b, const std::vector<int32_t>& pre, int n, bool hevc, hipStream_t st,
                 const int* var = nullptr) {
    for (int s = 0; s < rv::kStages; ++s) {
        int total = pre[(size_t)s * (n + 1) + n];
        hipError_t e = hevc ? rv::hevc_launch_stage(s, b, total, st, var ? var[s] : 0)
                            : rv::h264_launch_stage(s, b, total, st);
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
    }
    return RV_OK;
}

int depth_class_of(int bd) { return bd == 8 ? 1 : 2; }

// Return a job's buffers to the pool (after the work launched so far).
void job_release(rv_job* j) {
    std::lock_guard<std::mutex> lk(g_mu);
    pool_put(j->arena);
    pool_put(j->scratch);
    j->arena = DevBuf{};
    j->scratch = DevBuf{};
}


// ---- HEVC job preparation -------------------------------------------
// cu_at: the CU of each 4x4 luma block, filled (and returned non-empty)
// only when some CU is predicted on the device.
int validate_hevc(const rv_session* s, const rv_hevc_pic* p, std::vector<int32_t>& cu_at) {
    if (p->width != s->w || p->height != s->h || p->bit_depth != s->bd)
        return fail(RV_E_RANGE, "picture shape differs from the session");
    if (p->log2_ctb < 4 || p->log2_ctb > 6) return fail(RV_E_RANGE, "log2_ctb outside 4..6");
    for (int l = 0; l < 2; ++l) {
        if (p->num_ref[l] < 0 || p->num_ref[l] > RV_MAX_REFS) return fail(RV_E_RANGE, "num_ref outside 0..4");
        for (int r = 0; r < p->num_ref[l]; ++r)
            if (p->ref_slot[l][r] < 0 || p->ref_slot[l][r] >= s->slots) return fail(RV_E_RANGE, "ref_slot outside the DPB");
    }
    if (p->n_cus < 0 || p->n_pus < 0 || p->n_tus < 0) return fail(RV_E_INVALID, "negative list length");
    if ((p->n_cus && !p->cus) || (p->n_pus && !p->pus) || (p->n_tus && !p->tus) || (p->n_nz && !p->nz) ||
        (p->n_intra_pred && !p->intra_pred))
        return fail(RV_E_INVALID, "null work list");
    if ((p->sao_luma || p->sao_chroma) && !p->sao) return fail(RV_E_INVALID, "SAO enabled without parameters");
    if (p->weighted)
        for (int c = 0; c < 3; ++c)
            if (p->log2_wd[c] < 0 || p->log2_wd[c] > 14) return fail(RV_E_RANGE, "log2_wd outside 0..14");
    const int W = p->width, H = p->height;
    int64_t cu_area = 0, pred_area = 0, tu_area = 0;
    cu_at.clear();
    std::vector<int32_t> chroma_area;
    for (int i = 0; i < p->n_cus; ++i) {
        const rv_hevc_cu& c = p->cus[i];
        if (c.log2_size < 3 || c.log2_size > 6) return fail(RV_E_RANGE, "CU size outside 8..64");
        int sz = 1 << c.log2_size;
        if ((c.x & 7) || (c.y & 7) || c.x + sz > W || c.y + sz > H) return fail(RV_E_RANGE, "CU outside the picture");
        if (c.qp_y < -6 * (s->bd - 8) || c.qp_y > 51) return fail(RV_E_RANGE, "QpY outside -QpBdOffset..51");
        if (c.intra > RV_INTRA_DEVICE) return fail(RV_E_RANGE, "CU intra kind");
        cu_area += (int64_t)sz * sz;
        if (c.intra == RV_INTRA_HOST) {
            if ((uint64_t)c.pred_off + (uint64_t)sz * sz * 3 / 2 > p->n_intra_pred)
                return fail(RV_E_RANGE, "intra prediction outside intra_pred");
        }
        if (c.intra == RV_INTRA_DEVICE) {
            if (c.part_mode != 0 && c.part_mode != 3) return fail(RV_E_RANGE, "device-intra CU neither 2Nx2N nor NxN");
            for (int k = 0; k < (c.part_mode == 3 ? 4 : 1); ++k)
                if (c.luma_mode[k] > 34) return fail(RV_E_RANGE, "intra mode outside 0..34");
            if (c.chroma_mode > 34) return fail(RV_E_RANGE, "intra mode outside 0..34");
            if (cu_at.empty()) {
                cu_at.assign((size_t)(W / 4) * (H / 4), -1);
                chroma_area.assign(p->n_cus, 0);
            }
            for (int by = c.y / 4; by < (c.y + sz) / 4; ++by)
                for (int bx = c.x / 4; bx < (c.x + sz) / 4; ++bx) cu_at[(size_t)by * (W / 4) + bx] = i;
        }
        if (c.intra) pred_area += (int64_t)sz * sz;
    }
    for (int i = 0; i < p->n_pus; ++i) {
        const rv_hevc_pu& u = p->pus[i];
        if (u.w < 4 || u.h < 4 || u.w > 64 || u.h > 64 || (u.w & 3) || (u.h & 3) || (u.x & 3) || (u.y & 3) ||
            u.x + u.w > W || u.y + u.h > H)
            return fail(RV_E_RANGE, "PU outside the picture or of bad size");
        if (u.dir < 1 || u.dir > 3) return fail(RV_E_RANGE, "PU without a list");
        for (int l = 0; l < 2; ++l)
            if ((u.dir >> l) & 1)
                if (u.ref_idx[l] < 0 || u.ref_idx[l] >= p->num_ref[l]) return fail(RV_E_RANGE, "ref_idx outside the list");
        pred_area += (int64_t)u.w * u.h;
    }
    for (int i = 0; i < p->n_tus; ++i) {
        const rv_hevc_tu& t = p->tus[i];
        if (t.log2 < 2 || t.log2 > 5 || t.comp > 2) return fail(RV_E_RANGE, "TU size / component");
        int n = 1 << t.log2, pw = t.comp ? W / 2 : W, ph = t.comp ? H / 2 : H;
        if ((t.x & 3) || (t.y & 3) || t.x + n > pw || t.y + n > ph) return fail(RV_E_RANGE, "TU outside the plane");
        if ((t.flags & RV_TU_DST) && (t.log2 != 2 || t.comp != 0)) return fail(RV_E_RANGE, "DST on a non-4x4 luma TU");
        if (t.flags & ~(RV_TU_DST | RV_TU_CODED | RV_TU_LEVELS | RV_TU_INTRA)) return fail(RV_E_RANGE, "unknown TU flag");
        const bool levels = (t.flags & RV_TU_LEVELS) != 0;
        if (levels && t.qp > 51 + 6 * (s->bd - 8)) return fail(RV_E_RANGE, "TU qP outside 0..51 + QpBdOffset");
        if (t.flags & RV_TU_CODED) {
            if (t.max_x >= n || t.max_y >= n) return fail(RV_E_RANGE, "nonzero bounds outside the TU");
            if (t.n_nz == 0 || t.n_nz > n * n || (uint64_t)t.coef_off + t.n_nz > p->n_nz)
                return fail(RV_E_RANGE, "nonzero list outside nz");
            // The reference transforms (idct_scalar, src/dsp/hevc.rs:499-526)
            // and scales (scale_coefficients, src/hevc/residual.rs:1255-1256)
            // only inside the bounds; a nonzero outside them would be
            // dropped or left unscaled there and used here, so it is refused.
            // The DST transforms the whole 4x4 and needs no bounds unless
            // its levels are scaled.
            const bool bounded = levels || (t.flags & RV_TU_DST) == 0;
            for (uint32_t k = 0; k < t.n_nz; ++k) {
                const rv_nz& e = p->nz[t.coef_off + k];
                if (e.pos >= n * n || (bounded && ((e.pos & (n - 1)) > t.max_x || (e.pos >> t.log2) > t.max_y)))
                    return fail(RV_E_RANGE, "nonzero coefficient outside the TU or its bounds");
            }
        }
        if (t.comp == 0) tu_area += (int64_t)n * n;
        if (!cu_at.empty()) {
            const int sc = t.comp ? 1 : 0;
            const int ci = cu_at[(size_t)((t.y << sc) >> 2) * (W / 4) + ((t.x << sc) >> 2)];
            if (ci >= 0) {
                // A TU of a device-intra CU lies inside it, an NxN CU's luma
                // TUs inside one PU.
                const rv_hevc_cu& c = p->cus[ci];
                const int cs = (1 << c.log2_size) >> sc, cx = c.x >> sc, cy = c.y >> sc;
                if (t.x < cx || t.y < cy || t.x + n > cx + cs || t.y + n > cy + cs)
                    return fail(RV_E_RANGE, "TU across a device-intra CU");
                if (!sc && c.part_mode == 3 && n > cs / 2) return fail(RV_E_RANGE, "luma TU across NxN intra PUs");
                if (sc) chroma_area[ci] += n * n;
            }
        }
    }
    const int64_t area = (int64_t)W * H;
    if (cu_area != area || pred_area != area || tu_area != area)
        return fail(RV_E_RANGE, "CU / PU / luma TU lists do not tile the picture");
    for (size_t i = 0; i < chroma_area.size(); ++i) {
        const int cs = (1 << p->cus[i].log2_size) / 2;
        if (p->cus[i].intra == RV_INTRA_DEVICE && chroma_area[i] != 2 * cs * cs)
            return fail(RV_E_RANGE, "device-intra CU without all its chroma TUs");
    }
    return RV_OK;
}

// Dependency levels of the TUs of device-intra CUs (decoding order = list
// order): a TU's level is one more than the highest level among the TUs
// whose samples its references cover (its left column and corner,
// p[-1][-1 .. 2n-1], and its top row, p[0 .. 2n-1][-1]); samples of other
// CUs are ready before the intra kernel (level 0). Per component, a map of
// 4x4 cells. Returns the levels (n_levels = max + 1 of them, level 0
// unused) per TU index, -1 for TUs of other CUs.
int intra_levels(const rv_hevc_pic* p, const std::vector<int32_t>& cu_at, std::vector<int32_t>& lvl) {
    const int W = p->width, H = p->height;
    lvl.assign(p->n_tus, -1);
    std::vector<int32_t> map[3];
    map[0].assign((size_t)(W / 4) * (H / 4), 0);
    map[1].assign((size_t)(W / 8) * (H / 8), 0);
    map[2].assign(map[1].size(), 0);
    int top = 0;
    for (int i = 0; i < p->n_tus; ++i) {
        const rv_hevc_tu& t = p->tus[i];
        const int sc = t.comp ? 1 : 0;
        if (cu_at[(size_t)((t.y << sc) >> 2) * (W / 4) + ((t.x << sc) >> 2)] < 0) continue;
        const int n = 1 << t.log2, pw = W >> sc, ph = H >> sc, w4 = pw / 4;
        std::vector<int32_t>& m = map[t.comp];
        int l = 0;
        if (t.x > 0) {
            const int cx = (t.x - 1) >> 2;
            for (int cy = std::max(t.y - 1, 0) >> 2; cy <= (std::min(t.y + 2 * n, ph) - 1) >> 2; ++cy)
                l = std::max(l, m[(size_t)cy * w4 + cx]);
        }
        if (t.y > 0) {
            const int cy = (t.y - 1) >> 2;
            for (int cx = t.x >> 2; cx <= (std::min(t.x + 2 * n, pw) - 1) >> 2; ++cx) l = std::max(l, m[(size_t)cy * w4 + cx]);
        }
        lvl[i] = l + 1;
        top = std::max(top, l + 1);
        for (int cy = t.y >> 2; cy < (t.y + n) >> 2; ++cy)
            for (int cx = t.x >> 2; cx < (t.x + n) >> 2; ++cx) m[(size_t)cy * w4 + cx] = l + 1;
    }
    return top;
}


// implicit_pair (src/h264/recon.rs:214-233): weights at logWD 5 from POC
// distances; (32, 32) for long-term references, equal POCs or a ratio out
// of range.
void implicit_weights(int cur, int poc0, int poc1, bool lt0, bool lt1, int* w0, int* w1) {
    int tb = std::min(std::max(cur - poc0, -128), 127);
    int td = std::min(std::max(poc1 - poc0, -128), 127);
    *w0 = *w1 = 32;
    if (td == 0 || lt0 || lt1) return;
    int tx = (16384 + std::abs(td / 2)) / td;
    int dsf = std::min(std::max((tb * tx + 32) >> 6, -1024), 1023);
    int v1 = dsf >> 2;
    if (v1 < -64 || v1 > 128) return;
    *w0 = 64 - v1;
    *w1 = v1;
}

// Neighbouring macroblocks of H.264 macroblock addr that may supply intra
// prediction samples in a one-slice frame: bit 0 left, 1 above, 2
// above-right, 3 above-left; each present and, under constrained intra
// prediction, intra (intra_ok, src/h264/recon.rs:301-307).
uint32_t h264_neighbours(const rv_h264_pic* p, int addr) {
    const int mbw = p->width / 16, mbx = addr % mbw, mby = addr / mbw;
    auto ok = [&](int n) { return !p->constrained_intra_pred || p->mbs[n].intra != 0; };
    return (mbx > 0 && ok(addr - 1) ? 1u : 0u) | (mby > 0 && ok(addr - mbw) ? 2u : 0u) |
           (mby > 0 && mbx < mbw - 1 && ok(addr - mbw + 1) ? 4u : 0u) |
           (mbx > 0 && mby > 0 && ok(addr - mbw - 1) ? 8u : 0u);
}

int validate_h264(const rv_session* s, const rv_h264_pic* p, int out_slot) {
    if (p->width != s->w || p->height != s->h || p->bit_depth != 8)
        return fail(RV_E_RANGE, "picture shape differs from the session");
    const int W = p->width, H = p->height, mbw = W / 16, mbh = H / 16;
    if (p->n_mbs != mbw * mbh || !p->mbs) return fail(RV_E_RANGE, "n_mbs must cover the picture");
    for (int l = 0; l < 2; ++l) {
        if (p->num_ref[l] < 0 || p->num_ref[l] > RV_MAX_REFS) return fail(RV_E_RANGE, "num_ref outside 0..4");
        for (int r = 0; r < p->num_ref[l]; ++r) {
            if (p->ref_slot[l][r] < 0 || p->ref_slot[l][r] >= s->slots) return fail(RV_E_RANGE, "ref_slot outside the DPB");
            if (p->ref_slot[l][r] == out_slot) return fail(RV_E_RANGE, "out_slot is also a reference");
        }
    }
    if (p->weighting < RV_H264_DEFAULT || p->weighting > RV_H264_IMPLICIT) return fail(RV_E_INVALID, "weighting");
    if (p->weighting == RV_H264_EXPLICIT)
        for (int c = 0; c < 3; ++c)
            if (p->log_wd[c] < 0 || p->log_wd[c] > 7) return fail(RV_E_RANGE, "log_wd outside 0..7");
    if (p->n_parts < 0 || p->n_blks < 0 || (p->n_parts && !p->parts) || (p->n_blks && !p->blks) ||
        (p->n_coefs && !p->coefs) || (p->n_intra_pred && !p->intra_pred))
        return fail(RV_E_INVALID, "null work list");
    int64_t area = 0;
    for (int i = 0; i < p->n_mbs; ++i) {
        const rv_h264_mb& m = p->mbs[i];
        if (m.qp < 0 || m.qp > 51 || m.qpc[0] < 0 || m.qpc[0] > 51 || m.qpc[1] < 0 || m.qpc[1] > 51)
            return fail(RV_E_RANGE, "QP outside 0..51");
        if (m.intra == RV_INTRA_HOST) {
            if ((uint64_t)m.pred_off + 384 > p->n_intra_pred) return fail(RV_E_RANGE, "intra prediction outside intra_pred");
            area += 256;
        } else if (m.intra == RV_INTRA_DEVICE) {
            // The modes must have the neighbours they read, as the reference's
            // predictors require (src/h264/intra.rs:78-88, 109-325, 622-790).
            const uint32_t nb = h264_neighbours(p, i);
            const bool A = nb & 1, B = nb & 2, D = nb & 8;
            if (m.intra_kind > RV_H264_I16X16) return fail(RV_E_RANGE, "intra_kind");
            if ((m.intra_kind == RV_H264_I8X8) != (m.transform_8x8 != 0))
                return fail(RV_E_RANGE, "transform_8x8 differs from the intra kind");
            if (m.intra_kind == RV_H264_I16X16) {
                const int md = m.modes[0];
                if (md > 3 || (md == 0 && !B) || (md == 1 && !A) || (md == 3 && !(A && B && D)))
                    return fail(RV_E_INVALID, "Intra_16x16 mode without its neighbours");
            } else {
                const int step = m.intra_kind == RV_H264_I8X8 ? 2 : 1;
                for (int by = 0; by < 4; by += step)
                    for (int bx = 0; bx < 4; bx += step) {
                        const bool left = bx > 0 || A, top = by > 0 || B;
                        const bool tl = (bx > 0 && by > 0) || (bx == 0 ? (by == 0 ? D : A) : B);
                        const int md = m.modes[by * 4 + bx];
                        const bool ok = md == 2 || ((md == 0 || md == 3 || md == 7) && top) ||
                                        ((md == 1 || md == 8) && left) || (md >= 4 && md <= 6 && top && left && tl);
                        if (md > 8 || !ok) return fail(RV_E_INVALID, "Intra_NxN mode without its neighbours");
                    }
            }
            const int cm = m.chroma_mode;
            if (cm > 3 || (cm == 1 && !A) || (cm == 2 && !B) || (cm == 3 && !(A && B && D)))
                return fail(RV_E_INVALID, "chroma intra mode without its neighbours");
            area += 256;
        } else if (m.intra) {
            return fail(RV_E_RANGE, "intra outside 0..2");
        }
    }
    for (int i = 0; i < p->n_parts; ++i) {
        const rv_h264_part& t = p->parts[i];
        if ((t.w != 4 && t.w != 8 && t.w != 16) || (t.h != 4 && t.h != 8 && t.h != 16) || (t.x & 3) || (t.y & 3) ||
            t.x + t.w > W || t.y + t.h > H)
            return fail(RV_E_RANGE, "partition outside the picture or of bad size");
        if ((t.x >> 4) != ((t.x + t.w - 1) >> 4) || (t.y >> 4) != ((t.y + t.h - 1) >> 4))
            return fail(RV_E_RANGE, "partition crosses a macroblock");
        if (p->mbs[(t.y >> 4) * mbw + (t.x >> 4)].intra) return fail(RV_E_RANGE, "partition inside an intra macroblock");
        if (t.dir < 1 || t.dir > 3) return fail(RV_E_RANGE, "partition without a list");
        for (int l = 0; l < 2; ++l)
            if ((t.dir >> l) & 1)
                if (t.ref_idx[l] < 0 || t.ref_idx[l] >= p->num_ref[l]) return fail(RV_E_RANGE, "ref_idx outside the list");
        area += (int64_t)t.w * t.h;
    }
    if (area != (int64_t)W * H) return fail(RV_E_RANGE, "partitions and intra macroblocks do not tile the picture");
    for (int i = 0; i < p->n_blks; ++i) {
        const rv_h264_blk& b = p->blks[i];
        int pw = b.comp ? W / 2 : W, ph = b.comp ? H / 2 : H;
        if (b.comp > 2 || (b.size != 4 && b.size != 8) || (b.size == 8 && b.comp != 0) || (b.x & 3) || (b.y & 3) ||
            b.x + b.size > pw || b.y + b.size > ph)
            return fail(RV_E_RANGE, "residual block outside its plane or of bad size");
        if ((uint64_t)b.coef_off + (uint64_t)b.size * b.size > p->n_coefs) return fail(RV_E_RANGE, "coefficients outside coefs");
        if (b.flags & ~RV_BLK_LEVELS) return fail(RV_E_RANGE, "unknown block flag");
        if ((b.flags & RV_BLK_LEVELS) && b.has_dc) {
            // A DC group: the four chroma DCs of a macroblock's component, or
            // the sixteen luma DCs of an Intra_16x16 macroblock.
            const int group = b.comp ? 4 : 16;
            if (b.size != 4) return fail(RV_E_RANGE, "DC on an 8x8 block");
            if (!b.comp && !p->mbs[(b.y >> 4) * mbw + (b.x >> 4)].intra)
                return fail(RV_E_RANGE, "luma DC outside an intra macroblock");
            if (b.dc < 0 || (uint64_t)b.dc + group > p->n_coefs) return fail(RV_E_RANGE, "DC levels outside coefs");
        }
    }
    // The blocks of device-intra macroblocks: luma of the macroblock's
    // transform size, a luma DC only under Intra_16x16, each position
    // listed once (the device adds them in LDS by position).
    std::vector<uint32_t> seen;
    for (int i = 0; i < p->n_blks; ++i) {
        const rv_h264_blk& b = p->blks[i];
        const int mb = b.comp ? (b.y >> 3) * mbw + (b.x >> 3) : (b.y >> 4) * mbw + (b.x >> 4);
        const rv_h264_mb& m = p->mbs[mb];
        if (m.intra != RV_INTRA_DEVICE) continue;
        if (seen.empty()) seen.assign((size_t)p->n_mbs, 0u);
        uint32_t bits;
        if (b.comp == 0) {
            if (b.size != (m.intra_kind == RV_H264_I8X8 ? 8 : 4)) return fail(RV_E_RANGE, "block size differs from the intra kind");
            if (b.has_dc && m.intra_kind != RV_H264_I16X16) return fail(RV_E_RANGE, "luma DC outside Intra_16x16");
            const int bx = (b.x & 15) >> 2, by = (b.y & 15) >> 2, w4 = b.size / 4;
            bits = 0;
            for (int y = 0; y < w4; ++y)
                for (int x = 0; x < w4; ++x) bits |= 1u << ((by + y) * 4 + bx + x);
        } else {
            bits = 1u << (16 + (b.comp - 1) * 4 + ((b.y & 7) >> 2) * 2 + ((b.x & 7) >> 2));
        }
        if (seen[(size_t)mb] & bits) return fail(RV_E_RANGE, "block listed twice in a device-intra macroblock");
        seen[(size_t)mb] |= bits;
    }
    return RV_OK;
}

void h264_bytes(const rv_h264_pic* p, int64_t out[4]) {
    int64_t mc = 0;
    for (int i = 0; i < p->n_parts; ++i) {
        const rv_h264_part& t = p->parts[i];
        for (int l = 0; l < 2; ++l)
            if ((t.dir >> l) & 1) mc += (int64_t)(t.w + 5) * (t.h + 5) + 2 * (int64_t)(t.w / 2 + 1) * (t.h / 2 + 1);
    }
    int64_t planes = (int64_t)p->width * p->height * 3 / 2;
    // Intra macroblocks: a host prediction is read and written in stage 0;
    // a device one reads its neighbours (25 + 16 luma, 2 x 17 chroma) and
    // writes its 384 samples after the transform stage.
    int64_t intra_dev = 0;
    for (int i = 0; i < p->n_mbs; ++i) {
        if (p->mbs[i].intra == RV_INTRA_HOST) mc += 384 * 2;
        else if (p->mbs[i].intra == RV_INTRA_DEVICE) intra_dev += 384 + 75;
    }
    out[0] = mc + planes;
    int64_t coef = 0;
    for (int i = 0; i < p->n_blks; ++i) coef += (int64_t)p->blks[i].size * p->blks[i].size;
    out[1] = 4 * coef + 2 * coef + intra_dev;
    out[2] = p->deblock ? 2 * planes : 0;
    out[3] = 0;
}

}  // namespace

extern "C" {

int rv_abi_size(int which) {
    switch (which) {
        case 0: return (int)sizeof(rv_hevc_cu);
        case 1: return (int)sizeof(rv_hevc_pu);
        case 2: return (int)sizeof(rv_hevc_tu);
        case 3: return (int)sizeof(rv_sao);
        case 4: return (int)sizeof(rv_hevc_pic);
        case 5: return (int)sizeof(rv_h264_mb);
        case 6: return (int)sizeof(rv_h264_part);
        case 7: return (int)sizeof(rv_h264_blk);
        case 8: return (int)sizeof(rv_h264_pic);
        case 9: return (int)sizeof(rv_nz);
        case 10: return (int)sizeof(rv_hevc_scaling);
        case 11: return (int)sizeof(rv_h264_scaling);
        case 12: return (int)sizeof(rv_hevc_slice);
        default: return -1;
    }
}

int rv_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int rv_init(int device_ordinal) {
    try {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_device >= 0) {
            if (g_device == device_ordinal) return RV_OK;
            return fail(RV_E_STATE, "already initialised on another device");
        }
        // RIVET_HIP_DISABLE=1 (like H26X_NO_SIMD, src/dsp/mod.rs:295-350)
        // refuses the tier, so the caller keeps its own CPU tables.
        const char* off = std::getenv("RIVET_HIP_DISABLE");
        if (off && *off && *off != '0') return fail(RV_E_NODEVICE, "disabled by RIVET_HIP_DISABLE");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(RV_E_NODEVICE, "no HIP device");
        if (device_ordinal < 0 || device_ordinal >= n) return fail(RV_E_NODEVICE, "device ordinal out of range");
        RV_HIP(hipSetDevice(device_ordinal));
        hipDeviceProp_t prop;
        RV_HIP(hipGetDeviceProperties(&prop, device_ordinal));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            return fail(RV_E_NODEVICE, std::string("kernels are built for gfx950, device is ") + prop.gcnArchName);
        RV_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
        RV_HIP(hipStreamCreateWithFlags(&g_io, hipStreamNonBlocking));
        RV_HIP(hipStreamCreateWithFlags(&g_up, hipStreamNonBlocking));
        for (int i = 0; i < kMarks; ++i) RV_HIP(hipEventCreate(&g_marks[i]));
        RV_HIP(hipMalloc((void**)&g_dev_err, 256));
        RV_HIP(hipMemset(g_dev_err, 0, 256));
        g_err_seen = 0;
        if (device_ordinal < kMaxDevices) {
            void* ptr = nullptr;
            RV_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            g_scratch_dev[device_ordinal] = ptr;
        }
        const char* e = std::getenv("RIVET_HIP_NO_DEBLOCK");
        if (e && *e && *e != '0') g_stage_mask |= 1;
        e = std::getenv("RIVET_HIP_NO_SAO");
        if (e && *e && *e != '0') g_stage_mask |= 2;
        g_device = device_ordinal;
        return RV_OK;
    } catch (...) {
        return fail(RV_E_HIP, "exception in rv_init");
    }
}

const char* rv_tier_name(void) { return g_device >= 0 ? "hip-gfx950" : "hip-uninitialised"; }
const char* rv_last_error(void) { return g_err.c_str(); }

int rv_set_stage_mask(int mask) {
    if (mask < 0 || mask > 3) return fail(RV_E_INVALID, "stage mask is 0..3");
    g_stage_mask = mask;
    return RV_OK;
}

int64_t rv_session_footprint(int codec, int width, int height, int bit_depth, int dpb_slots) {
    if (width <= 0 || height <= 0 || dpb_slots <= 0) return -1;
    PlaneGeom g = plane_geom(width, height, bit_depth);
    int64_t pic = (int64_t)(g.bytes[0] + g.bytes[1] + g.bytes[2]);
    int64_t n4 = (int64_t)(width / 4) * (height / 4);
    // DPB + one job's reconstruction planes and side maps + work lists
    // (bounded by the coefficients of a fully coded picture).
    int64_t lists = (int64_t)width * height * 3 / 2 * (codec == RV_CODEC_H264 ? 4 : 2) / 4 + n4 * 32;
    return (int64_t)dpb_slots * (pic + 3 * (int64_t)kPlaneSlack) + pic + n4 * 16 + lists;
}

int rv_session_create(int codec, int width, int height, int bit_depth, int dpb_slots, rv_session** out) {
    try {
        if (!out) return fail(RV_E_INVALID, "null out");
        *out = nullptr;
        if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
        if (codec != RV_CODEC_HEVC && codec != RV_CODEC_H264) return fail(RV_E_INVALID, "unknown codec");
        int m = codec == RV_CODEC_HEVC ? 8 : 16;
        if (width <= 0 || height <= 0 || width % m || height % m || width > 8192 || height > 8192)
            return fail(RV_E_RANGE, "picture size");
        if (bit_depth < 8 || bit_depth > 12 || (codec == RV_CODEC_H264 && bit_depth != 8))
            return fail(RV_E_RANGE, "bit depth");
        if (dpb_slots < 1 || dpb_slots > 64) return fail(RV_E_RANGE, "dpb_slots outside 1..64");
        rv_session* s = new rv_session();
        s->codec = codec;
        s->w = width;
        s->h = height;
        s->bd = bit_depth;
        s->slots = dpb_slots;
        s->g = plane_geom(width, height, bit_depth);
        s->planes.assign((size_t)dpb_slots * 3, nullptr);
        s->bufs.assign((size_t)dpb_slots * 3, DevBuf{});
        for (int i = 0; i < dpb_slots * 3; ++i) {
            if (!pool_get(s->g.bytes[i % 3] + kPlaneSlack, &s->bufs[i])) {
                std::lock_guard<std::mutex> lk(g_mu);
                for (const DevBuf& b : s->bufs) pool_put(b);
                delete s;
                return fail(RV_E_NOMEM, "DPB allocation failed");
            }
            s->planes[i] = s->bufs[i].p;
        }
        // Zeroed DPB slots (a slot read before any picture is written there
        // reads zeros), on the upload stream, waited for here.
        hipError_t e = hipSuccess;
        for (int i = 0; i < dpb_slots * 3 && e == hipSuccess; ++i)
            e = hipMemsetAsync(s->planes[i], 0, s->g.bytes[i % 3], g_up);
        if (e == hipSuccess) e = hipStreamSynchronize(g_up);
        if (e != hipSuccess) {
            std::lock_guard<std::mutex> lk(g_mu);
            for (const DevBuf& b : s->bufs) pool_put(b);
            delete s;
            return hip_fail(e, "DPB clear");
        }
        *out = s;
        return RV_OK;
    } catch (...) {
        return fail(RV_E_NOMEM, "exception in rv_session_create");
    }
}

int rv_session_destroy(rv_session* s) {
    if (!s) return fail(RV_E_INVALID, "null session");
    // Never waits: the planes return to the pool once the work launched so
    // far has completed (plane uploads and downloads are synchronous).
    std::lock_guard<std::mutex> lk(g_mu);
    for (const DevBuf& b : s->bufs) pool_put(b);
    delete s;
    return RV_OK;
}

int rv_dpb_upload(rv_session* s, int slot, const void* y, const void* cb, const void* cr) {
    if (!s || !y || !cb || !cr) return fail(RV_E_INVALID, "null argument");
    if (slot < 0 || slot >= s->slots) return fail(RV_E_RANGE, "slot outside the DPB");
    if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
    // A slot may be rewritten only once its readers have run: wait for the
    // latest launch that covered this session (none for a new session),
    // never for other sessions' work. The planes go through the calling
    // thread's pinned staging in one DMA transfer per plane.
    RV_HIP(wait_session(s));
    size_t row[3], off[3], total = 0;
    for (int c = 0; c < 3; ++c) {
        row[c] = (size_t)s->g.pw[c] * s->g.ss;
        off[c] = total;
        total += row[c] * s->g.ph[c];
    }
    char* st = t_stage.get(total);
    if (!st) return fail(RV_E_NOMEM, "pinned staging allocation failed");
    const void* src[3] = {y, cb, cr};
    for (int c = 0; c < 3; ++c) memcpy(st + off[c], src[c], row[c] * s->g.ph[c]);
    for (int c = 0; c < 3; ++c)
        RV_HIP(hipMemcpy2DAsync(s->planes[(size_t)slot * 3 + c], (size_t)s->g.pitch[c] * s->g.ss, st + off[c],
                                row[c], row[c], s->g.ph[c], hipMemcpyHostToDevice, g_up));
    if (!t_stage.ev) RV_HIP(hipEventCreateWithFlags(&t_stage.ev, hipEventDisableTiming));
    RV_HIP(hipEventRecord(t_stage.ev, g_up));
    RV_HIP(hipEventSynchronize(t_stage.ev));
    return RV_OK;
}

int rv_plane_download(rv_session* s, int slot, void* y, void* cb, void* cr) {
    if (!s || !y || !cb || !cr) return fail(RV_E_INVALID, "null argument");
    if (slot < 0 || slot >= s->slots) return fail(RV_E_RANGE, "slot outside the DPB");
    if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
    // The slot's writer is at latest this session's last launch.
    RV_HIP(wait_session(s));
    void* dst[3] = {y, cb, cr};
    for (int c = 0; c < 3; ++c)
        RV_HIP(hipMemcpy2DAsync(dst[c], (size_t)s->g.pw[c] * s->g.ss, s->planes[(size_t)slot * 3 + c],
                                (size_t)s->g.pitch[c] * s->g.ss, (size_t)s->g.pw[c] * s->g.ss, s->g.ph[c],
                                hipMemcpyDeviceToHost, g_io));
    RV_HIP(hipStreamSynchronize(g_io));
    return check_device_errors();
}

int rv_picture_hash(rv_session* s, int slot, int type, void* out) {
    try {
        if (!s || !out) return fail(RV_E_INVALID, "null argument");
        if (type < RV_HASH_MD5 || type > RV_HASH_CHECKSUM) return fail(RV_E_INVALID, "unknown hash type");
        if (slot < 0 || slot >= s->slots) return fail(RV_E_RANGE, "slot outside the DPB");
        if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
        rv::HashPlanes hp;
        for (int c = 0; c < 3; ++c) {
            hp.p[c] = (const uint8_t*)s->planes[(size_t)slot * 3 + c];
            hp.w[c] = s->g.pw[c];
            hp.h[c] = s->g.ph[c];
            hp.pitch[c] = s->g.pitch[c] * s->g.ss;
        }
        hp.ss = s->g.ss;
        DevBuf buf;
        if (!pool_get(rv::hash_scratch_bytes(type, hp), &buf)) return fail(RV_E_NOMEM, "hash scratch allocation failed");
        hipEvent_t ev = nullptr;
        int rc = RV_OK;
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            rc = hip_fail(e, "event create");
        } else {
            // On the kernel stream, after the launch that wrote the slot.
            std::lock_guard<std::mutex> lk(g_mu);
            e = rv::hash_launch(type, hp, buf.p, g_stream);
            if (e == hipSuccess) e = hipEventRecord(ev, g_stream);
            if (e != hipSuccess) rc = hip_fail(e, "picture hash launch");
        }
        if (rc == RV_OK) {
            const size_t n = type == RV_HASH_MD5 ? 48 : type == RV_HASH_CRC ? 6 : 12;
            e = hipEventSynchronize(ev);
            if (e == hipSuccess) e = hipMemcpy(out, buf.p, n, hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = hip_fail(e, "picture hash result");
        }
        {
            std::lock_guard<std::mutex> lk(g_mu);
            pool_put(buf);
        }
        if (ev) (void)hipEventDestroy(ev);
        return rc;
    } catch (...) {
        return fail(RV_E_NOMEM, "exception in rv_picture_hash");
    }
}

int rv_frame_md5(int n, rv_session* const* sessions, const int* slots, const int* widths, const int* heights,
                 uint8_t* out) {
    try {
        if (n < 0) return fail(RV_E_INVALID, "negative picture count");
        if (n == 0) return RV_OK;
        if (!sessions || !slots || !widths || !heights || !out) return fail(RV_E_INVALID, "null argument");
        if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
        std::vector<rv::FrameMd5> frames((size_t)n);
        for (int i = 0; i < n; ++i) {
            const rv_session* s = sessions[i];
            if (!s) return fail(RV_E_INVALID, "null session");
            if (slots[i] < 0 || slots[i] >= s->slots) return fail(RV_E_RANGE, "slot outside the DPB");
            if (widths[i] <= 0 || heights[i] <= 0 || widths[i] > s->w || heights[i] > s->h)
                return fail(RV_E_RANGE, "output size outside the session's pictures");
            rv::FrameMd5& f = frames[(size_t)i];
            for (int c = 0; c < 3; ++c) {
                f.p[c] = (const uint8_t*)s->planes[(size_t)slots[i] * 3 + c];
                f.w[c] = c ? (widths[i] + 1) / 2 : widths[i];
                f.h[c] = c ? (heights[i] + 1) / 2 : heights[i];
                f.pitch[c] = s->g.pitch[c] * s->g.ss;
            }
            f.ss = s->g.ss;
        }
        const size_t tbl = align_up((size_t)n * sizeof(rv::FrameMd5), 256);
        DevBuf buf;
        if (!pool_get(tbl + 16 * (size_t)n, &buf)) return fail(RV_E_NOMEM, "frame hash buffer allocation failed");
        hipEvent_t ev = nullptr;
        int rc = RV_OK;
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            rc = hip_fail(e, "event create");
        } else {
            // On the kernel stream, after the launches that wrote the slots.
            std::lock_guard<std::mutex> lk(g_mu);
            e = hipMemcpyAsync(buf.p, frames.data(), (size_t)n * sizeof(rv::FrameMd5), hipMemcpyHostToDevice, g_stream);
            if (e == hipSuccess)
                e = rv::frame_md5_launch((const rv::FrameMd5*)buf.p, n, (uint8_t*)buf.p + tbl, g_stream);
            if (e == hipSuccess) e = hipEventRecord(ev, g_stream);
            if (e != hipSuccess) rc = hip_fail(e, "frame hash launch");
        }
        if (rc == RV_OK) {
            e = hipEventSynchronize(ev);
            if (e == hipSuccess) e = hipMemcpy(out, (uint8_t*)buf.p + tbl, 16 * (size_t)n, hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = hip_fail(e, "frame hash result");
        }
        {
            std::lock_guard<std::mutex> lk(g_mu);
            pool_put(buf);
        }
        if (ev) (void)hipEventDestroy(ev);
        return rc;
    } catch (...) {
        return fail(RV_E_NOMEM, "exception in rv_frame_md5");
    }
}

int rv_picture_hash_noalloc(rv_session* s, int slot, int type, void* out) {
    try {
        if (!s || !out) return fail(RV_E_INVALID, "null argument");
        if (type < RV_HASH_MD5 || type > RV_HASH_CHECKSUM) return fail(RV_E_INVALID, "unknown hash type");
        if (slot < 0 || slot >= s->slots) return fail(RV_E_RANGE, "slot outside the DPB");
        if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
        rv::HashPlanes hp;
        for (int c = 0; c < 3; ++c) {
            hp.p[c] = (const uint8_t*)s->planes[(size_t)slot * 3 + c];
            hp.w[c] = s->g.pw[c];
            hp.h[c] = s->g.ph[c];
            hp.pitch[c] = s->g.pitch[c] * s->g.ss;
        }
        hp.ss = s->g.ss;
        if (rv::hash_scratch_bytes(type, hp) > sizeof(scratch))
            return fail(RV_E_RANGE, "picture taller than RV_HASH_NOALLOC_MAX_ROWS rows");
        void* buf = g_device < kMaxDevices ? g_scratch_dev[g_device] : nullptr;
        if (!buf) return fail(RV_E_STATE, "no device address for the hash array");
        const size_t n = type == RV_HASH_MD5 ? 48 : type == RV_HASH_CRC ? 6 : 12;
        std::lock_guard<std::mutex> hk(g_noalloc_mu);
        {
            // After the launch that wrote the slot; the result comes back on
            // the same stream, so no event is needed.
            std::lock_guard<std::mutex> lk(g_mu);
            hipError_t e = rv::hash_launch(type, hp, buf, g_stream);
            if (e == hipSuccess) e = hipMemcpyAsync(out, buf, n, hipMemcpyDeviceToHost, g_stream);
            if (e != hipSuccess) return hip_fail(e, "picture hash launch");
        }
        hipError_t e = hipStreamSynchronize(g_stream);
        if (e != hipSuccess) return hip_fail(e, "picture hash result");
        return RV_OK;
    } catch (...) {
        return fail(RV_E_INVALID, "exception in rv_picture_hash_noalloc");
    }
}

// Per-CTB loop-filter words (rv::kCtb*) of a picture with several slices
// or tiles, or with CUs exempt from the loop filters: which CTB boundaries
// deblocking may cross and the slice's offsets (src/hevc/deblock.rs:
// 105-165, 186-201), which neighbouring CTBs SAO may read (Neighbours::of,
// src/hevc/sao.rs:193-238) and whether the CTB's SAO takes the per-sample
// path (a neighbour inside the picture it may not read, or an exempt
// block, :140-151). Empty: one slice, one tile, no exempt CU.
static int hevc_ctb_words(const rv_hevc_pic* p, std::vector<uint32_t>& out) {
    out.clear();
    bool any_exempt = false;
    for (int i = 0; i < p->n_cus; ++i) any_exempt |= p->cus[i].exempt != 0;
    if (!p->ctb_slice && !any_exempt) return RV_OK;
    const int ctb = 1 << p->log2_ctb;
    const int wc = (p->width + ctb - 1) / ctb, hc = (p->height + ctb - 1) / ctb;
    const size_t nc = (size_t)wc * hc;
    const bool multi = p->ctb_slice != nullptr;
    if (multi) {
        if (p->n_slices <= 0 || p->n_slices > 65535 || !p->slices || !p->ctb_slice_addr || !p->ctb_tile || !p->ctb_ts)
            return fail(RV_E_INVALID, "ctb_slice without slices, ctb_slice_addr, ctb_tile or ctb_ts");
        for (size_t c = 0; c < nc; ++c)
            if (p->ctb_slice[c] >= p->n_slices || p->ctb_ts[c] >= nc)
                return fail(RV_E_RANGE, "ctb_slice or ctb_ts outside the slices / CTBs");
        for (int k = 0; k < p->n_slices; ++k) {
            const rv_hevc_slice& sl = p->slices[k];
            if (sl.beta_offset < -128 || sl.beta_offset > 127 || sl.tc_offset < -128 || sl.tc_offset > 127)
                return fail(RV_E_RANGE, "slice beta_offset / tc_offset outside -128..127");
        }
    } else if (p->beta_offset < -128 || p->beta_offset > 127 || p->tc_offset < -128 || p->tc_offset > 127) {
        return fail(RV_E_RANGE, "beta_offset / tc_offset outside -128..127");
    }
    std::vector<uint8_t> exempt(nc, 0);
    for (int i = 0; i < p->n_cus; ++i) {
        const rv_hevc_cu& cu = p->cus[i];
        if (!(cu.exempt & RV_CU_LF_EXEMPT)) continue;
        const int sz = 1 << cu.log2_size;
        for (int cy = cu.y >> p->log2_ctb; cy <= (cu.y + sz - 1) >> p->log2_ctb; ++cy)
            for (int cx = cu.x >> p->log2_ctb; cx <= (cu.x + sz - 1) >> p->log2_ctb; ++cx)
                exempt[(size_t)cy * wc + cx] = 1;
    }
    auto slice = [&](size_t c) -> const rv_hevc_slice* { return multi ? &p->slices[p->ctb_slice[c]] : nullptr; };
    // deblock.rs:125-137: the edge between CTB cq (q side) and cp.
    auto crosses = [&](size_t cq, size_t cp) {
        if (!multi) return true;
        if (p->ctb_tile[cq] != p->ctb_tile[cp] && !p->loop_filter_across_tiles) return false;
        return p->ctb_slice_addr[cq] == p->ctb_slice_addr[cp] || slice(cq)->loop_filter_across_slices != 0;
    };
    static const int DX[8] = {-1, 1, 0, 0, -1, 1, -1, 1}, DY[8] = {0, 0, -1, 1, -1, -1, 1, 1};
    out.assign(nc, 0);
    for (int ry = 0; ry < hc; ++ry)
        for (int rx = 0; rx < wc; ++rx) {
            const size_t c = (size_t)ry * wc + rx;
            uint32_t nb = 0, inside = 0;
            for (int k = 0; k < 8; ++k) {
                const int nx = rx + DX[k], ny = ry + DY[k];
                if (nx < 0 || ny < 0 || nx >= wc || ny >= hc) continue;
                inside |= 1u << k;
                const size_t n = (size_t)ny * wc + nx;
                bool ok = true;
                if (multi && p->ctb_slice_addr[n] != p->ctb_slice_addr[c]) {
                    if (p->ctb_ts[n] < p->ctb_ts[c] && !slice(c)->loop_filter_across_slices) ok = false;
                    if (p->ctb_ts[c] < p->ctb_ts[n] && !slice(n)->loop_filter_across_slices) ok = false;
                }
                if (multi && !p->loop_filter_across_tiles && p->ctb_tile[n] != p->ctb_tile[c]) ok = false;
                if (ok) nb |= 1u << k;
            }
            uint32_t w = nb;
            if (nb != inside || exempt[c]) w |= rv::kCtbSaoPerSample;
            if (multi && slice(c)->deblocking_disabled) w |= rv::kCtbDbkOff;
            if (rx > 0 && !crosses(c, c - 1)) w |= rv::kCtbLeftOff;
            if (ry > 0 && !crosses(c, c - wc)) w |= rv::kCtbTopOff;
            const int bo = multi ? slice(c)->beta_offset : p->beta_offset;
            const int to = multi ? slice(c)->tc_offset : p->tc_offset;
            w |= (uint32_t)(uint8_t)(int8_t)bo << 16 | (uint32_t)(uint8_t)(int8_t)to << 24;
            out[c] = w;
        }
    return RV_OK;
}

int rv_hevc_prepare(rv_session* s, const rv_hevc_pic* p, int out_slot, rv_job** out) {
    try {
        if (!s || !p || !out) return fail(RV_E_INVALID, "null argument");
        *out = nullptr;
        if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
        if (s->codec != RV_CODEC_HEVC) return fail(RV_E_INVALID, "not an HEVC session");
        if (out_slot < 0 || out_slot >= s->slots) return fail(RV_E_RANGE, "out_slot outside the DPB");
        std::vector<int32_t> cu_at;
        int rc = validate_hevc(s, p, cu_at);
        if (rc) return rc;
        std::vector<uint32_t> ctbw;
        rc = hevc_ctb_words(p, ctbw);
        if (rc) return rc;
        const int W = p->width, H = p->height;
        // Device intra: the TUs of RV_INTRA_DEVICE CUs leave the stage-1
        // lists for the intra kernel, grouped by dependency level.
        std::vector<int32_t> ilvl;
        const int n_ilevels = cu_at.empty() ? 0 : intra_levels(p, cu_at, ilvl);
        std::vector<int32_t> lvl_cnt(n_ilevels + 1, 0);
        size_t n_itus = 0, n_inz = 0, n_ires = 0;
        int64_t b_intra = 0;
        const int64_t ssz = p->bit_depth > 8 ? 2 : 1;
        for (size_t i = 0; i < ilvl.size(); ++i) {
            if (ilvl[i] < 0) continue;
            const rv_hevc_tu& t = p->tus[i];
            ++lvl_cnt[ilvl[i]];
            ++n_itus;
            const int n = 1 << t.log2;
            if (t.flags & RV_TU_CODED) {
                n_inz += t.n_nz;
                n_ires += (size_t)n * n;
            }
            b_intra += ssz * ((int64_t)n * n + 4 * n + 1) + ((t.flags & RV_TU_CODED) ? 4 * (int64_t)t.n_nz : 0);
        }
        auto is_itu = [&](int i) { return !ilvl.empty() && ilvl[i] >= 0; };
        // Pass 1: counts. Every list is then written once, straight into the
        // pinned staging buffer, in the order the kernels read it:
        //  - MC: the 16-aligned core of each PU as 16x16 tiles, and the rest
        //    as blocks of at most 8x8, padded with empty blocks to a multiple
        //    of four (a wave reads four records together), in PU order;
        //  - CUs by size, PUs by shape and luma TUs by size, largest first
        //    (side data writes one class of equal shapes per work item, lane
        //    by cell; picture order within a class; the PU index stored per
        //    cell refers to this order, and the deblocking filter reads
        //    motion through it);
        //  - coded TUs by size class, smallest first, their nonzero
        //    coefficients repacked in that order so a wave's TUs read one run
        //    of memory.
        size_t n_tiles = 0, n_tiles16 = 0;
        int64_t b_mc = 0;
        const bool weighted = p->weighted != 0;
        // PUs by shape (w, h descending), the side-data classes.
        // Side-data keys, band-major (bands of 1 << kTxBandShift luma rows):
        // CUs by size, PUs by shape (w, h descending), luma TUs by size.
        const int n_sbands = (H + (1 << rv::kTxBandShift) - 1) >> rv::kTxBandShift;
        auto sband = [](int y) { return (size_t)(y >> rv::kTxBandShift); };
        auto pu_bucket = [&](const rv_hevc_pu& u) {
            return sband(u.y) * 256 + (size_t)((16 - u.w / 4) * 16 + (16 - u.h / 4));
        };
        auto cu_bucket = [&](const rv_hevc_cu& c) { return sband(c.y) * 4 + (size_t)(6 - c.log2_size); };
        auto ltu_bucket = [&](const rv_hevc_tu& t) { return sband(t.y) * 4 + (size_t)(5 - t.log2); };
        std::vector<int> pu_cnt((size_t)n_sbands * 256 + 1, 0);
        // MC blocks of a PU: the 16-aligned core in 16x16 tiles, the right
        // and bottom strips in blocks of at most 8x8. A block of an 8-bit
        // picture whose windows (luma 8-tap, chroma 4-tap, every list) lie
        // inside the planes goes to the strip kernel, the others to the
        // tile kernel.
        const int cw_pl = W / 2, ch_pl = H / 2;
        // Strips serve 8-bit pictures only (deeper ones take the tile kernel).
        const bool strips_ok = p->bit_depth == 8;
        auto interior = [&](const rv_hevc_pu& u, int dx, int dy, int tw, int th) {
            if (!strips_ok) return false;
            for (int l = 0; l < 2; ++l) {
                if (!((u.dir >> l) & 1)) continue;
                const int mvx = u.mv[l].x, mvy = u.mv[l].y;
                const int lx = u.x + dx + (mvx >> 2) - 3, ly = u.y + dy + (mvy >> 2) - 3;
                if (lx < 0 || lx + tw + 7 > W || ly < 0 || ly + th + 7 > H) return false;
                const int cx = (u.x + dx) / 2 + (mvx >> 3) - 1, cy = (u.y + dy) / 2 + (mvy >> 3) - 1;
                if (cx < 0 || cx + tw / 2 + 3 > cw_pl || cy < 0 || cy + th / 2 + 3 > ch_pl) return false;
            }
            return true;
        };
        auto for_each_block = [](const rv_hevc_pu& u, auto&& f) {
            const int w16 = u.w & ~15, h16 = u.h & ~15;
            for (int dy = 0; dy < h16; dy += 16)
                for (int dx = 0; dx < w16; dx += 16) f(dx, dy, 16, 16);
            for (int dy = 0; dy < u.h; dy += 8)
                for (int dx = dy < h16 ? w16 : 0; dx < u.w; dx += 8) f(dx, dy, std::min(8, u.w - dx), std::min(8, u.h - dy));
        };
        auto luma_group = [](int th, bool bi) { return (th == 16 ? 0 : th == 8 ? 2 : 4) + (bi ? 0 : 1); };
        auto chroma_group = [](int th, bool bi) { return 6 + (th == 16 ? 0 : th == 8 ? 2 : 4) + (bi ? 0 : 1); };
        // Strip segments: band of 128 luma rows (one band below 1024 rows),
        // then group.
        const int band_sh = H >= 1024 ? 7 : 14, n_bands = ((H - 1) >> band_sh) + 1;
        std::vector<size_t> s_cnt((size_t)n_bands * rv::kStripGroups, 0);
        auto seg = [&](int y, int g) { return (size_t)(y >> band_sh) * rv::kStripGroups + g; };
        for (int i = 0; i < p->n_pus; ++i) {
            const rv_hevc_pu& u = p->pus[i];
            for_each_block(u, [&](int dx, int dy, int tw, int th) {
                if (interior(u, dx, dy, tw, th)) {
                    s_cnt[seg(u.y + dy, luma_group(th, u.dir == 3))] += (size_t)(tw / 4);
                    s_cnt[seg(u.y + dy, chroma_group(th, u.dir == 3))] += (size_t)2 * ((tw / 2 + 3) / 4);
                } else if (tw == 16 && th == 16) {
                    ++n_tiles16;
                } else {
                    ++n_tiles;
                }
            });
            ++pu_cnt[pu_bucket(u) + 1];
            // Algorithmic bytes of motion compensation (the reference window
            // per list; whole-sample default-weighted uni reads the block).
            const bool uni_default = u.dir != 3 && !weighted;
            for (int l = 0; l < 2; ++l) {
                if (!((u.dir >> l) & 1)) continue;
                const int mvx = u.mv[l].x, mvy = u.mv[l].y;
                const bool dl = uni_default && (mvx & 3) == 0 && (mvy & 3) == 0;
                const bool dc = uni_default && (mvx & 7) == 0 && (mvy & 7) == 0;
                b_mc += dl ? (int64_t)u.w * u.h : (int64_t)(u.w + 7) * (u.h + 7);
                b_mc += 2 * (dc ? (int64_t)(u.w / 2) * (u.h / 2) : (int64_t)(u.w / 2 + 3) * (u.h / 2 + 3));
            }
        }
        const size_t n_tiles_pad = (n_tiles + 3) & ~(size_t)3;
        std::vector<int32_t> s_wave0(s_cnt.size() + 1, 0);  // first wave of each segment
        for (size_t k = 0; k < s_cnt.size(); ++k) s_wave0[k + 1] = s_wave0[k] + (int32_t)((s_cnt[k] + 63) / 64);
        const int32_t n_swaves = s_wave0.back();
        const size_t n_strips = (size_t)n_swaves * 64;
        std::vector<int> cu_cnt((size_t)n_sbands * 4 + 1, 0);
        int n_intra = 0;
        int64_t intra_area = 0;
        for (int i = 0; i < p->n_cus; ++i) {
            const rv_hevc_cu& c = p->cus[i];
            ++cu_cnt[cu_bucket(c) + 1];
            if (c.intra == RV_INTRA_HOST) {
                ++n_intra;
                intra_area += (int64_t)1 << (2 * c.log2_size);
            }
        }
        std::vector<int> ltu_cnt((size_t)n_sbands * 4 + 1, 0);
        size_t n_ltus = 0, n_ctus = 0, n_nzs = 0;
        int64_t coded_area = 0;
        // Stage-1 segments: (band of 1 << kTxBandShift luma rows, size
        // class), band-major.
        const int n_tbands = (H + (1 << rv::kTxBandShift) - 1) >> rv::kTxBandShift;
        std::vector<int32_t> seg_cnt((size_t)n_tbands * 4, 0);
        auto tx_seg = [&](const rv_hevc_tu& t) {
            return (size_t)(((t.comp ? 2 * t.y : t.y) >> rv::kTxBandShift) * 4 + (t.log2 - 2));
        };
        for (int i = 0; i < p->n_tus; ++i) {
            const rv_hevc_tu& t = p->tus[i];
            if (t.comp == 0) {
                ++n_ltus;
                ++ltu_cnt[ltu_bucket(t) + 1];
            }
            if ((t.flags & RV_TU_CODED) && !is_itu(i)) {
                ++seg_cnt[tx_seg(t)];
                ++n_ctus;
                n_nzs += t.n_nz;
                coded_area += (int64_t)1 << (2 * t.log2);
            }
        }
        // At most 8192 x 8192: 6.3 M TUs, inside txw's 24-bit index.
        if (n_ctus >= (1u << 24)) return fail(RV_E_RANGE, "more coded TUs than a stage-1 item can index");
        size_t n_txw = 0;
        for (size_t k = 0; k < seg_cnt.size(); ++k) {
            const int tpw = rv::tus_per_item((int)(k & 3));
            n_txw += (size_t)((seg_cnt[k] + tpw - 1) / tpw);
        }
        for (size_t k = 0; k + 1 < pu_cnt.size(); ++k) pu_cnt[k + 1] += pu_cnt[k];
        for (size_t k = 0; k + 1 < cu_cnt.size(); ++k) cu_cnt[k + 1] += cu_cnt[k];
        for (size_t k = 0; k + 1 < ltu_cnt.size(); ++k) ltu_cnt[k + 1] += ltu_cnt[k];
        // Side-data classes over the sorted lists, band by band: CUs (64 ..
        // 8), PUs by shape, luma TUs (32 .. 4); empty classes are left out.
        std::vector<rv::SideClass> sclass;
        int n_side = 0;
        auto add_class = [&](int kind, int first, int count, int cw, int ch) {
            if (count <= 0) return;
            sclass.push_back(rv::SideClass{kind, first, count, cw, ch, n_side});
            n_side += (int)(((int64_t)count * cw * ch + rv::kSideCellsPerItem - 1) / rv::kSideCellsPerItem);
        };
        for (int b = 0; b < n_sbands; ++b) {
            const size_t c0 = (size_t)b * 4, p0 = (size_t)b * 256;
            for (int k = 0; k < 4; ++k)
                add_class(rv::SIDE_CU, cu_cnt[c0 + k], cu_cnt[c0 + k + 1] - cu_cnt[c0 + k], 16 >> k, 16 >> k);
            for (int u = 0; u < 256; ++u)
                add_class(rv::SIDE_PU, pu_cnt[p0 + u], pu_cnt[p0 + u + 1] - pu_cnt[p0 + u], 16 - u / 16, 16 - u % 16);
            for (int k = 0; k < 4; ++k)
                add_class(rv::SIDE_TU, ltu_cnt[c0 + k], ltu_cnt[c0 + k + 1] - ltu_cnt[c0 + k], 8 >> k, 8 >> k);
        }
        const int n_sclass = (int)sclass.size();
        if (n_sclass > 65535) return fail(RV_E_RANGE, "too many side-data classes");
        const int ctb = 1 << p->log2_ctb, wc = (W + ctb - 1) / ctb, hc = (H + ctb - 1) / ctb;
        const size_t n_sao = (size_t)wc * hc * 3;

        rv_job* j = new rv_job();
        j->s = s;
        j->codec = RV_CODEC_HEVC;
        j->out_slot = out_slot;
        {
            // Algorithmic bytes per kernel class (mc, tx, dbk, sao).
            const int64_t ss = p->bit_depth > 8 ? 2 : 1, luma = (int64_t)W * H, planes = luma * 3 / 2;
            j->bytes[0] = ss * (b_mc + planes + intra_area * 3 / 2 * 2);
            j->bytes[1] = 4 * (int64_t)n_nzs + 2 * ss * coded_area + b_intra;
            j->bytes[2] = p->deblock ? 2 * (2 * ss * (luma + (n_intra ? luma / 2 : 0))) : 0;
            j->bytes[3] = 2 * ss * planes;
        }
        // Device arena.
        size_t sz = carve_size<rv::HevcPicDev>(1) + carve_size<rv_hevc_pu>(p->n_pus) + carve_size<rv_hevc_cu>(p->n_cus) +
                    carve_size<rv_hevc_tu>(n_ltus) + carve_size<rv_hevc_tu>(n_ctus) + carve_size<uint32_t>(n_txw) +
                    carve_size<rv::McTile>(n_tiles_pad) + carve_size<rv::McTile>(n_tiles16) +
                    carve_size<rv::McStrip>(n_strips) + carve_size<uint8_t>(n_swaves) + carve_size<int32_t>(2) +
                    carve_size<rv::SideClass>(n_sclass) + carve_size<uint16_t>(n_side) +
                    carve_size<rv_nz>(n_nzs + n_inz) + carve_size<rv::IntraTu>(n_itus) +
                    carve_size<int32_t>(n_ilevels + 1) + carve_size<int32_t>(2 * (size_t)(n_ilevels + 1)) +
                    carve_size<uint16_t>(p->n_intra_pred) + carve_size<rv_sao>(n_sao) +
                    carve_size<int32_t>(n_intra) + carve_size<rv_hevc_scaling>(p->scaling ? 1 : 0) +
                    carve_size<uint32_t>(ctbw.size()) + carve_size<const rv::HevcPicDev*>(1) + carve_size<int32_t>(rv::kStages * 2) + 256;
        size_t n4 = (size_t)(W / 4) * (H / 4);
        size_t scr = s->g.bytes[0] + s->g.bytes[1] + s->g.bytes[2] + carve_size<uint8_t>(n4) * 3 +
                     carve_size<int8_t>(n4) + carve_size<uint32_t>(n4) + carve_size<int16_t>(n_ires) +
                     carve_size<rv::IntraTask>(n_itus) + 1024;
        char* hbase = t_stage.get(sz);
        if (!hbase || !pool_get(sz, &j->arena) || !pool_get(scr, &j->scratch)) {
            job_release(j);
            delete j;
            return fail(RV_E_NOMEM, "job allocation failed");
        }
        char* hp = hbase;
        char* base = (char*)j->arena.p;
        auto dev = [&](char* hptr) { return base + (hptr - hbase); };
        rv::HevcPicDev* hdesc = carve<rv::HevcPicDev>(hp, 1);
        rv_hevc_pu* h_pus = carve<rv_hevc_pu>(hp, p->n_pus);
        rv_hevc_cu* h_cus = carve<rv_hevc_cu>(hp, p->n_cus);
        rv_hevc_tu* h_ltus = carve<rv_hevc_tu>(hp, n_ltus);
        rv_hevc_tu* h_ctus = carve<rv_hevc_tu>(hp, n_ctus);
        uint32_t* h_txw = carve<uint32_t>(hp, n_txw);
        rv::McTile* h_tiles = carve<rv::McTile>(hp, n_tiles_pad);
        rv::McTile* h_tiles16 = carve<rv::McTile>(hp, n_tiles16);
        rv::McStrip* h_strips = carve<rv::McStrip>(hp, n_strips);
        uint8_t* h_swg = carve<uint8_t>(hp, n_swaves);
        int32_t* h_spre = carve<int32_t>(hp, 2);
        rv::SideClass* h_sclass = carve<rv::SideClass>(hp, n_sclass);
        if (n_sclass) std::memcpy(h_sclass, sclass.data(), n_sclass * sizeof(rv::SideClass));
        uint16_t* h_sidx = carve<uint16_t>(hp, n_side);
        uint32_t* h_ctbw = carve<uint32_t>(hp, ctbw.size());
        if (!ctbw.empty()) std::memcpy(h_ctbw, ctbw.data(), ctbw.size() * sizeof(uint32_t));
        for (int k = 0; k < n_sclass; ++k) {
            const int end = k + 1 < n_sclass ? sclass[(size_t)k + 1].item0 : n_side;
            for (int i = sclass[(size_t)k].item0; i < end; ++i) h_sidx[i] = (uint16_t)k;
        }
        rv_nz* h_coefs = carve<rv_nz>(hp, n_nzs + n_inz);
        rv::IntraTu* h_itus = carve<rv::IntraTu>(hp, n_itus);
        int32_t* h_ilevel = carve<int32_t>(hp, n_ilevels + 1);
        // {0, all TUs}, then per level {0, count}.
        int32_t* h_ipre = carve<int32_t>(hp, 2 * (size_t)(n_ilevels + 1));
        uint16_t* h_ip = carve<uint16_t>(hp, p->n_intra_pred);
        rv_sao* h_sao = carve<rv_sao>(hp, n_sao);
        int32_t* h_intra = carve<int32_t>(hp, n_intra);
        rv_hevc_scaling* h_scaling = carve<rv_hevc_scaling>(hp, p->scaling ? 1 : 0);
        if (p->scaling) *h_scaling = *p->scaling;
        const rv::HevcPicDev** h_ptr = carve<const rv::HevcPicDev*>(hp, 1);
        int32_t* h_pre = carve<int32_t>(hp, rv::kStages * 2);
        // Pass 2: write.
        size_t nt = 0, nt16 = 0;
        std::vector<size_t> s_at(s_cnt.size());
        for (size_t k = 0; k < s_cnt.size(); ++k) s_at[k] = (size_t)s_wave0[k] * 64;
        for (int i = 0; i < p->n_pus; ++i) {
            const rv_hevc_pu& u = p->pus[i];
            h_pus[pu_cnt[pu_bucket(u)]++] = u;
            uint8_t refs = 0;
            for (int l = 0; l < 2; ++l)
                if ((u.dir >> l) & 1) refs |= (uint8_t)(u.ref_idx[l] << (4 * l));
            auto tile = [&](rv::McTile& t, int dx, int dy, int tw, int th) {
                t.x = (uint16_t)(u.x + dx);
                t.y = (uint16_t)(u.y + dy);
                t.tw = (uint8_t)tw;
                t.th = (uint8_t)th;
                t.dir = u.dir;
                t.refs = refs;
                t.mv[0] = u.mv[0];
                t.mv[1] = u.mv[1];
            };
            auto strip = [&](size_t k, int x, int y, int cols, int comp) {
                rv::McStrip& t = h_strips[s_at[k]++];
                t.x = (uint16_t)x;
                t.y = (uint16_t)y;
                t.cols = (uint8_t)cols;
                t.dir = u.dir;
                t.refs = refs;
                t.comp = (uint8_t)comp;
                t.mv[0] = u.mv[0];
                t.mv[1] = u.mv[1];
            };
            for_each_block(u, [&](int dx, int dy, int tw, int th) {
                if (interior(u, dx, dy, tw, th)) {
                    const size_t gl = seg(u.y + dy, luma_group(th, u.dir == 3)),
                                 gc = seg(u.y + dy, chroma_group(th, u.dir == 3));
                    for (int sx = 0; sx < tw; sx += 4) strip(gl, u.x + dx + sx, u.y + dy, 4, 0);
                    for (int comp = 1; comp <= 2; ++comp)
                        for (int sx = 0; sx < tw / 2; sx += 4)
                            strip(gc, (u.x + dx) / 2 + sx, (u.y + dy) / 2, std::min(4, tw / 2 - sx), comp);
                } else if (tw == 16 && th == 16) {
                    tile(h_tiles16[nt16++], dx, dy, 16, 16);
                } else {
                    tile(h_tiles[nt++], dx, dy, tw, th);
                }
            });
        }
        if (n_tiles_pad > nt) std::memset(h_tiles + nt, 0, (n_tiles_pad - nt) * sizeof(rv::McTile));
        // Each segment padded to whole waves with its first strip, no
        // columns written; each wave's group.
        for (size_t k = 0; k < s_cnt.size(); ++k) {
            const size_t g0 = (size_t)s_wave0[k] * 64;
            for (size_t i = s_at[k]; i < (size_t)s_wave0[k + 1] * 64; ++i) {
                h_strips[i] = h_strips[g0];
                h_strips[i].cols = 0;
            }
            for (int32_t w = s_wave0[k]; w < s_wave0[k + 1]; ++w) h_swg[w] = (uint8_t)(k % rv::kStripGroups);
        }
        for (int i = 0, ni = 0; i < p->n_cus; ++i) {
            const rv_hevc_cu& c = p->cus[i];
            const int at = cu_cnt[cu_bucket(c)]++;
            h_cus[at] = c;
            if (c.intra == RV_INTRA_HOST) h_intra[ni++] = at;
        }
        // Segment starts in ctus, and the items of each segment.
        std::vector<size_t> seg_at(seg_cnt.size());
        for (size_t k = 0, a = 0, w = 0; k < seg_cnt.size(); a += (size_t)seg_cnt[k], ++k) {
            seg_at[k] = a;
            const int tpw = rv::tus_per_item((int)(k & 3));
            for (int f = 0; f < seg_cnt[k]; f += tpw)
                h_txw[w++] = (uint32_t)(a + (size_t)f) | (uint32_t)(std::min(tpw, seg_cnt[k] - f) - 1) << 24 |
                             (uint32_t)(k & 3) << 30;
        }
        // Device-intra TUs by level (list order within a level), their
        // coefficients after the stage-1 ones.
        if (n_ilevels) {
            // Levels run from 1: level l fills [h_ilevel[l - 1], h_ilevel[l]).
            h_ilevel[0] = 0;
            for (int l = 1; l <= n_ilevels; ++l) h_ilevel[l] = h_ilevel[l - 1] + lvl_cnt[l];
            h_ipre[0] = 0;
            h_ipre[1] = (int32_t)n_itus;
            for (int l = 0; l < n_ilevels; ++l) {
                h_ipre[2 * l + 2] = 0;
                h_ipre[2 * l + 3] = lvl_cnt[l + 1];
            }
            std::vector<int32_t> put(h_ilevel, h_ilevel + n_ilevels + 1);
            size_t nz_at = n_nzs, res_at = 0;
            for (int i = 0; i < p->n_tus; ++i) {
                if (!is_itu(i)) continue;
                const rv_hevc_tu& t = p->tus[i];
                rv::IntraTu& it = h_itus[put[ilvl[i] - 1]++];
                std::memset(&it, 0, sizeof(it));
                it.t = t;
                it.level = ilvl[i] - 1;
                it.res = -1;
                const int sc = t.comp ? 1 : 0;
                const rv_hevc_cu& cu = p->cus[cu_at[(size_t)((t.y << sc) >> 2) * (W / 4) + ((t.x << sc) >> 2)]];
                if (t.comp) {
                    it.mode = cu.chroma_mode;
                } else {
                    const int h = 1 << (cu.log2_size - 1);
                    it.mode = cu.part_mode == 3 ? cu.luma_mode[(t.y >= cu.y + h ? 2 : 0) + (t.x >= cu.x + h ? 1 : 0)]
                                                : cu.luma_mode[0];
                }
                if (t.flags & RV_TU_CODED) {
                    it.t.coef_off = (uint32_t)nz_at;
                    std::memcpy(h_coefs + nz_at, p->nz + t.coef_off, t.n_nz * sizeof(rv_nz));
                    nz_at += t.n_nz;
                    // Residual offsets are sums of N x N, multiples of 16.
                    it.res = (int32_t)res_at;
                    res_at += (size_t)1 << (2 * t.log2);
                }
            }
        }
        for (int i = 0; i < p->n_tus; ++i) {
            const rv_hevc_tu& t = p->tus[i];
            if (t.comp == 0) h_ltus[ltu_cnt[ltu_bucket(t)]++] = t;
            if ((t.flags & RV_TU_CODED) && !is_itu(i)) h_ctus[seg_at[tx_seg(t)]++] = t;
        }
        // Coefficients in ctus order, so a wave's TUs read one run.
        for (size_t k = 0, at = 0; k < n_ctus; ++k) {
            rv_hevc_tu& c = h_ctus[k];
            std::memcpy(h_coefs + at, p->nz + c.coef_off, c.n_nz * sizeof(rv_nz));
            c.coef_off = (uint32_t)at;
            at += c.n_nz;
        }
        if (p->n_intra_pred) std::memcpy(h_ip, p->intra_pred, p->n_intra_pred * sizeof(uint16_t));
        if (p->sao) std::memcpy(h_sao, p->sao, n_sao * sizeof(rv_sao));
        else std::memset(h_sao, 0, n_sao * sizeof(rv_sao));
        // Scratch: reconstruction planes and side maps.
        char* sp = (char*)j->scratch.p;
        rv::HevcPicDev& d = j->hd;
        std::memset(&d, 0, sizeof(d));
        d.w = W;
        d.h = H;
        d.bd = p->bit_depth;
        d.log2_ctb = p->log2_ctb;
        d.wc = wc;
        d.hc = hc;
        d.w4 = W / 4;
        d.h4 = H / 4;
        for (int c = 0; c < 3; ++c) {
            d.pitch[c] = s->g.pitch[c];
            d.pw[c] = s->g.pw[c];
            d.ph[c] = s->g.ph[c];
            d.cur[c] = sp;
            sp += s->g.bytes[c];
            d.out[c] = s->planes[(size_t)out_slot * 3 + c];
        }
        for (int l = 0; l < 2; ++l)
            for (int r = 0; r < RV_MAX_REFS; ++r) {
                int slot = r < p->num_ref[l] ? p->ref_slot[l][r] : 0;
                for (int c = 0; c < 3; ++c) d.ref[l][r][c] = s->planes[(size_t)slot * 3 + c];
                d.ref_delta[l][r] = r < p->num_ref[l] ? p->poc - p->ref_poc[l][r] : 0;
                d.ref_lt[l][r] = r < p->num_ref[l] ? (p->ref_long_term[l][r] != 0) : 0;
                for (int c = 0; c < 3; ++c) {
                    d.wp_w[l][r][c] = p->wp_w[l][r][c];
                    d.wp_o[l][r][c] = p->wp_o[l][r][c];
                }
            }
        d.weighted = p->weighted;
        for (int c = 0; c < 3; ++c) d.log2_wd[c] = p->log2_wd[c];
        d.deblock = (g_stage_mask & 1) ? 0 : p->deblock;
        d.beta_offset = p->beta_offset;
        d.tc_offset = p->tc_offset;
        d.cb_qp_offset = p->cb_qp_offset;
        d.cr_qp_offset = p->cr_qp_offset;
        d.sao_luma = (g_stage_mask & 2) ? 0 : p->sao_luma;
        d.sao_chroma = (g_stage_mask & 2) ? 0 : p->sao_chroma;
        d.n_cus = p->n_cus;
        d.n_pus = p->n_pus;
        d.n_ltus = (int)n_ltus;
        d.n_tiles = (int)n_tiles_pad;
        d.n_tiles16 = (int)n_tiles16;
        d.n_side = n_side;
        d.n_sclass = n_sclass;
        d.sclass = (const rv::SideClass*)dev((char*)h_sclass);
        d.sidx = (const uint16_t*)dev((char*)h_sidx);
        d.ctbw = ctbw.empty() ? nullptr : (const uint32_t*)dev((char*)h_ctbw);
        d.n_intra = n_intra;
        d.intra_idx = (const int32_t*)dev((char*)h_intra);
        d.n_txw = (int32_t)n_txw;
        d.txw = (const uint32_t*)dev((char*)h_txw);
        d.cus = (const rv_hevc_cu*)dev((char*)h_cus);
        d.pus = (const rv_hevc_pu*)dev((char*)h_pus);
        d.ltus = (const rv_hevc_tu*)dev((char*)h_ltus);
        d.ctus = (const rv_hevc_tu*)dev((char*)h_ctus);
        d.tiles = (const rv::McTile*)dev((char*)h_tiles);
        d.tiles16 = (const rv::McTile*)dev((char*)h_tiles16);
        d.strips = (const rv::McStrip*)dev((char*)h_strips);
        d.swg = (const uint8_t*)dev((char*)h_swg);
        d.n_swaves = n_swaves;
        d.nz = (const rv_nz*)dev((char*)h_coefs);
        d.intra_pred = (const uint16_t*)dev((char*)h_ip);
        d.sao = (const rv_sao*)dev((char*)h_sao);
        d.scaling = p->scaling ? (const rv_hevc_scaling*)dev((char*)h_scaling) : nullptr;
        d.n_ilevels = n_ilevels;
        d.ilevel = (const int32_t*)dev((char*)h_ilevel);
        d.itus = (const rv::IntraTu*)dev((char*)h_itus);
        d.strong = p->strong_intra_smoothing != 0;
        d.cip = p->constrained_intra_pred != 0;
        d.cuinfo = (uint8_t*)sp;
        sp += carve_size<uint8_t>(n4);
        d.tuinfo = (uint8_t*)sp;
        sp += carve_size<uint8_t>(n4);
        d.bs = (uint8_t*)sp;
        sp += carve_size<uint8_t>(n4);
        d.qp = (int8_t*)sp;
        sp += carve_size<int8_t>(n4);
        d.pumap = (uint32_t*)sp;
        sp += carve_size<uint32_t>(n4);
        d.ires = (int16_t*)sp;
        sp += carve_size<int16_t>(n_ires);
        rv::IntraTask* itask = (rv::IntraTask*)sp;
        rv::hevc_stage_items(d);
        *hdesc = d;
        j->desc_dev = dev((char*)hdesc);
        *h_ptr = (const rv::HevcPicDev*)j->desc_dev;
        for (int st = 0; st < rv::kStages; ++st) {
            h_pre[st * 2] = 0;
            h_pre[st * 2 + 1] = d.items[st];
        }
        j->single.n = 1;
        j->single.hevc = (const rv::HevcPicDev* const*)dev((char*)h_ptr);
        j->single.h264 = nullptr;
        j->itotal.assign(lvl_cnt.begin() + (n_ilevels ? 1 : 0), lvl_cnt.end());
        j->n_itus = (int32_t)n_itus;
        j->single.ilevels = n_ilevels;
        j->single.ipre = (const int32_t*)dev((char*)h_ipre);
        j->single.ibase = (const int32_t*)dev((char*)h_ilevel);  // one picture: its own level starts
        j->single.itask = itask;
        j->single.iall = (int32_t)n_itus;
        j->single.itotal = j->itotal.data();
        j->single.prefix = (const int32_t*)dev((char*)h_pre);
        h_spre[0] = 0;
        h_spre[1] = n_swaves;
        j->single.spre = (const int32_t*)dev((char*)h_spre);
        j->single.stotal = h_spre[1];
        // The per-4x4 maps need no clearing: for work lists whose CUs, PUs
        // and luma TUs tile the picture (checked by area above) stage 0
        // writes every cell and stage 1 every strength, and a cell left
        // unwritten by other lists can only give wrong strengths, never an
        // access outside the job (the PU index is bounds-checked, QP-indexed
        // tables clamp). Clearing them here would queue a fill kernel behind
        // the running stage kernels and block this call until it ran.
        hipError_t e = stage_upload(j->arena.p, sz, nullptr, 0);
        if (e != hipSuccess) {
            job_release(j);
            delete j;
            return hip_fail(e, "job upload");
        }
        *out = j;
        return RV_OK;
    } catch (...) {
        return fail(RV_E_NOMEM, "exception in rv_hevc_prepare");
    }
}

int rv_h264_prepare(rv_session* s, const rv_h264_pic* p, int out_slot, rv_job** out) {
    try {
        if (!s || !p || !out) return fail(RV_E_INVALID, "null argument");
        *out = nullptr;
        if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
        if (s->codec != RV_CODEC_H264) return fail(RV_E_INVALID, "not an H.264 session");
        if (out_slot < 0 || out_slot >= s->slots) return fail(RV_E_RANGE, "out_slot outside the DPB");
        int rc = validate_h264(s, p, out_slot);
        if (rc) return rc;
        const int W = p->width, H = p->height, mbw = W / 16, mbh = H / 16;
        // Blocks: those of device-intra macroblocks grouped by macroblock
        // (iblks); the others 4x4 first, then 8x8.
        std::vector<rv_h264_blk> b4, b8;
        std::vector<std::vector<int>> own;
        for (int i = 0; i < p->n_blks; ++i) {
            const rv_h264_blk& bk = p->blks[i];
            const int mb = bk.comp ? (bk.y >> 3) * mbw + (bk.x >> 3) : (bk.y >> 4) * mbw + (bk.x >> 4);
            if (p->mbs[mb].intra == RV_INTRA_DEVICE) {
                if (own.empty()) own.resize((size_t)p->n_mbs);
                own[(size_t)mb].push_back(i);
                continue;
            }
            (bk.size == 8 ? b8 : b4).push_back(bk);
        }
        std::vector<rv_h264_blk> blks(b4);
        blks.insert(blks.end(), b8.begin(), b8.end());
        // Device-intra macroblocks by dependency level: one more than the
        // highest level among the device-intra neighbours it may read (left,
        // above, above-right, above-left), raster order within a level.
        std::vector<int32_t> lvl((size_t)p->n_mbs, -1), lvl_cnt;
        int n_imbs = 0, n_iblks = 0;
        for (int a = 0; a < p->n_mbs; ++a) {
            if (p->mbs[a].intra != RV_INTRA_DEVICE) continue;
            const int mbx = a % mbw, mby = a / mbw;
            int l = 0;
            auto dep = [&](bool in, int n) {
                if (in && lvl[(size_t)n] >= 0) l = std::max(l, lvl[(size_t)n] + 1);
            };
            dep(mbx > 0, a - 1);
            dep(mby > 0, a - mbw);
            dep(mby > 0 && mbx < mbw - 1, a - mbw + 1);
            dep(mbx > 0 && mby > 0, a - mbw - 1);
            lvl[(size_t)a] = l;
            if ((int)lvl_cnt.size() <= l) lvl_cnt.resize((size_t)l + 1, 0);
            ++lvl_cnt[(size_t)l];
            ++n_imbs;
            if (!own.empty()) n_iblks += (int)own[(size_t)a].size();
        }
        const int n_ilevels = (int)lvl_cnt.size();
        rv_job* j = new rv_job();
        j->s = s;
        j->codec = RV_CODEC_H264;
        j->out_slot = out_slot;
        h264_bytes(p, j->bytes);
        int n_intra = 0;
        for (int a = 0; a < p->n_mbs; ++a) n_intra += p->mbs[a].intra ? 1 : 0;
        size_t sz = carve_size<rv::H264PicDev>(1) + carve_size<rv_h264_mb>(p->n_mbs) +
                    carve_size<rv_h264_part>(p->n_parts) + carve_size<uint32_t>(p->n_parts + n_intra) +
                    carve_size<rv_h264_blk>(blks.size()) + carve_size<rv::H264IntraMb>(n_imbs) +
                    carve_size<rv_h264_blk>(n_iblks) + carve_size<int32_t>(n_ilevels + 1) +
                    carve_size<int32_t>(2 * (size_t)(n_ilevels + 1)) +
                    carve_size<int32_t>(p->n_coefs) + carve_size<uint16_t>(p->n_intra_pred) +
                    carve_size<rv_h264_scaling>(p->scaling ? 1 : 0) +
                    carve_size<const rv::H264PicDev*>(1) + carve_size<int32_t>(rv::kStages * 2) + 256;
        size_t n4 = (size_t)(W / 4) * (H / 4), nmb = (size_t)mbw * mbh;
        const size_t nx = rv::h264_deblock_xch_words(mbw, mbh);
        size_t scr = carve_size<uint32_t>(n4) + carve_size<uint32_t>(nmb * rv::kH264DbkWords) + carve_size<uint64_t>(nx) +
                     1024;
        char* hbase = t_stage.get(sz);
        if (!hbase || !pool_get(sz, &j->arena) || !pool_get(scr, &j->scratch)) {
            job_release(j);
            delete j;
            return fail(RV_E_NOMEM, "job allocation failed");
        }
        char* hp = hbase;
        char* base = (char*)j->arena.p;
        auto dev = [&](char* hptr) { return base + (hptr - hbase); };
        rv::H264PicDev* hdesc = carve<rv::H264PicDev>(hp, 1);
        rv_h264_mb* h_mbs = carve<rv_h264_mb>(hp, p->n_mbs);
        rv_h264_part* h_parts = carve<rv_h264_part>(hp, p->n_parts);
        uint32_t* h_order = carve<uint32_t>(hp, p->n_parts + n_intra);
        rv_h264_blk* h_blks = carve<rv_h264_blk>(hp, blks.size());
        rv::H264IntraMb* h_imbs = carve<rv::H264IntraMb>(hp, n_imbs);
        rv_h264_blk* h_iblks = carve<rv_h264_blk>(hp, n_iblks);
        int32_t* h_ilevel = carve<int32_t>(hp, n_ilevels + 1);
        // {0, all macroblocks}, then per level {0, count}: the prefix rows
        // of a one-picture batch.
        int32_t* h_ipre = carve<int32_t>(hp, 2 * (size_t)(n_ilevels + 1));
        {
            h_ilevel[0] = 0;
            for (int l = 0; l < n_ilevels; ++l) h_ilevel[l + 1] = h_ilevel[l] + lvl_cnt[(size_t)l];
            std::vector<int32_t> at(h_ilevel, h_ilevel + n_ilevels);
            int32_t nb = 0;
            for (int a = 0; a < p->n_mbs; ++a) {
                if (lvl[(size_t)a] < 0) continue;
                rv::H264IntraMb& r = h_imbs[at[(size_t)lvl[(size_t)a]]++];
                r.addr = a;
                r.avail = h264_neighbours(p, a);
                r.blk0 = nb;
                r.nblk = own.empty() ? 0 : (int32_t)own[(size_t)a].size();
                for (int i = 0; i < r.nblk; ++i) h_iblks[nb++] = p->blks[own[(size_t)a][(size_t)i]];
            }
            h_ipre[0] = 0;
            h_ipre[1] = n_imbs;
            for (int l = 0; l < n_ilevels; ++l) {
                h_ipre[2 * (l + 1)] = 0;
                h_ipre[2 * (l + 1) + 1] = lvl_cnt[(size_t)l];
            }
        }
        int32_t* h_coefs = carve<int32_t>(hp, p->n_coefs);
        uint16_t* h_ip = carve<uint16_t>(hp, p->n_intra_pred);
        rv_h264_scaling* h_scaling = carve<rv_h264_scaling>(hp, p->scaling ? 1 : 0);
        if (p->scaling) *h_scaling = *p->scaling;
        const rv::H264PicDev** h_ptr = carve<const rv::H264PicDev*>(hp, 1);
        int32_t* h_pre = carve<int32_t>(hp, rv::kStages * 2);
        std::memcpy(h_mbs, p->mbs, p->n_mbs * sizeof(rv_h264_mb));
        if (p->n_parts) std::memcpy(h_parts, p->parts, p->n_parts * sizeof(rv_h264_part));
        // Motion compensation takes partitions above 8 x 8 one per
        // workgroup and the rest four per workgroup: the larger first.
        int n_big = 0;
        for (int i = 0; i < p->n_parts; ++i)
            if (p->parts[i].w > 8 || p->parts[i].h > 8) h_order[n_big++] = (uint32_t)i;
        for (int i = 0, k = n_big; i < p->n_parts; ++i)
            if (!(p->parts[i].w > 8 || p->parts[i].h > 8)) h_order[k++] = (uint32_t)i;
        // Each lane takes w h / (lanes per partition) samples, so the
        // partitions that share a workgroup are grouped by area: largest
        // first in each class, picture order within an area.
        auto by_area = [&](uint32_t a, uint32_t b) {
            return p->parts[a].w * p->parts[a].h > p->parts[b].w * p->parts[b].h;
        };
        std::stable_sort(h_order, h_order + n_big, by_area);
        std::stable_sort(h_order + n_big, h_order + p->n_parts, by_area);
        // Then the intra macroblocks, one workgroup each.
        for (int a = 0, k = p->n_parts; a < p->n_mbs; ++a)
            if (p->mbs[a].intra) h_order[k++] = (uint32_t)a;
        if (!blks.empty()) std::memcpy(h_blks, blks.data(), blks.size() * sizeof(rv_h264_blk));
        if (p->n_coefs) std::memcpy(h_coefs, p->coefs, p->n_coefs * sizeof(int32_t));
        if (p->n_intra_pred) std::memcpy(h_ip, p->intra_pred, p->n_intra_pred * sizeof(uint16_t));
        rv::H264PicDev& d = j->h4;
        std::memset(&d, 0, sizeof(d));
        d.w = W;
        d.h = H;
        d.mbw = mbw;
        d.mbh = mbh;
        d.w4 = W / 4;
        d.h4 = H / 4;
        for (int c = 0; c < 3; ++c) {
            d.pitch[c] = s->g.pitch[c];
            d.pw[c] = s->g.pw[c];
            d.ph[c] = s->g.ph[c];
            // Reconstructed and deblocked in place in the output slot.
            d.cur[c] = d.out[c] = s->planes[(size_t)out_slot * 3 + c];
        }
        for (int l = 0; l < 2; ++l)
            for (int r = 0; r < RV_MAX_REFS; ++r) {
                int slot = r < p->num_ref[l] ? p->ref_slot[l][r] : 0;
                for (int c = 0; c < 3; ++c) d.ref[l][r][c] = s->planes[(size_t)slot * 3 + c];
                d.ref_id[l][r] = r < p->num_ref[l] ? slot : -1;
                for (int c = 0; c < 3; ++c) {
                    d.wp_w[l][r][c] = p->wp_w[l][r][c];
                    d.wp_o[l][r][c] = p->wp_o[l][r][c];
                }
            }
        d.weighting = p->weighting;
        for (int c = 0; c < 3; ++c) d.log_wd[c] = p->log_wd[c];
        for (int r0 = 0; r0 < RV_MAX_REFS; ++r0)
            for (int r1 = 0; r1 < RV_MAX_REFS; ++r1) {
                int w0 = 32, w1 = 32;
                if (r0 < p->num_ref[0] && r1 < p->num_ref[1])
                    implicit_weights(p->poc, p->ref_poc[0][r0], p->ref_poc[1][r1], p->ref_long_term[0][r0] != 0,
                                     p->ref_long_term[1][r1] != 0, &w0, &w1);
                d.imp_w[r0][r1][0] = w0;
                d.imp_w[r0][r1][1] = w1;
            }
        d.deblock = (g_stage_mask & 1) ? 0 : p->deblock;
        d.offset_a = p->offset_a;
        d.offset_b = p->offset_b;
        d.n_mbs = p->n_mbs;
        d.n_parts = p->n_parts;
        d.n_blks4 = (int)b4.size();
        d.n_blks8 = (int)b8.size();
        d.mbs = (const rv_h264_mb*)dev((char*)h_mbs);
        d.parts = (const rv_h264_part*)dev((char*)h_parts);
        d.part_order = (const uint32_t*)dev((char*)h_order);
        d.n_big = n_big;
        d.n_intra = n_intra;
        d.blks = (const rv_h264_blk*)dev((char*)h_blks);
        d.n_ilevels = n_ilevels;
        d.ilevel = (const int32_t*)dev((char*)h_ilevel);
        d.imbs = (const rv::H264IntraMb*)dev((char*)h_imbs);
        d.iblks = (const rv_h264_blk*)dev((char*)h_iblks);
        d.coefs = (const int32_t*)dev((char*)h_coefs);
        d.intra_pred = (const uint16_t*)dev((char*)h_ip);
        d.scaling = p->scaling ? (const rv_h264_scaling*)dev((char*)h_scaling) : nullptr;
        char* sp = (char*)j->scratch.p;
        d.partmap = (uint32_t*)sp;
        sp += carve_size<uint32_t>(n4);
        d.dbk = (uint32_t*)sp;
        sp += carve_size<uint32_t>(nmb * rv::kH264DbkWords);
        d.xch = (uint64_t*)sp;
        d.n_xch = (int64_t)nx;
        // The exchange's poll bound and awaited-tag XOR. A job prepared
        // while the environment sets RIVET_HIP_TEST_EXCHANGE_MISS=1 awaits
        // tags that are never published and gives up after 16 polls, so the
        // error path can be tested with kernels that still finish normally.
        const char* miss = std::getenv("RIVET_HIP_TEST_EXCHANGE_MISS");
        const bool force_miss = miss && *miss && *miss != '0';
        d.xch_spin = force_miss ? 16 : 1 << 18;
        d.xch_tag_xor = force_miss ? 0x80000000u : 0u;
        d.err = g_dev_err;
        rv::h264_stage_items(d);
        *hdesc = d;
        j->desc_dev = dev((char*)hdesc);
        *h_ptr = (const rv::H264PicDev*)j->desc_dev;
        for (int st = 0; st < rv::kStages; ++st) {
            h_pre[st * 2] = 0;
            h_pre[st * 2 + 1] = d.items[st];
        }
        j->single.n = 1;
        j->single.hevc = nullptr;
        j->single.h264 = (const rv::H264PicDev* const*)dev((char*)h_ptr);
        j->single.prefix = (const int32_t*)dev((char*)h_pre);
        j->itotal = lvl_cnt;
        j->n_itus = n_imbs;
        j->single.ilevels = n_ilevels;
        j->single.ipre = (const int32_t*)dev((char*)h_ipre);
        j->single.itotal = j->itotal.data();
        // No clearing of the scratch (see rv_hevc_prepare): stage 0 writes
        // the partition map, stage 2 the edge words and, every launch, the
        // exchange; partition indices are bounds-checked where read.
        hipError_t e = stage_upload(j->arena.p, sz, nullptr, 0);
        if (e != hipSuccess) {
            job_release(j);
            delete j;
            return hip_fail(e, "job upload");
        }
        *out = j;
        return RV_OK;
    } catch (...) {
        return fail(RV_E_NOMEM, "exception in rv_h264_prepare");
    }
}

int rv_job_free(rv_job* j) {
    if (!j) return fail(RV_E_INVALID, "null job");
    // Never waits: the buffers return to the pool once the work launched
    // so far (which includes every launch of this job) has completed.
    job_release(j);
    delete j;
    return RV_OK;
}

int rv_job_bytes(const rv_job* j, int64_t out[4]) {
    if (!j || !out) return fail(RV_E_INVALID, "null argument");
    for (int i = 0; i < 4; ++i) out[i] = j->bytes[i];
    return RV_OK;
}

int rv_batch_create(rv_job* const* jobs, int n, rv_batch** out) {
    try {
        if (!jobs || !out || n <= 0) return fail(RV_E_INVALID, "empty batch");
        *out = nullptr;
        rv_batch* b = new rv_batch();
        for (int i = 0; i < n; ++i) {
            if (!jobs[i]) {
                delete b;
                return fail(RV_E_INVALID, "null job in batch");
            }
            b->jobs.push_back(jobs[i]);
            (jobs[i]->codec == RV_CODEC_HEVC ? b->hevc : b->h264).push_back(jobs[i]);
        }
        int nh = (int)b->hevc.size(), n4 = (int)b->h264.size();
        for (int i = 0; i < nh; ++i) {
            const int c = depth_class_of(b->hevc[i]->hd.bd);
            b->depth_cls = i == 0 ? c : (b->depth_cls == c ? c : 0);
        }
        {
            std::lock_guard<std::mutex> lk(g_mu);
            if (g_tuned[b->depth_cls]) std::memcpy(b->var, g_tuned_var[b->depth_cls], sizeof(b->var));
        }
        b->pre_h.assign((size_t)rv::kStages * (nh + 1), 0);
        b->pre_4.assign((size_t)rv::kStages * (n4 + 1), 0);
        for (int s = 0; s < rv::kStages; ++s) {
            for (int i = 0; i < nh; ++i)
                b->pre_h[(size_t)s * (nh + 1) + i + 1] = b->pre_h[(size_t)s * (nh + 1) + i] + b->hevc[i]->hd.items[s];
            for (int i = 0; i < n4; ++i)
                b->pre_4[(size_t)s * (n4 + 1) + i + 1] = b->pre_4[(size_t)s * (n4 + 1) + i] + b->h264[i]->h4.items[s];
        }
        // Device-intra levels: the prefix over the HEVC pictures of their TU
        // counts, all levels, then per level; each level's first task.
        int ilevels = 0;
        for (int i = 0; i < nh; ++i) ilevels = std::max(ilevels, b->hevc[i]->hd.n_ilevels);
        std::vector<int32_t> ipre((size_t)(ilevels + 1) * (nh + 1), 0), ibase((size_t)ilevels + 1, 0);
        b->itotal.assign(ilevels, 0);
        for (int i = 0; i < nh; ++i) ipre[(size_t)i + 1] = ipre[(size_t)i] + b->hevc[i]->n_itus;
        for (int l = 0; l < ilevels; ++l) {
            int32_t* row = ipre.data() + (size_t)(l + 1) * (nh + 1);
            for (int i = 0; i < nh; ++i) {
                const std::vector<int32_t>& t = b->hevc[i]->itotal;
                row[i + 1] = row[i] + (l < (int)t.size() ? t[l] : 0);
            }
            b->itotal[l] = row[nh];
            ibase[(size_t)l + 1] = ibase[(size_t)l] + row[nh];
        }
        // H.264 device-intra macroblocks: the same prefix rows over the
        // H.264 pictures (no task table: the level kernel locates its
        // picture by the prefix).
        int ilevels4 = 0;
        for (int i = 0; i < n4; ++i) ilevels4 = std::max(ilevels4, b->h264[i]->h4.n_ilevels);
        std::vector<int32_t> ipre4((size_t)(ilevels4 + 1) * (n4 + 1), 0);
        b->itotal4.assign(ilevels4, 0);
        for (int i = 0; i < n4; ++i) ipre4[(size_t)i + 1] = ipre4[(size_t)i] + b->h264[i]->n_itus;
        for (int l = 0; l < ilevels4; ++l) {
            int32_t* row = ipre4.data() + (size_t)(l + 1) * (n4 + 1);
            for (int i = 0; i < n4; ++i) {
                const std::vector<int32_t>& t = b->h264[i]->itotal;
                row[i + 1] = row[i] + (l < (int)t.size() ? t[l] : 0);
            }
            b->itotal4[l] = row[n4];
        }
        // Motion-compensation strips: the prefix over the HEVC pictures of
        // their waves.
        std::vector<int32_t> spre((size_t)nh + 1, 0);
        for (int i = 0; i < nh; ++i) spre[(size_t)i + 1] = spre[(size_t)i] + b->hevc[i]->hd.n_swaves;
        const int32_t iall = ipre[(size_t)nh];
        if (iall > 0 && !pool_get((size_t)iall * sizeof(rv::IntraTask), &b->itask)) {
            delete b;
            return fail(RV_E_NOMEM, "batch intra tasks");
        }
        size_t sz = carve_size<void*>(nh) + carve_size<void*>(n4) + carve_size<int32_t>(b->pre_h.size()) +
                    carve_size<int32_t>(b->pre_4.size()) + carve_size<int32_t>(ipre.size()) +
                    carve_size<int32_t>(ibase.size()) + carve_size<int32_t>(spre.size()) +
                    carve_size<int32_t>(ipre4.size());
        std::vector<uint32_t> host(sz / 4 + 1, 0u);
        char* hbase = (char*)host.data();
        if (!pool_get(sz, &b->tables)) {
            pool_put(b->itask);
            delete b;
            return fail(RV_E_NOMEM, "batch tables");
        }
        char* hp = hbase;
        char* base = (char*)b->tables.p;
        auto dev = [&](char* h) { return base + (h - hbase); };
        void** ph = carve<void*>(hp, nh);
        void** p4 = carve<void*>(hp, n4);
        int32_t* preh = carve<int32_t>(hp, b->pre_h.size());
        int32_t* pre4 = carve<int32_t>(hp, b->pre_4.size());
        int32_t* hipre = carve<int32_t>(hp, ipre.size());
        std::memcpy(hipre, ipre.data(), ipre.size() * 4);
        int32_t* hibase = carve<int32_t>(hp, ibase.size());
        std::memcpy(hibase, ibase.data(), ibase.size() * 4);
        int32_t* hspre = carve<int32_t>(hp, spre.size());
        std::memcpy(hspre, spre.data(), spre.size() * 4);
        int32_t* hipre4 = carve<int32_t>(hp, ipre4.size());
        std::memcpy(hipre4, ipre4.data(), ipre4.size() * 4);
        for (int i = 0; i < nh; ++i) ph[i] = b->hevc[i]->desc_dev;
        for (int i = 0; i < n4; ++i) p4[i] = b->h264[i]->desc_dev;
        std::memcpy(preh, b->pre_h.data(), b->pre_h.size() * 4);
        std::memcpy(pre4, b->pre_4.data(), b->pre_4.size() * 4);
        b->bh.n = nh;
        b->bh.hevc = (const rv::HevcPicDev* const*)dev((char*)ph);
        b->bh.prefix = (const int32_t*)dev((char*)preh);
        b->bh.ilevels = ilevels;
        b->bh.ipre = (const int32_t*)dev((char*)hipre);
        b->bh.ibase = (const int32_t*)dev((char*)hibase);
        b->bh.itask = (rv::IntraTask*)b->itask.p;
        b->bh.iall = iall;
        b->bh.itotal = b->itotal.data();
        b->bh.spre = (const int32_t*)dev((char*)hspre);
        b->bh.stotal = spre[(size_t)nh];
        b->b4.n = n4;
        b->b4.h264 = (const rv::H264PicDev* const*)dev((char*)p4);
        b->b4.prefix = (const int32_t*)dev((char*)pre4);
        b->b4.ilevels = ilevels4;
        b->b4.ipre = (const int32_t*)dev((char*)hipre4);
        b->b4.itotal = b->itotal4.data();
        // Written on the library stream, ahead of the batch's launches.
        hipError_t e = hipSuccess;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            const int nw = (int)(sz / 4);
            for (int w0 = 0; w0 < nw && e == hipSuccess; w0 += kFillWords) {
                WordChunk c;
                c.dst = (uint32_t*)b->tables.p + w0;
                c.n = std::min(kFillWords, nw - w0);
                std::memcpy(c.w, host.data() + w0, (size_t)c.n * 4);
                hipLaunchKernelGGL(rv_fill_words, dim3(1), dim3(256), 0, g_stream, c);
                e = hipGetLastError();
            }
            if (e == hipSuccess) (void)note_stream_work(&e);
            if (e != hipSuccess) {
                pool_put(b->tables);
                pool_put(b->itask);
            }
        }
        if (e != hipSuccess) {
            delete b;
            return hip_fail(e, "batch tables");
        }
        *out = b;
        return RV_OK;
    } catch (...) {
        return fail(RV_E_NOMEM, "exception in rv_batch_create");
    }
}

int rv_batch_free(rv_batch* b) {
    if (!b) return fail(RV_E_INVALID, "null batch");
    std::lock_guard<std::mutex> lk(g_mu);
    // A captured graph is destroyed only once it has finished; batches
    // launched without one free at once, their tables returning to the
    // pool when the work launched so far has completed.
    if (b->exec) {
        (void)hipStreamSynchronize(g_stream);
        (void)hipGraphExecDestroy(b->exec);
    }
    if (b->graph) (void)hipGraphDestroy(b->graph);
    pool_put(b->tables);
    pool_put(b->itask);
    delete b;
    return RV_OK;
}

// Caller holds g_mu. Unless chosen for this batch, the variants are the
// tuned ones of its depth class (if any), taken at each launch and capture.
static int launch_batched(rv_batch* b, hipStream_t st) {
    if (!b->var_set && g_tuned[b->depth_cls]) std::memcpy(b->var, g_tuned_var[b->depth_cls], sizeof(b->var));
    int nh = (int)b->hevc.size(), n4 = (int)b->h264.size();
    if (nh) {
        int rc = launch_codec(b->bh, b->pre_h, nh, true, st, b->var);
        if (rc) return rc;
    }
    if (n4) {
        int rc = launch_codec(b->b4, b->pre_4, n4, false, st);
        if (rc) return rc;
    }
    return RV_OK;
}

// Times every variant of each multi-variant HEVC stage that the batch's
// pictures allow, inside full reconstruction runs of the batch (stage 0
// rewrites every sample, so runs repeat exactly): one untimed run, then
// `reps` timed ones, the minimum of the stage's event time counting. Keeps
// the fastest per stage for the batch and its depth class. Caller holds
// g_mu; nothing is captured meanwhile.
static int tune_batch(rv_batch* b, int reps) {
    const int nh = (int)b->hevc.size();
    if (!nh) return RV_OK;
    int var[rv::kStages];
    std::memcpy(var, b->var, sizeof(var));
    hipEvent_t ev[2];
    RV_HIP(hipEventCreate(&ev[0]));
    if (hipEventCreate(&ev[1]) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return fail(RV_E_HIP, "hipEventCreate");
    }
    int rc = RV_OK;
    for (int s = 0; s < rv::kStages && rc == RV_OK; ++s) {
        const int nv = rv::hevc_stage_variants(s);
        if (nv <= 1) continue;
        int best = 0;
        float best_ms = 1e30f;
        for (int v = 0; v < nv && rc == RV_OK; ++v) {
            const int d = rv::hevc_variant_depth(s, v);
            if (d != 0 && d != b->depth_cls) continue;
            var[s] = v;
            float t_min = 1e30f;
            for (int r = 0; r <= reps && rc == RV_OK; ++r) {
                for (int k = 0; k < rv::kStages && rc == RV_OK; ++k) {
                    if (k == s && hipEventRecord(ev[0], g_stream) != hipSuccess) rc = fail(RV_E_HIP, "event record");
                    hipError_t e = rv::hevc_launch_stage(k, b->bh, b->pre_h[(size_t)k * (nh + 1) + nh], g_stream, var[k]);
                    if (e != hipSuccess) rc = hip_fail(e, "kernel launch");
                    if (k == s && rc == RV_OK && hipEventRecord(ev[1], g_stream) != hipSuccess)
                        rc = fail(RV_E_HIP, "event record");
                }
                if (rc == RV_OK && hipEventSynchronize(ev[1]) != hipSuccess) rc = fail(RV_E_HIP, "event wait");
                float ms = 0;
                if (rc == RV_OK && hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) rc = fail(RV_E_HIP, "event time");
                if (r > 0) t_min = std::min(t_min, ms);
            }
            if (t_min < best_ms) {
                best_ms = t_min;
                best = v;
            }
        }
        var[s] = best;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc) return rc;
    std::memcpy(b->var, var, sizeof(var));
    b->var_set = true;
    std::memcpy(g_tuned_var[b->depth_cls], var, sizeof(var));
    g_tuned[b->depth_cls] = true;
    return note_launch(b->hevc);
}

int rv_batch_tune(rv_batch* b, int reps) {
    try {
        if (!b) return fail(RV_E_INVALID, "null batch");
        if (reps < 1 || reps > 100) return fail(RV_E_RANGE, "reps outside 1..100");
        if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
        std::lock_guard<std::mutex> lk(g_mu);
        const int rc = tune_batch(b, reps);
        if (rc == RV_OK && b->exec) {  // re-captured at the next graph launch
            (void)hipStreamSynchronize(g_stream);
            (void)hipGraphExecDestroy(b->exec);
            (void)hipGraphDestroy(b->graph);
            b->exec = nullptr;
            b->graph = nullptr;
        }
        return rc;
    } catch (...) {
        return fail(RV_E_HIP, "exception in rv_batch_tune");
    }
}

int rv_batch_variants(const rv_batch* b, int out[5]) {
    if (!b || !out) return fail(RV_E_INVALID, "null argument");
    for (int s = 0; s < rv::kStages; ++s) out[s] = b->var[s];
    return RV_OK;
}

int rv_batch_set_variant(rv_batch* b, int stage, int variant) {
    if (!b) return fail(RV_E_INVALID, "null batch");
    if (stage < 0 || stage >= rv::kStages || variant < 0 || variant >= rv::hevc_stage_variants(stage))
        return fail(RV_E_RANGE, "stage or variant out of range");
    const int d = rv::hevc_variant_depth(stage, variant);
    if (d != 0 && (b->hevc.empty() || d != b->depth_cls))
        return fail(RV_E_RANGE, "variant not valid for the batch's sample depths");
    std::lock_guard<std::mutex> lk(g_mu);
    if (b->exec && b->var[stage] != variant) {  // re-captured at the next graph launch
        (void)hipStreamSynchronize(g_stream);
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
    }
    b->var[stage] = variant;
    b->var_set = true;
    return RV_OK;
}

int rv_variant_count(int codec, int stage) {
    if (stage < 0 || stage >= rv::kStages) return 0;
    if (codec == RV_CODEC_HEVC) return rv::hevc_stage_variants(stage);
    return codec == RV_CODEC_H264 ? 1 : 0;
}

const char* rv_variant_name(int codec, int stage, int variant) {
    if (codec == RV_CODEC_HEVC) return rv::hevc_variant_name(stage, variant);
    return codec == RV_CODEC_H264 && stage >= 0 && stage < rv::kStages && variant == 0 ? "any" : "";
}

int rv_batch_launch(rv_batch* b, int mode) {
    try {
        if (!b) return fail(RV_E_INVALID, "null batch");
        if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
        std::lock_guard<std::mutex> lk(g_mu);
        if (mode == RV_LAUNCH_PLAIN) {
            for (rv_job* j : b->jobs) {
                std::vector<int32_t> pre((size_t)rv::kStages * 2);
                const int32_t* items = j->codec == RV_CODEC_HEVC ? j->hd.items : j->h4.items;
                for (int s = 0; s < rv::kStages; ++s) {
                    pre[s * 2] = 0;
                    pre[s * 2 + 1] = items[s];
                }
                const int cls = j->codec == RV_CODEC_HEVC ? depth_class_of(j->hd.bd) : 0;
                int rc = launch_codec(j->single, pre, 1, j->codec == RV_CODEC_HEVC, g_stream,
                                      g_tuned[cls] ? g_tuned_var[cls] : nullptr);
                if (rc) return rc;
            }
            return note_launch(b->jobs);
        }
        if (mode == RV_LAUNCH_BATCHED) {
            if (!b->var_set && tuning_enabled() && !b->hevc.empty() && !g_tuned[b->depth_cls]) {
                const int rc = tune_batch(b, 2);
                if (rc) return rc;
            }
            int rc = launch_batched(b, g_stream);
            return rc ? rc : note_launch(b->jobs);
        }
        if (mode == RV_LAUNCH_GRAPH) {
            if (!b->exec && !b->var_set && tuning_enabled() && !b->hevc.empty() && !g_tuned[b->depth_cls]) {
                const int rc = tune_batch(b, 3);
                if (rc) return rc;
            }
            if (!b->exec) {
                RV_HIP(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
                int rc = launch_batched(b, g_stream);
                hipGraph_t graph = nullptr;
                hipError_t e = hipStreamEndCapture(g_stream, &graph);
                if (rc || e != hipSuccess) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return rc ? rc : hip_fail(e, "hipStreamEndCapture");
                }
                b->graph = graph;  // destroyed with the batch from here on
                RV_HIP(hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0));
            }
            RV_HIP(hipGraphLaunch(b->exec, g_stream));
            return note_launch(b->jobs);
        }
        return fail(RV_E_INVALID, "unknown launch mode");
    } catch (...) {
        return fail(RV_E_HIP, "exception in rv_batch_launch");
    }
}

int rv_batch_profile(rv_batch* b, int reps, float out_ms[5]) {
    try {
        if (!b || !out_ms || reps <= 0) return fail(RV_E_INVALID, "bad argument");
        if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
        std::lock_guard<std::mutex> lk(g_mu);
        hipEvent_t ev[rv::kStages + 1] = {};
        auto destroy = on_exit([&] {
            for (auto& e : ev)
                if (e) (void)hipEventDestroy(e);
        });
        for (auto& e : ev) RV_HIP(hipEventCreate(&e));
        double acc[rv::kStages] = {0, 0, 0, 0, 0};
        int rc = RV_OK;
        for (int r = 0; r < reps && rc == RV_OK; ++r) {
            RV_HIP(hipEventRecord(ev[0], g_stream));
            for (int s = 0; s < rv::kStages && rc == RV_OK; ++s) {
                int nh = (int)b->hevc.size(), n4 = (int)b->h264.size();
                if (nh) {
                    hipError_t e =
                        rv::hevc_launch_stage(s, b->bh, b->pre_h[(size_t)s * (nh + 1) + nh], g_stream, b->var[s]);
                    if (e != hipSuccess) rc = hip_fail(e, "kernel launch");
                }
                if (n4 && rc == RV_OK) {
                    hipError_t e = rv::h264_launch_stage(s, b->b4, b->pre_4[(size_t)s * (n4 + 1) + n4], g_stream);
                    if (e != hipSuccess) rc = hip_fail(e, "kernel launch");
                }
                RV_HIP(hipEventRecord(ev[s + 1], g_stream));
            }
            RV_HIP(hipEventSynchronize(ev[rv::kStages]));
            for (int s = 0; s < rv::kStages; ++s) {
                float ms = 0;
                RV_HIP(hipEventElapsedTime(&ms, ev[s], ev[s + 1]));
                acc[s] += ms;
            }
        }
        for (int s = 0; s < rv::kStages; ++s) out_ms[s] = (float)(acc[s] / reps);
        const int rn = note_launch(b->jobs);
        return rc ? rc : rn;
    } catch (...) {
        return fail(RV_E_HIP, "exception in rv_batch_profile");
    }
}

int rv_probe_copy(int64_t bytes, int reps, double* gbs) {
    try {
        if (!gbs || bytes < 4096 || reps <= 0) return fail(RV_E_INVALID, "bad argument");
        if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
        bytes &= ~(int64_t)4095;
        void *a = nullptr, *b = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto release = on_exit([&] {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
        });
        if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess)
            return fail(RV_E_NOMEM, "probe buffers");
        std::lock_guard<std::mutex> lk(g_mu);
        RV_HIP(hipEventCreate(&e0));
        RV_HIP(hipEventCreate(&e1));
        RV_HIP(hipMemsetAsync(a, 1, bytes, g_stream));
        RV_HIP(rv::launch_copy_probe(b, a, bytes, g_stream));
        RV_HIP(hipEventRecord(e0, g_stream));
        for (int r = 0; r < reps; ++r) RV_HIP(rv::launch_copy_probe(b, a, bytes, g_stream));
        RV_HIP(hipEventRecord(e1, g_stream));
        RV_HIP(hipEventSynchronize(e1));
        float ms = 0;
        RV_HIP(hipEventElapsedTime(&ms, e0, e1));
        *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
        return RV_OK;
    } catch (...) {
        return fail(RV_E_HIP, "exception in rv_probe_copy");
    }
}

int rv_sync(void) {
    if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
    RV_HIP(hipStreamSynchronize(g_stream));
    return check_device_errors();
}

int rv_mark(int* id_out) {
    if (!id_out) return fail(RV_E_INVALID, "null out");
    if (!ready()) return fail(RV_E_NODEVICE, "rv_init not called");
    std::lock_guard<std::mutex> lk(g_mu);
    int id = g_next_mark;
    g_next_mark = (g_next_mark + 1) % kMarks;
    RV_HIP(hipEventRecord(g_marks[id], g_stream));
    *id_out = id;
    return RV_OK;
}

int rv_mark_query(int id) {
    if (id < 0 || id >= kMarks || !ready()) return fail(RV_E_INVALID, "bad marker");
    hipError_t e = hipEventQuery(g_marks[id]);
    if (e == hipSuccess) return 0;
    if (e == hipErrorNotReady) return 1;
    return hip_fail(e, "hipEventQuery");
}

int rv_mark_wait(int id) {
    if (id < 0 || id >= kMarks || !ready()) return fail(RV_E_INVALID, "bad marker");
    RV_HIP(hipEventSynchronize(g_marks[id]));
    return check_device_errors();
}

int rv_mark_elapsed(int id_start, int id_end, float* ms) {
    if (!ms || id_start < 0 || id_start >= kMarks || id_end < 0 || id_end >= kMarks || !ready())
        return fail(RV_E_INVALID, "bad marker");
    RV_HIP(hipEventElapsedTime(ms, g_marks[id_start], g_marks[id_end]));
    return RV_OK;
}

}  // extern "C"
