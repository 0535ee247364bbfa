// Device-side descriptors shared by the HIP kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/rivet_hip.h"

namespace rv {

// One motion-compensation block of one prediction unit, carrying the PU's
// motion: either a 16x16 tile of the PU's 16-aligned core (one per wave,
// one shared reference window) or an at most 8x8 block of what remains (a
// wave serves four), each with its two 4:2:0 chroma blocks. Splitting PUs
// on the host keeps work of even size.
struct alignas(16) McTile {
    uint16_t x, y;    // block top-left, luma samples
    uint8_t tw, th;   // block size, luma samples (16, or 4 or 8)
    uint8_t dir;      // bit 0: list 0, bit 1: list 1
    uint8_t refs;     // ref_idx of list 0 in bits 0..3, list 1 in bits 4..7
    rv_mv mv[2];
};
static_assert(sizeof(McTile) == 16, "one 16-byte load per tile");

// One strip of motion compensation: 4 output columns (2 for the chroma of
// 4-wide blocks) by H rows of one component of a block whose reference
// windows all lie inside the reference planes. A lane computes a strip's
// whole first stage (H + 7 luma or H + 3 chroma rows) in registers, so no
// window is shared and no sample is clamped. Strips are grouped by kind,
// height and list use; each group's strips of one band of the picture are
// padded to whole waves with copies of their first strip that have
// cols = 0.
struct alignas(16) McStrip {
    uint16_t x, y;  // top-left output sample, in samples of the component
    uint8_t cols;   // output columns: 4, 2, or 0 for padding
    uint8_t dir;    // bit 0: list 0, bit 1: list 1
    uint8_t refs;   // ref_idx of list 0 in bits 0..3, list 1 in bits 4..7
    uint8_t comp;   // 0 luma, 1 Cb, 2 Cr
    rv_mv mv[2];    // the PU's luma quarter-sample vectors
};
static_assert(sizeof(McStrip) == 16, "one 16-byte load per strip");
// Groups: luma heights 16, 8, 4 then chroma heights 8, 4, 2, each bi then
// uni (group = kind * 6 + height class * 2 + (uni ? 1 : 0)).
constexpr int kStripGroups = 12;
// Coded TUs per stage-1 work item (one wave) of each size class (4, 8,
// 16, 32): 64 / TPW lanes per TU.
__host__ __device__ constexpr int tus_per_item(int cls) { return cls == 0 ? 64 : cls == 1 ? 32 : cls == 2 ? 16 : 4; }
// Bands of the stage-1 item order and of the side-data classes: 1 <<
// kTxBandShift luma rows.
constexpr int kTxBandShift = 7;

// One transform block of an RV_INTRA_DEVICE CU: the TU (coef_off rebased
// into the job's nonzero list), its intra prediction mode, where its
// residual goes and its dependency level.
struct alignas(16) IntraTu {
    rv_hevc_tu t;
    int32_t mode;   // IntraPredModeY of its PU, or IntraPredModeC
    int32_t res;    // coded: first int16 of its N x N residual in ires; else -1
    int32_t level;  // dependency level, from 0
    int32_t pad;
};
static_assert(sizeof(IntraTu) == 32, "two 16-byte loads per intra TU");

// Side data (the per-4x4 maps) is written per cell: the CUs, PUs and luma
// TUs of one shape form a class (consecutive in their sorted lists), and
// one work item covers kSideCellsPerItem consecutive cells of a class,
// lane by lane, so a row of an entry is written by adjacent lanes.
enum SideKind : int32_t { SIDE_CU, SIDE_PU, SIDE_TU };
struct SideClass {
    int32_t kind;    // SideKind
    int32_t first;   // index of the class's first entry in its list
    int32_t count;   // entries
    int32_t cw, ch;  // entry size in 4x4 cells
    int32_t item0;   // the class's first work item
};
constexpr int kSideCellsPerItem = 256;
// Classes, per band of 1 << kTxBandShift luma rows: CU sizes 64..8, PU
// shapes (w, h in 4..64), luma TU sizes 32..4 (the empty ones left out).

// Per-picture descriptor of an HEVC job, resident in device memory.
// Flags of the per-CTB loop-filter words (HevcPicDev::ctbw).
constexpr uint32_t kCtbDbkOff = 1u << 8;         // the CTB's slice has deblocking disabled
constexpr uint32_t kCtbLeftOff = 1u << 9;        // no filtering across the left CTB boundary
constexpr uint32_t kCtbTopOff = 1u << 10;        // nor across the top one
constexpr uint32_t kCtbSaoPerSample = 1u << 11;  // SAO per sample: unusable neighbour or exempt block
// Exemption bits of HevcPicDev::cuinfo (set only with ctbw).
constexpr uint8_t kCuSaoExempt = 8;   // filter_exempt bit 0: SAO and deblocking leave it alone
constexpr uint8_t kCuDbkExempt = 16;  // filter_exempt & 5: deblocking leaves it alone

struct HevcPicDev {
    int32_t w, h, bd, log2_ctb, wc, hc, w4, h4;
    int32_t pitch[3];  // samples
    int32_t pw[3], ph[3];
    void* cur[3];  // reconstruction target (deblocked in place)
    void* out[3];  // SAO output: the job's DPB slot
    const void* ref[2][RV_MAX_REFS][3];
    int32_t weighted, log2_wd[3];
    int32_t wp_w[2][RV_MAX_REFS][3], wp_o[2][RV_MAX_REFS][3];
    int32_t ref_delta[2][RV_MAX_REFS], ref_lt[2][RV_MAX_REFS];
    int32_t deblock, beta_offset, tc_offset, cb_qp_offset, cr_qp_offset;
    int32_t sao_luma, sao_chroma;
    // Work lists.
    int32_t n_cus, n_pus, n_ltus, n_tiles;  // n_tiles: small blocks, a multiple of 4
    int32_t n_tiles16;                      // 16x16 tiles
    int32_t n_side, n_intra;  // stage-0 side-data items, intra CUs
    // Side-data classes band by band (the cells of one band are written
    // close in time, so their partial lines meet in L2); sidx[i] is the
    // class of side item i.
    int32_t n_sclass;
    const SideClass* sclass;
    const uint16_t* sidx;
    // Per-CTB loop-filter words (raster CTBs) of pictures with several
    // slices or tiles or with exempt CUs, null otherwise: bits 0..7 the
    // neighbouring CTBs SAO may read (l, r, t, b, tl, tr, bl, br), then
    // the kCtb* flags, the slice's beta_offset (bits 16..23) and tc_offset
    // (bits 24..31) as signed bytes. With them, cuinfo carries the
    // exemption bits kCuSaoExempt / kCuDbkExempt.
    const uint32_t* ctbw;
    const int32_t* intra_idx; // indices of the intra CUs in cus
    // Stage-1 transform items: item i takes ctus[first .. first + count)
    // of one size class cls (4, 8, 16, 32; 64 / N TUs per item), packed as
    // txw[i] = first | (count - 1) << 24 | cls << 30.
    int32_t n_txw;
    const uint32_t* txw;
    const rv_hevc_cu* cus;
    const rv_hevc_pu* pus;
    const rv_hevc_tu* ltus;  // luma TUs (coded or not): side data
    const McTile* tiles;     // small blocks
    const McTile* tiles16;   // 16x16 tiles
    // Strips of the blocks whose windows lie inside the reference planes
    // (the tiles above hold the others), 64 per wave: wave w takes
    // strips[64 w ..] of group swg[w]. Waves go band by band of 128 luma
    // rows, all groups of a band before the next band, so a band's
    // reference lines are reused from L2 by every group.
    const McStrip* strips;
    const uint8_t* swg;
    int32_t n_swaves;
    // Coded TUs band by band (128 luma rows; chroma TUs by the luma rows
    // they cover), by size class within a band, list order within a
    // class: the items of one band run together, so the partial lines of
    // neighbouring TUs of different sizes meet in L2 rather than going to
    // memory once per TU.
    const rv_hevc_tu* ctus;
    const rv_nz* nz;         // nonzero coefficients of ctus, in ctus order
    const uint16_t* intra_pred;
    const rv_sao* sao;
    const rv_hevc_scaling* scaling;  // levels' scaling lists; null = flat 16
    // Intra prediction on the device: the TUs of RV_INTRA_DEVICE CUs by
    // dependency level; level k holds itus[ilevel[k] .. ilevel[k + 1]).
    int32_t n_ilevels;
    const int32_t* ilevel;
    const IntraTu* itus;
    int32_t strong, cip;  // strong intra smoothing, constrained intra pred
    int16_t* ires;        // residuals of the coded itus (job scratch)
    // Per-4x4 side data (job scratch).
    uint8_t* cuinfo;   // bit 0 intra, bit 1 CU left edge, bit 2 CU top edge
    int8_t* qp;        // QpY
    uint32_t* pumap;   // PU index (bits 0..29), bit 30 PU left edge, bit 31 top
    uint8_t* tuinfo;   // bit 0 cbf, bit 1 TU left edge, bit 2 TU top edge
    uint8_t* bs;       // bits 0..1 bS of the left edge, bits 2..3 top edge
    // Item counts per stage (derived).
    int32_t items[5];
};

// Deblocking parameters of one H.264 macroblock: per edge and 4-line
// segment one word alpha | beta << 8 | bS-4 flag << 16 | s << 24, s = 0
// where the segment is left alone, else tC0 + 1. Luma vertical edges as
// [segment][edge 0..3], luma horizontal likewise, then chroma vertical as
// [plane][segment][edge 0, 4], chroma horizontal likewise.
constexpr int kH264DbkWords = 64;

// Macroblock rows per deblocking band (one workgroup each).
// Macroblock rows per band of the H.264 deblocking wavefront (one
// workgroup of 32 threads per row).
constexpr int kH264DbkRows = 16;
__host__ __device__ inline int h264_deblock_bands(int mbh) { return (mbh + kH264DbkRows - 1) / kH264DbkRows; }

// One RV_INTRA_DEVICE H.264 macroblock: its address, which neighbouring
// macroblocks may supply samples (bit 0 left A, 1 above B, 2 above-right
// C, 3 above-left D: present and, under constrained intra prediction,
// intra) and its residual blocks iblks[blk0 .. blk0 + nblk).
struct alignas(16) H264IntraMb {
    int32_t addr;
    uint32_t avail;
    int32_t blk0, nblk;
};

// Per-picture descriptor of an H.264 job.
struct H264PicDev {
    int32_t w, h, mbw, mbh, w4, h4;
    int32_t pitch[3];
    int32_t pw[3], ph[3];
    void* cur[3];
    void* out[3];
    const void* ref[2][RV_MAX_REFS][3];
    int32_t ref_id[2][RV_MAX_REFS];
    int32_t weighting;  // RV_H264_*
    int32_t log_wd[3];
    int32_t wp_w[2][RV_MAX_REFS][3], wp_o[2][RV_MAX_REFS][3];
    int32_t imp_w[RV_MAX_REFS][RV_MAX_REFS][2];  // implicit (w0, w1)
    int32_t deblock, offset_a, offset_b;
    int32_t n_mbs, n_parts, n_blks4, n_blks8;
    const rv_h264_mb* mbs;
    const rv_h264_part* parts;
    // Partition indices, above 8 x 8 first (n_big), then the rest; then
    // the indices of the n_intra intra macroblocks.
    const uint32_t* part_order;
    int32_t n_big, n_intra;
    const rv_h264_blk* blks;  // 4x4 first, then 8x8 (not of RV_INTRA_DEVICE macroblocks)
    const int32_t* coefs;
    // Macroblocks predicted on the device, by dependency level: level k
    // holds imbs[ilevel[k] .. ilevel[k + 1]); their blocks in iblks.
    int32_t n_ilevels;
    const int32_t* ilevel;
    const H264IntraMb* imbs;
    const rv_h264_blk* iblks;
    const uint16_t* intra_pred;
    const rv_h264_scaling* scaling;  // levels' scaling lists; null = flat 16
    uint32_t* partmap;   // per 4x4: partition index (0xFFFFFFFF = intra)
    uint32_t* dbk;       // per MB: kH264DbkWords words of edge parameters
    uint64_t* xch;       // deblocking: band ticket, then the strips crossing bands
    int64_t n_xch;       // words of xch
    // The exchange poll's bound, and a value XORed into the awaited tag (0;
    // non-zero only where a test forces every wait to give up). A wait that
    // gives up adds 1 to *err, the library's device error counter, which
    // the host reads at its next synchronisation and reports.
    int32_t xch_spin;
    uint32_t xch_tag_xor;
    uint32_t* err;
    int32_t items[5];
};

// Work item of this workgroup: consecutive workgroups go round-robin to the
// eight XCDs, so workgroup b takes item (b % 8) * chunk + b / 8 — each XCD
// walks one contiguous run of items (in a batch, mostly one picture) and
// its L2 sees that run's reference windows and output lines. The grid is
// 8 * chunk; items past `total` exit.
__device__ __forceinline__ int xcd_item() {
    const int chunk = gridDim.x >> 3;
    return (int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3);
}

// One call of one kernel-table entry (the per-entry API). Pointers are
// device copies of the caller's spans, rebased so the entry's own offsets
// (off, origin) index them as the caller's slices would.
enum HevcOp : int32_t {
    HE_IDCT, HE_IDST, HE_ADD_RES, HE_COPY, HE_FILT_H, HE_FILT_V, HE_FILT_V2, HE_UNI, HE_BI, HE_WUNI,
    HE_WBI, HE_FUSED, HE_SAO_BAND, HE_SAO_EDGE, HE_DBK_LUMA, HE_DBK_CHROMA
};
struct HevcEntry {
    int32_t op, ss;       // HevcOp; bytes per sample (1, 2)
    int32_t luma;         // 8-tap (1) or 4-tap (0) filters
    int32_t vertical;     // deblocking: _v (1) or _h (0)
    void* dst;
    const void* src;
    const int16_t* other; // bi: the second 14-bit input
    int64_t dst_stride, src_stride;
    int64_t off;          // deblocking: q0 of line 0; SAO edge: origin
    int64_t na, nb;       // SAO edge neighbour offsets
    int32_t w, h;         // block size (idct: n)
    int32_t frac, fx, fy, shift, max, bd;
    int32_t log2_wd, w0, w1, o0, o1;
    int32_t max_x, max_y;
    int32_t beta[2], tc[4];
    uint8_t no_p[4], no_q[4];
    int16_t table[32];    // SAO band table / edge offsets (5)
};

enum H264Op : int32_t {
    HA_QPEL, HA_CHROMA, HA_COPY, HA_AVG, HA_WUNI, HA_WBI, HA_DBK, HA_IDCT4, HA_IDCT8, HA_DC, HA_RES4, HA_RES8
};
struct H264Entry {
    int32_t op, pos;
    uint8_t* dst;
    const uint8_t* src;   // or the first operand
    const uint8_t* b;     // second operand (avg, weighted_bi)
    const void* coefs;    // int16 (idct*_add) or int32 (residual*)
    int64_t stride, src_stride, off;
    int32_t w, h, xf, yf, max;
    int32_t log_wd, w0, w1, o0, o1;
    int32_t alpha, beta, chroma, vertical, intra;
    int32_t dc, has_dc, n;
    int16_t tc0[4];
};

// Per-batch tables: the descriptors and, per stage, prefix sums of the
// per-picture work-item counts.
constexpr int kStages = 5;
struct BatchDev {
    int32_t n;
    const HevcPicDev* const* hevc;  // per picture (null for H.264)
    const H264PicDev* const* h264;
    const int32_t* prefix;          // kStages x (n + 1)
    // HEVC motion-compensation strips: prefix over pictures of their waves
    // (device memory, n + 1) and the total (host: read by the launcher).
    const int32_t* spre;
    int32_t stotal;
    // Device intra (HEVC): the most dependency levels of any picture; the
    // prefix over pictures of their TU counts, all levels first, then per
    // level (device memory, (levels + 1) x (n + 1)); the first task of each
    // level in itask (device, levels + 1); the task table, level by level,
    // written by the preparation kernel (device); the TUs of all levels and
    // each level's count (host: read by the launcher only).
    int32_t ilevels;
    const int32_t* ipre;
    const int32_t* ibase;
    struct IntraTask* itask;
    int32_t iall;
    const int32_t* itotal;
};

// Task of a device-intra level kernel: one TU of one picture.
struct alignas(16) IntraTask {
    const HevcPicDev* pic;
    const IntraTu* tu;
};
static_assert(sizeof(IntraTask) == 16, "one 16-byte load per task");

}  // namespace rv
