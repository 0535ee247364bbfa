// Seeded synthetic work lists (see include/rivet_gen.h).
//
// Geometry follows what a conforming stream can carry, because the kernels
// and the reference's drivers assume it: HEVC CUs split where they cross
// the picture edge, AMP only on CUs of 16 and up, no bi-prediction on 8x4 /
// 4x8 PUs, a 64x64 CU's luma transform tree starts at 32x32, and a luma
// 8x8 split into four 4x4 TUs keeps one 4x4 chroma TU per component (the
// 4:2:0 rule). H.264: 8x8 transforms only when every partition is 8x8 or
// larger.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/rivet_gen.h"

namespace {

uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// The reference test LCG (src/dsp/hevc.rs:1444-1449): 31 bits per step.
struct Lcg {
    uint64_t s;
    explicit Lcg(uint64_t seed) : s(seed) {}
    uint32_t next() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (uint32_t)(s >> 33);
    }
    int uniform(int lo, int hi) { return lo + (int)(next() % (uint32_t)(hi - lo + 1)); }
    bool permille(int p) { return (int)(next() % 1000u) < p; }
    bool percent(int p) { return (int)(next() % 100u) < p; }
    // Discrete Laplacian magnitude >= 1: P(m = k) proportional to e^(-k/b),
    // by Bernoulli trials against round(2^31 e^(-1/b)).
    int laplace(int b) {
        uint32_t thr;
        switch (b) {
            case 8: thr = 1895147668u; break;
            case 12: thr = 1975780336u; break;
            case 16: thr = 2017374191u; break;
            default: thr = 1895147668u; break;
        }
        int m = 1;
        while (m < 32767 && next() < thr) ++m;
        return m;
    }
};

enum Purpose : uint64_t { STRUCT = 1, COEF = 2, SAO = 3, INTRA = 4, WEIGHTS = 5, SCALING = 6, SLICES = 8 };

// Seeded scaling factors, 4..64 each.
void fill_factors(Lcg& r, uint8_t* p, size_t n) {
    for (size_t i = 0; i < n; ++i) p[i] = (uint8_t)r.uniform(4, 64);
}

// QpC as a function of qPi for 4:2:0 (Table 8-10 of H.265).
int hevc_chroma_qp(int qpi) {
    static const int T[14] = {29, 30, 31, 32, 33, 33, 34, 34, 35, 35, 36, 36, 37, 37};
    if (qpi < 30) return qpi;
    if (qpi > 43) return qpi - 6;
    return T[qpi - 30];
}

}  // namespace

struct rv_gen {
    int codec;
    rv_hevc_pic hevc;
    rv_h264_pic h264;
    std::vector<rv_hevc_cu> cus;
    std::vector<rv_hevc_pu> pus;
    std::vector<rv_hevc_tu> tus;
    std::vector<rv_nz> nz;
    std::vector<uint16_t> intra_pred;
    std::vector<rv_sao> sao;
    std::vector<rv_h264_mb> mbs;
    std::vector<rv_h264_part> parts;
    std::vector<rv_h264_blk> blks;
    std::vector<int32_t> coefs32;
    rv_hevc_scaling hevc_scaling;
    rv_h264_scaling h264_scaling;
    std::vector<rv_hevc_slice> slices;
    std::vector<uint16_t> ctb_slice, ctb_tile;
    std::vector<uint32_t> ctb_slice_addr, ctb_ts;
};

namespace {

struct HevcBuilder {
    const rv_gen_hevc_cfg& c;
    rv_gen& g;
    Lcg rs, rc, ri;
    int W, H, max;

    HevcBuilder(const rv_gen_hevc_cfg& cfg, rv_gen& gen, uint64_t seed)
        : c(cfg), g(gen), rs(rv_gen_subseed(seed, 0, 0, STRUCT)), rc(rv_gen_subseed(seed, 0, 0, COEF)),
          ri(rv_gen_subseed(seed, 0, 0, INTRA)), W(cfg.width), H(cfg.height), max((1 << cfg.bit_depth) - 1) {}

    // Coefficients of one n x n TU (nonzero list, raster order): returns
    // whether any is nonzero.
    bool coefs(int n, rv_hevc_tu& tu) {
        const uint32_t off = (uint32_t)g.nz.size();
        int mx = 0, my = 0;
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                if (!rc.permille(c.nz_permille)) continue;
                int m = rc.laplace(c.lap_b);
                int v = (rc.next() & 1) ? -m : m;
                g.nz.push_back(rv_nz{(uint16_t)(y * n + x), (int16_t)v});
                if (x > mx) mx = x;
                if (y > my) my = y;
            }
        const uint32_t cnt = (uint32_t)g.nz.size() - off;
        tu.coef_off = cnt ? off : 0;
        tu.n_nz = (uint16_t)cnt;
        tu.max_x = (uint8_t)(cnt ? mx : 0);
        tu.max_y = (uint8_t)(cnt ? my : 0);
        return cnt != 0;
    }

    void add_tu(int comp, int x, int y, int log2, bool dst, bool always, bool intra, int qp_y) {
        rv_hevc_tu tu;
        std::memset(&tu, 0, sizeof(tu));
        tu.x = (uint16_t)x;
        tu.y = (uint16_t)y;
        tu.comp = (uint8_t)comp;
        tu.log2 = (uint8_t)log2;
        bool coded = coefs(1 << log2, tu);
        tu.flags = (uint8_t)((dst ? RV_TU_DST : 0) | (coded ? RV_TU_CODED : 0));
        if (coded && c.levels) {
            // qP of the TU: Qp'Y, or Qp'C from qPi = Clip3(-QpBdOffsetC, 57,
            // QpY + pps_cb_qp_offset) (the generated pictures' offsets are 0).
            const int off = 6 * (c.bit_depth - 8);
            const int qpi = qp_y < -off ? -off : (qp_y > 57 ? 57 : qp_y);
            tu.qp = (uint8_t)((comp ? hevc_chroma_qp(qpi) : qp_y) + off);
            tu.flags |= (uint8_t)(RV_TU_LEVELS | (intra ? RV_TU_INTRA : 0));
        }
        if (coded || always) g.tus.push_back(tu);
    }

    // Luma TU geometry of a CU: leaves as (x, y, log2); reports a 4x4 split.
    void tu_tree(int x, int y, int log2, std::vector<int>& leaves) {
        bool split = log2 > 2 && rs.percent(50);
        if (!split) {
            leaves.push_back(x);
            leaves.push_back(y);
            leaves.push_back(log2);
            return;
        }
        int h = 1 << (log2 - 1);
        for (int k = 0; k < 4; ++k) tu_tree(x + (k & 1) * h, y + (k >> 1) * h, log2 - 1, leaves);
    }

    void cu(int x, int y, int log2) {
        int s = 1 << log2;
        if (x >= W || y >= H) return;
        bool must = x + s > W || y + s > H;
        if (log2 > 3 && (must || rs.percent(50))) {
            int h = s / 2;
            for (int k = 0; k < 4; ++k) cu(x + (k & 1) * h, y + (k >> 1) * h, log2 - 1);
            return;
        }
        rv_hevc_cu cu{};
        std::memset(&cu, 0, sizeof(cu));
        cu.x = (uint16_t)x;
        cu.y = (uint16_t)y;
        cu.log2_size = (uint8_t)log2;
        cu.qp_y = (int8_t)rs.uniform(c.qp_min, c.qp_max);
        // Transform tree: from min(CU, 32) (a 64x64 CU is four 32x32 TUs).
        std::vector<int> leaves;
        int tl = log2 > 5 ? 5 : log2;
        for (int ty = 0; ty < s; ty += 1 << tl)
            for (int tx = 0; tx < s; tx += 1 << tl) tu_tree(x + tx, y + ty, tl, leaves);
        bool has4 = false, all4 = true;
        for (size_t i = 0; i < leaves.size(); i += 3) {
            if (leaves[i + 2] == 2) has4 = true;
            else all4 = false;
        }
        bool intra;
        if (c.intra_device) {
            intra = rs.percent(c.intra_percent);
            cu.intra = intra ? RV_INTRA_DEVICE : 0;
        } else {
            intra = log2 == 3 && has4 && rs.percent(c.intra_percent);
            cu.intra = intra ? RV_INTRA_HOST : 0;
        }
        if (intra && c.intra_device) {
            const bool nxn = log2 == 3 && all4;
            cu.part_mode = nxn ? 3 : 0;
            for (int k = 0; k < (nxn ? 4 : 1); ++k) cu.luma_mode[k] = (uint8_t)rs.uniform(0, 34);
            cu.chroma_mode = (uint8_t)rs.uniform(0, 34);
        } else if (intra) {
            cu.part_mode = 0;
            cu.pred_off = (uint32_t)g.intra_pred.size();
            for (int i = 0; i < s * s + 2 * (s / 2) * (s / 2); ++i) g.intra_pred.push_back((uint16_t)ri.uniform(0, max));
        } else {
            // Partitions: 2Nx2N, 2NxN, Nx2N, and AMP from 16 up.
            int nmodes = log2 >= 4 ? 7 : 3;
            int pm = rs.uniform(0, nmodes - 1);
            static const int PMAP[7] = {0, 1, 2, 4, 5, 6, 7};  // PartMode codes (NxN = 3 unused)
            cu.part_mode = (uint8_t)PMAP[pm];
            int q = s / 4, hs = s / 2;
            int geo[2][4];
            int n = 2;
            switch (pm) {
                case 0: n = 1; geo[0][0] = 0, geo[0][1] = 0, geo[0][2] = s, geo[0][3] = s; break;
                case 1: geo[0][0] = 0, geo[0][1] = 0, geo[0][2] = s, geo[0][3] = hs;
                        geo[1][0] = 0, geo[1][1] = hs, geo[1][2] = s, geo[1][3] = hs; break;
                case 2: geo[0][0] = 0, geo[0][1] = 0, geo[0][2] = hs, geo[0][3] = s;
                        geo[1][0] = hs, geo[1][1] = 0, geo[1][2] = hs, geo[1][3] = s; break;
                case 3: geo[0][0] = 0, geo[0][1] = 0, geo[0][2] = s, geo[0][3] = q;  // 2NxnU
                        geo[1][0] = 0, geo[1][1] = q, geo[1][2] = s, geo[1][3] = s - q; break;
                case 4: geo[0][0] = 0, geo[0][1] = 0, geo[0][2] = s, geo[0][3] = s - q;  // 2NxnD
                        geo[1][0] = 0, geo[1][1] = s - q, geo[1][2] = s, geo[1][3] = q; break;
                case 5: geo[0][0] = 0, geo[0][1] = 0, geo[0][2] = q, geo[0][3] = s;  // nLx2N
                        geo[1][0] = q, geo[1][1] = 0, geo[1][2] = s - q, geo[1][3] = s; break;
                default: geo[0][0] = 0, geo[0][1] = 0, geo[0][2] = s - q, geo[0][3] = s;  // nRx2N
                         geo[1][0] = s - q, geo[1][1] = 0, geo[1][2] = q, geo[1][3] = s; break;
            }
            for (int k = 0; k < n; ++k) {
                rv_hevc_pu pu;
                std::memset(&pu, 0, sizeof(pu));
                pu.x = (uint16_t)(x + geo[k][0]);
                pu.y = (uint16_t)(y + geo[k][1]);
                pu.w = (uint8_t)geo[k][2];
                pu.h = (uint8_t)geo[k][3];
                bool small = pu.w + pu.h == 12;  // 8x4 / 4x8
                int dir;
                if (!c.b_picture) dir = 1;
                else if (!small && rs.percent(c.bi_percent)) dir = 3;
                else dir = rs.percent(50) ? 1 : 2;
                pu.dir = (uint8_t)dir;
                for (int l = 0; l < 2; ++l) {
                    if ((dir >> l) & 1) {
                        pu.ref_idx[l] = 0;
                        pu.mv[l].x = (int16_t)rs.uniform(-c.mv_range, c.mv_range);
                        pu.mv[l].y = (int16_t)rs.uniform(-c.mv_range, c.mv_range);
                    } else {
                        pu.ref_idx[l] = -1;
                    }
                }
                g.pus.push_back(pu);
            }
        }
        g.cus.push_back(cu);
        // Transform units: luma (all, coded or not), chroma (coded only).
        for (size_t i = 0; i < leaves.size(); i += 3) {
            int tx = leaves[i], ty = leaves[i + 1], l2 = leaves[i + 2];
            // Device-intra CUs list every chroma TU, coded or not.
            const bool all_c = intra && c.intra_device;
            add_tu(0, tx, ty, l2, intra && l2 == 2, true, intra, cu.qp_y);
            if (l2 > 2) {
                add_tu(1, tx / 2, ty / 2, l2 - 1, false, all_c, intra, cu.qp_y);
                add_tu(2, tx / 2, ty / 2, l2 - 1, false, all_c, intra, cu.qp_y);
            } else if ((tx & 7) == 0 && (ty & 7) == 0) {
                // 4:2:0: the four 4x4 luma TUs of an 8x8 share one 4x4
                // chroma TU per component (carried with the first).
                add_tu(1, tx / 2, ty / 2, 2, false, all_c, intra, cu.qp_y);
                add_tu(2, tx / 2, ty / 2, 2, false, all_c, intra, cu.qp_y);
            }
        }
    }

    void run(uint64_t seed) {
        rv_hevc_pic& p = g.hevc;
        std::memset(&p, 0, sizeof(p));
        p.width = W;
        p.height = H;
        p.bit_depth = c.bit_depth;
        p.log2_ctb = 6;
        p.poc = c.poc;
        p.num_ref[0] = 1;
        p.num_ref[1] = c.b_picture ? 1 : 0;
        for (int l = 0; l < 2; ++l) {
            p.ref_slot[l][0] = c.ref_slot[l];
            p.ref_poc[l][0] = c.ref_poc[l];
        }
        if (c.explicit_wp) {
            Lcg rw(rv_gen_subseed(seed, 0, 0, WEIGHTS));
            int bd = c.bit_depth;
            int shift1 = 14 - bd > 2 ? 14 - bd : 2;
            p.weighted = 1;
            for (int comp = 0; comp < 3; ++comp) p.log2_wd[comp] = 6 + shift1;
            for (int l = 0; l < 2; ++l)
                for (int comp = 0; comp < 3; ++comp) {
                    p.wp_w[l][0][comp] = rw.uniform(32, 96);
                    p.wp_o[l][0][comp] = rw.uniform(-32, 32) * (1 << (bd - 8));
                }
        }
        p.deblock = c.deblock;
        p.sao_luma = p.sao_chroma = c.sao;
        int ctb = 64, wc = (W + ctb - 1) / ctb, hc = (H + ctb - 1) / ctb;
        for (int ry = 0; ry < hc; ++ry)
            for (int rx = 0; rx < wc; ++rx) cu(rx * ctb, ry * ctb, 6);
        // SAO per CTB: luma type, chroma type shared by Cb / Cr (class too).
        Lcg rsao(rv_gen_subseed(seed, 0, 0, SAO));
        g.sao.assign((size_t)wc * hc * 3, rv_sao{0, 0, {0, 0, 0, 0}});
        if (c.sao) {
            for (int i = 0; i < wc * hc; ++i) {
                int ty = rsao.uniform(0, 2), tc = rsao.uniform(0, 2), cls_c = rsao.uniform(0, 3);
                for (int comp = 0; comp < 3; ++comp) {
                    rv_sao& s = g.sao[(size_t)i * 3 + comp];
                    s.type_idx = (uint8_t)(comp == 0 ? ty : tc);
                    if (s.type_idx == 1) s.band_or_class = (uint8_t)rsao.uniform(0, 31);
                    else if (s.type_idx == 2) s.band_or_class = (uint8_t)(comp == 0 ? rsao.uniform(0, 3) : cls_c);
                    if (s.type_idx)
                        for (int k = 0; k < 4; ++k) s.offsets[k] = (int16_t)rsao.uniform(-7, 7);
                }
            }
        }
        p.n_cus = (int)g.cus.size();
        p.n_pus = (int)g.pus.size();
        p.n_tus = (int)g.tus.size();
        p.cus = g.cus.data();
        p.pus = g.pus.data();
        p.tus = g.tus.data();
        p.nz = g.nz.data();
        p.n_nz = (uint32_t)g.nz.size();
        p.intra_pred = g.intra_pred.data();
        p.n_intra_pred = (uint32_t)g.intra_pred.size();
        p.sao = g.sao.data();
        p.strong_intra_smoothing = c.strong_intra_smoothing;
        p.constrained_intra_pred = c.constrained_intra_pred;
        if (c.scaling) {
            Lcg rsl(rv_gen_subseed(seed, 0, 0, SCALING));
            fill_factors(rsl, &g.hevc_scaling.list4[0][0], sizeof(g.hevc_scaling.list4));
            fill_factors(rsl, &g.hevc_scaling.list8[0][0][0], sizeof(g.hevc_scaling.list8));
            fill_factors(rsl, &g.hevc_scaling.dc[0][0], sizeof(g.hevc_scaling.dc));
            p.scaling = &g.hevc_scaling;
        }
        slices_and_tiles(wc, hc, seed);
    }

    // Uniformly spaced tiles and the CTB raster-to-tile-scan map (6.5.1),
    // slices as runs of tile-scan order, per-slice filter settings, and
    // exempt CUs: all from their own seeded stream.
    void slices_and_tiles(int wc, int hc, uint64_t seed) {
        rv_hevc_pic& p = g.hevc;
        Lcg r(rv_gen_subseed(seed, 0, 0, SLICES));
        if (c.exempt_percent > 0)
            for (rv_hevc_cu& cu : g.cus) {
                if (!r.percent(c.exempt_percent)) continue;
                const int k = r.uniform(0, 2);
                cu.exempt = (uint8_t)(k == 0 ? RV_CU_LF_EXEMPT : k == 1 ? (RV_CU_LF_EXEMPT | RV_CU_BYPASS) : RV_CU_DBK_EXEMPT);
            }
        const int cols = std::max(1, c.tile_cols), rows = std::max(1, c.tile_rows);
        const int ns = std::max(1, c.n_slices);
        if (cols == 1 && rows == 1 && ns == 1) return;
        std::vector<int> cb(cols + 1), rb(rows + 1);
        for (int i = 0; i <= cols; ++i) cb[i] = i * wc / cols;
        for (int i = 0; i <= rows; ++i) rb[i] = i * hc / rows;
        const int n = wc * hc;
        g.ctb_ts.assign(n, 0);
        g.ctb_tile.assign(n, 0);
        std::vector<uint32_t> ts_to_rs(n);
        for (int rs = 0; rs < n; ++rs) {
            const int tx = rs % wc, ty = rs / wc;
            int i = 0, j = 0;
            while (tx >= cb[i + 1]) ++i;
            while (ty >= rb[j + 1]) ++j;
            int ts = 0;
            for (int jj = 0; jj < j; ++jj) ts += wc * (rb[jj + 1] - rb[jj]);
            for (int ii = 0; ii < i; ++ii) ts += (rb[j + 1] - rb[j]) * (cb[ii + 1] - cb[ii]);
            ts += (ty - rb[j]) * (cb[i + 1] - cb[i]) + tx - cb[i];
            g.ctb_ts[rs] = (uint32_t)ts;
            g.ctb_tile[rs] = (uint16_t)(j * cols + i);
            ts_to_rs[ts] = (uint32_t)rs;
        }
        g.slices.assign(ns, rv_hevc_slice{0, 0, 0, 1});
        for (int k = 0; k < ns; ++k) {
            rv_hevc_slice& sl = g.slices[k];
            sl.loop_filter_across_slices = c.lf_across_slices < 0 ? (int)(r.next() & 1) : c.lf_across_slices;
            if (c.slice_vary) {
                sl.deblocking_disabled = r.percent(20) ? 1 : 0;
                sl.beta_offset = 2 * r.uniform(-3, 3);
                sl.tc_offset = 2 * r.uniform(-3, 3);
            } else {
                sl.beta_offset = p.beta_offset;
                sl.tc_offset = p.tc_offset;
            }
        }
        g.ctb_slice.assign(n, 0);
        g.ctb_slice_addr.assign(n, 0);
        for (int k = 0; k < ns; ++k) {
            const int t0 = k * n / ns, t1 = (k + 1) * n / ns;
            for (int t = t0; t < t1; ++t) {
                g.ctb_slice[ts_to_rs[t]] = (uint16_t)k;
                g.ctb_slice_addr[ts_to_rs[t]] = ts_to_rs[t0];
            }
        }
        p.n_slices = ns;
        p.slices = g.slices.data();
        p.ctb_slice = g.ctb_slice.data();
        p.ctb_slice_addr = g.ctb_slice_addr.data();
        p.ctb_tile = g.ctb_tile.data();
        p.ctb_ts = g.ctb_ts.data();
        p.loop_filter_across_tiles = c.lf_across_tiles < 0 ? (int)(r.next() & 1) : c.lf_across_tiles;
    }
};

const int CHROMA_QP[52] = {0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16, 17,
                           18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 29, 30, 31, 32, 32, 33,
                           34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};

struct H264Builder {
    const rv_gen_h264_cfg& c;
    rv_gen& g;
    Lcg rs, rc, ri, rsl;
    H264Builder(const rv_gen_h264_cfg& cfg, rv_gen& gen, uint64_t seed)
        : c(cfg), g(gen), rs(rv_gen_subseed(seed, 0, 0, STRUCT)), rc(rv_gen_subseed(seed, 0, 0, COEF)),
          ri(rv_gen_subseed(seed, 0, 0, INTRA)), rsl(rv_gen_subseed(seed, 0, 0, SCALING)) {}

    int level() {
        int m = rc.laplace(c.lap_b);
        return (rc.next() & 1) ? -m : m;
    }

    // A DC group of n levels (levels mode): its offset in coefs; *any tells
    // whether one is nonzero.
    uint32_t dc_group(int n, bool* any) {
        const uint32_t off = (uint32_t)g.coefs32.size();
        *any = false;
        for (int i = 0; i < n; ++i) {
            int v = rc.permille(c.nz_permille) ? level() : 0;
            g.coefs32.push_back(v);
            *any |= v != 0;
        }
        return off;
    }

    // One block of levels whose position 0 comes from the DC group at
    // group_off (has_dc) or is a level itself; kept when any level of the
    // block or its group is nonzero.
    bool level_block(int comp, int x, int y, int n, bool has_dc, uint32_t group_off, bool group_any) {
        uint32_t off = (uint32_t)g.coefs32.size();
        g.coefs32.resize(off + n * n, 0);
        bool any = false;
        for (int i = has_dc ? 1 : 0; i < n * n; ++i) {
            if (!rc.permille(c.nz_permille)) continue;
            g.coefs32[off + i] = level();
            any = true;
        }
        if (!any && !(has_dc && group_any)) {
            g.coefs32.resize(off);
            return false;
        }
        rv_h264_blk b;
        std::memset(&b, 0, sizeof(b));
        b.x = (uint16_t)x;
        b.y = (uint16_t)y;
        b.comp = (uint8_t)comp;
        b.size = (uint8_t)n;
        b.has_dc = has_dc ? 1 : 0;
        b.flags = RV_BLK_LEVELS;
        b.dc = has_dc ? (int32_t)group_off : 0;
        b.coef_off = off;
        g.blks.push_back(b);
        return any;
    }

    // One coded block (or nothing); returns whether any coefficient is set.
    bool block(int comp, int x, int y, int n, bool chroma_dc) {
        if (c.levels) return level_block(comp, x, y, n, false, 0, false);
        uint32_t off = (uint32_t)g.coefs32.size();
        g.coefs32.resize(off + n * n, 0);
        bool any = false;
        for (int i = chroma_dc ? 1 : 0; i < n * n; ++i) {
            if (!rc.permille(c.nz_permille)) continue;
            int m = rc.laplace(c.lap_b);
            g.coefs32[off + i] = (rc.next() & 1) ? -m : m;
            any = true;
        }
        int dc = 0;
        if (chroma_dc && rc.permille(c.nz_permille)) {
            int m = rc.laplace(c.lap_b);
            dc = (rc.next() & 1) ? -m : m;
        }
        if (!any && dc == 0) {
            g.coefs32.resize(off);
            return false;
        }
        rv_h264_blk b;
        std::memset(&b, 0, sizeof(b));
        b.x = (uint16_t)x;
        b.y = (uint16_t)y;
        b.comp = (uint8_t)comp;
        b.size = (uint8_t)n;
        b.has_dc = chroma_dc ? 1 : 0;
        b.dc = dc;
        b.coef_off = off;
        g.blks.push_back(b);
        return true;
    }

    void part(int x, int y, int w, int h) {
        rv_h264_part p;
        std::memset(&p, 0, sizeof(p));
        p.x = (uint16_t)x;
        p.y = (uint16_t)y;
        p.w = (uint8_t)w;
        p.h = (uint8_t)h;
        int dir;
        if (!c.b_picture) dir = 1;
        else if (rs.percent(c.bi_percent)) dir = 3;
        else dir = rs.percent(50) ? 1 : 2;
        p.dir = (uint8_t)dir;
        for (int l = 0; l < 2; ++l) {
            if ((dir >> l) & 1) {
                p.ref_idx[l] = 0;
                p.mv[l].x = (int16_t)rs.uniform(-c.mv_range, c.mv_range);
                p.mv[l].y = (int16_t)rs.uniform(-c.mv_range, c.mv_range);
            } else {
                p.ref_idx[l] = -1;
            }
        }
        g.parts.push_back(p);
    }

    // Availability of a device-intra macroblock's neighbours (left, above,
    // above-right, above-left: present and, under constrained intra
    // prediction, intra) and of one block's neighbours, as the decoder
    // derives them for a one-slice frame (src/h264/recon.rs:301-307,
    // 936-1011); the modes drawn are among those these allow.
    struct Nb {
        bool a, b, c, d;
    };
    Nb mb_nb(int addr, int mbw) {
        const int mbx = addr % mbw, mby = addr / mbw;
        auto ok = [&](int n) { return !c.constrained_intra_pred || g.mbs[(size_t)n].intra != 0; };
        return Nb{mbx > 0 && ok(addr - 1), mby > 0 && ok(addr - mbw), mby > 0 && mbx < mbw - 1 && ok(addr - mbw + 1),
                  mbx > 0 && mby > 0 && ok(addr - mbw - 1)};
    }
    int pick(const bool (&allowed)[9], int n) {
        int k = 0, list[9];
        for (int i = 0; i < n; ++i)
            if (allowed[i]) list[k++] = i;
        return list[ri.uniform(0, k - 1)];
    }
    void intra_modes(rv_h264_mb& m, int addr, int mbw) {
        const Nb nb = mb_nb(addr, mbw);
        m.intra_kind = (uint8_t)ri.uniform(0, 2);
        if (m.intra_kind == RV_H264_I16X16) {
            const bool ok[9] = {nb.b, nb.a, true, nb.a && nb.b && nb.d};
            m.modes[0] = (uint8_t)pick(ok, 4);
        } else {
            const int step = m.intra_kind == RV_H264_I8X8 ? 2 : 1;
            for (int by = 0; by < 4; by += step)
                for (int bx = 0; bx < 4; bx += step) {
                    const bool left = bx > 0 || nb.a, top = by > 0 || nb.b;
                    const bool tl = (bx > 0 && by > 0) || (bx == 0 ? (by == 0 ? nb.d : nb.a) : nb.b);
                    const bool ok[9] = {top, left, true, top, top && left && tl, top && left && tl, top && left && tl,
                                        top, left};
                    m.modes[by * 4 + bx] = (uint8_t)pick(ok, 9);
                }
        }
        const bool okc[9] = {true, nb.a, nb.b, nb.a && nb.b && nb.d};
        m.chroma_mode = (uint8_t)pick(okc, 4);
    }

    void run() {
        rv_h264_pic& p = g.h264;
        std::memset(&p, 0, sizeof(p));
        int W = c.width, H = c.height, mbw = W / 16, mbh = H / 16;
        p.width = W;
        p.height = H;
        p.bit_depth = 8;
        p.poc = c.poc;
        p.num_ref[0] = 1;
        p.num_ref[1] = c.b_picture ? 1 : 0;
        for (int l = 0; l < 2; ++l) {
            p.ref_slot[l][0] = c.ref_slot[l];
            p.ref_poc[l][0] = c.ref_poc[l];
        }
        p.weighting = c.weighting;
        p.deblock = c.deblock;
        p.constrained_intra_pred = c.constrained_intra_pred;
        g.mbs.resize((size_t)mbw * mbh);
        for (int a = 0; a < mbw * mbh; ++a) {
            int mx = (a % mbw) * 16, my = (a / mbw) * 16;
            rv_h264_mb m;
            std::memset(&m, 0, sizeof(m));
            m.qp = (int8_t)rs.uniform(c.qp_min, c.qp_max);
            m.qpc[0] = m.qpc[1] = (int8_t)CHROMA_QP[m.qp];
            m.intra = rs.percent(c.intra_percent) ? (c.intra_device ? RV_INTRA_DEVICE : RV_INTRA_HOST) : 0;
            bool all8 = true;
            if (m.intra == RV_INTRA_DEVICE) {
                intra_modes(m, a, mbw);
            } else if (m.intra) {
                m.pred_off = (uint32_t)g.intra_pred.size();
                for (int i = 0; i < 256 + 128; ++i) g.intra_pred.push_back((uint16_t)ri.uniform(0, 255));
            } else {
                int t = rs.uniform(0, 3);
                if (t == 0) part(mx, my, 16, 16);
                else if (t == 1) { part(mx, my, 16, 8); part(mx, my + 8, 16, 8); }
                else if (t == 2) { part(mx, my, 8, 16); part(mx + 8, my, 8, 16); }
                else {
                    for (int s = 0; s < 4; ++s) {
                        int sx = mx + (s & 1) * 8, sy = my + (s >> 1) * 8;
                        int st = rs.uniform(0, 3);
                        if (st == 0) part(sx, sy, 8, 8);
                        else if (st == 1) { part(sx, sy, 8, 4); part(sx, sy + 4, 8, 4); all8 = false; }
                        else if (st == 2) { part(sx, sy, 4, 8); part(sx + 4, sy, 4, 8); all8 = false; }
                        else {
                            for (int k = 0; k < 4; ++k) part(sx + (k & 1) * 4, sy + (k >> 1) * 4, 4, 4);
                            all8 = false;
                        }
                    }
                }
            }
            m.transform_8x8 = (all8 && rs.percent(50)) ? 1 : 0;
            if (m.intra == RV_INTRA_DEVICE) m.transform_8x8 = m.intra_kind == RV_H264_I8X8 ? 1 : 0;
            const bool i16 = m.intra == RV_INTRA_DEVICE ? m.intra_kind == RV_H264_I16X16
                                                        : c.levels && m.intra && !m.transform_8x8 && rs.percent(50);
            uint16_t nz = 0;
            if (m.transform_8x8) {
                for (int b8 = 0; b8 < 4; ++b8) {
                    int bx = (b8 & 1) * 2, by = (b8 >> 1) * 2;
                    if (block(0, mx + bx * 4, my + by * 4, 8, false))
                        nz |= (uint16_t)((1u << (by * 4 + bx)) | (1u << (by * 4 + bx + 1)) |
                                         (1u << ((by + 1) * 4 + bx)) | (1u << ((by + 1) * 4 + bx + 1)));
                }
            } else if (i16 && c.levels) {
                // Intra_16x16: a luma DC group and sixteen blocks taking it.
                bool gany;
                const uint32_t go = dc_group(16, &gany);
                for (int r = 0; r < 16; ++r)
                    if (level_block(0, mx + (r % 4) * 4, my + (r / 4) * 4, 4, true, go, gany)) nz |= (uint16_t)(1u << r);
            } else if (i16) {
                // Intra_16x16 with scaled values: each block's DC given.
                for (int r = 0; r < 16; ++r)
                    if (block(0, mx + (r % 4) * 4, my + (r / 4) * 4, 4, true)) nz |= (uint16_t)(1u << r);
            } else {
                for (int r = 0; r < 16; ++r)
                    if (block(0, mx + (r % 4) * 4, my + (r / 4) * 4, 4, false)) nz |= (uint16_t)(1u << r);
            }
            m.nz_mask = nz;
            for (int comp = 1; comp < 3; ++comp) {
                if (c.levels) {
                    bool gany;
                    const uint32_t go = dc_group(4, &gany);
                    for (int b = 0; b < 4; ++b)
                        level_block(comp, mx / 2 + (b & 1) * 4, my / 2 + (b >> 1) * 4, 4, true, go, gany);
                } else {
                    for (int b = 0; b < 4; ++b) block(comp, mx / 2 + (b & 1) * 4, my / 2 + (b >> 1) * 4, 4, true);
                }
            }
            g.mbs[a] = m;
        }
        p.n_mbs = mbw * mbh;
        p.n_parts = (int)g.parts.size();
        p.n_blks = (int)g.blks.size();
        p.mbs = g.mbs.data();
        p.parts = g.parts.data();
        p.blks = g.blks.data();
        p.coefs = g.coefs32.data();
        p.n_coefs = (uint32_t)g.coefs32.size();
        p.intra_pred = g.intra_pred.data();
        p.n_intra_pred = (uint32_t)g.intra_pred.size();
        if (c.scaling) {
            fill_factors(rsl, &g.h264_scaling.list4[0][0], sizeof(g.h264_scaling.list4));
            fill_factors(rsl, &g.h264_scaling.list8[0][0], sizeof(g.h264_scaling.list8));
            p.scaling = &g.h264_scaling;
        }
    }
};

}  // namespace

extern "C" {

uint64_t rv_gen_subseed(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    return splitmix(splitmix(splitmix(splitmix(seed) ^ a) ^ (b * 0x100000001B3ull)) ^ (c * 0xC2B2AE3D27D4EB4Full));
}

int rv_gen_hevc(const rv_gen_hevc_cfg* cfg, uint64_t seed, rv_gen** out) {
    if (!cfg || !out) return RV_E_INVALID;
    if (cfg->width <= 0 || cfg->height <= 0 || (cfg->width & 7) || (cfg->height & 7) || cfg->width > 8192 ||
        cfg->height > 8192 || cfg->bit_depth < 8 || cfg->bit_depth > 12 || cfg->qp_min > cfg->qp_max)
        return RV_E_RANGE;
    rv_gen* g = new (std::nothrow) rv_gen();
    if (!g) return RV_E_NOMEM;
    g->codec = RV_CODEC_HEVC;
    HevcBuilder b(*cfg, *g, seed);
    b.run(seed);
    *out = g;
    return RV_OK;
}

int rv_gen_h264(const rv_gen_h264_cfg* cfg, uint64_t seed, rv_gen** out) {
    if (!cfg || !out) return RV_E_INVALID;
    if (cfg->width <= 0 || cfg->height <= 0 || (cfg->width & 15) || (cfg->height & 15) || cfg->width > 8192 ||
        cfg->height > 8192 || cfg->qp_min > cfg->qp_max || cfg->qp_min < 0 || cfg->qp_max > 51)
        return RV_E_RANGE;
    rv_gen* g = new (std::nothrow) rv_gen();
    if (!g) return RV_E_NOMEM;
    g->codec = RV_CODEC_H264;
    H264Builder b(*cfg, *g, seed);
    b.run();
    *out = g;
    return RV_OK;
}

const rv_hevc_pic* rv_gen_hevc_pic(const rv_gen* g) { return g && g->codec == RV_CODEC_HEVC ? &g->hevc : nullptr; }
const rv_h264_pic* rv_gen_h264_pic(const rv_gen* g) { return g && g->codec == RV_CODEC_H264 ? &g->h264 : nullptr; }
void rv_gen_free(rv_gen* g) { delete g; }

int rv_gen_planes(int width, int height, int bit_depth, uint64_t seed, void* y, void* cb, void* cr) {
    if (width <= 0 || height <= 0 || (width & 1) || (height & 1) || bit_depth < 8 || bit_depth > 16 || !y || !cb ||
        !cr)
        return RV_E_INVALID;
    void* planes[3] = {y, cb, cr};
    for (int c = 0; c < 3; ++c) {
        Lcg r(rv_gen_subseed(seed, 0, (uint64_t)c, 7));
        size_t n = c ? (size_t)(width / 2) * (height / 2) : (size_t)width * height;
        int sh = 31 - bit_depth;
        if (bit_depth == 8) {
            uint8_t* p = (uint8_t*)planes[c];
            for (size_t i = 0; i < n; ++i) p[i] = (uint8_t)(r.next() >> sh);
        } else {
            uint16_t* p = (uint16_t*)planes[c];
            for (size_t i = 0; i < n; ++i) p[i] = (uint16_t)(r.next() >> sh);
        }
    }
    return RV_OK;
}

}  // extern "C"
