"""Pins of the oracle's HEVC picture drivers (CPU, no GPU).

The kernel-level tests pin the oracle's dsp entries; these pin what the
whole-picture drivers add on top (predict_block, src/hevc/inter.rs:555-742;
the residual dispatch, src/hevc/ctu.rs:2031-2088; boundary_strengths and
deblock_rows, src/hevc/deblock.rs:15-515; the SAO driver,
src/hevc/sao.rs:110-431) against statements written here from H.265 itself,
never from the oracle's code:

- prediction: whole-sample vectors give the clamped reference crop; every
  PU equals the standard's separable interpolation (8.5.3.3.3, taps of
  Tables 8-12 / 8-13 typed here) over clamped reference coordinates, with
  the default (8.5.3.3.4.2) or explicit (8.5.3.3.4.3) combination, written
  here in numpy; bi-prediction of one reference and vector equals
  uni-prediction; explicit weights of 2^denom with offset 0 equal the
  default weighting;
- residual: the reconstruction minus the prediction is each coded TU's
  inverse transform (the kernel-level-pinned entry) at the TU's place;
- boundary strengths of hand-built pictures against 8.7.2.4 by hand;
- deblocking: thresholds forced to zero leave the picture alone; samples
  change only next to edges with bS > 0; the whole picture equals the
  pinned edge filters applied to every edge of the bS map in the
  standard's order (all vertical edges, then all horizontal), with beta and
  tC from Table 8-12 and the chroma QP from Table 8-10, typed here;
- SAO: all-off is the deblocked picture; every CTB equals 8.7.3 written
  here in numpy over the deblocked picture (band table, edge classes,
  picture-border samples untouched);
- the 32-point transform matrix against the standard's listing.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import _abi
from rivet_hip import configs as cf

# ---------------------------------------------------------------- the standard

# Table 8-12 (luma, 1/4 sample) and Table 8-13 (chroma, 1/8 sample).
QPEL = [[0, 0, 0, 64, 0, 0, 0, 0], [-1, 4, -10, 58, 17, -5, 1, 0], [-1, 4, -11, 40, 40, -11, 4, -1],
        [0, 1, -5, 17, 58, -10, 4, -1]]
EPEL = [[0, 64, 0, 0], [-2, 58, 10, -2], [-4, 54, 16, -2], [-6, 46, 28, -4], [-4, 36, 36, -4], [-4, 28, 46, -6],
        [-2, 16, 54, -4], [-2, 10, 58, -2]]
# Table 8-12 of the deblocking clause: beta' by Q = 0..51, tC' by Q = 0..53.
BETA_P = [0] * 16 + list(range(6, 19)) + list(range(20, 65, 2))
TC_P = [0] * 18 + [1] * 9 + [2] * 4 + [3] * 4 + [4] * 3 + [5, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 22, 24]
# Table 8-10 (ChromaArrayType 1): QpC from qPi for 30 <= qPi <= 43.
QPC_TABLE = {30: 29, 31: 30, 32: 31, 33: 32, 34: 33, 35: 33, 36: 34, 37: 34, 38: 35, 39: 35, 40: 36, 41: 36,
             42: 37, 43: 37}


def qpc_420(qpi):
    return qpi if qpi < 30 else (qpi - 6 if qpi > 43 else QPC_TABLE[qpi])


def test_typed_tables_have_the_standards_shape():
    assert len(BETA_P) == 52 and len(TC_P) == 54
    assert BETA_P[16] == 6 and BETA_P[28] == 18 and BETA_P[29] == 20 and BETA_P[51] == 64
    assert TC_P[17] == 0 and TC_P[18] == 1 and TC_P[27] == 2 and TC_P[53] == 24
    assert all(sum(t) == 64 for t in QPEL + EPEL)


def test_oracle_filter_taps_are_the_standards():
    """Impulse responses of the oracle's first-stage entries."""
    for frac in range(1, 4):
        got = []
        for k in range(8):
            src = np.zeros(8, np.uint8)
            src[k] = 1
            out = np.zeros(1, np.int16)
            oracle.lib.orc_hevc_qpel_h(1, out.ctypes.data_as(C.POINTER(C.c_int16)), src.ctypes.data, 8, 1, 1,
                                       frac, 0)
            got.append(int(out[0]))
        assert got == QPEL[frac], frac
    for frac in range(1, 8):
        got = []
        for k in range(4):
            src = np.zeros(4, np.uint8)
            src[k] = 1
            out = np.zeros(1, np.int16)
            oracle.lib.orc_hevc_epel_h(1, out.ctypes.data_as(C.POINTER(C.c_int16)), src.ctypes.data, 4, 1, 1,
                                       frac, 0)
            got.append(int(out[0]))
        assert got == EPEL[frac], frac


# ---------------------------------------------------------------- helpers

def gen(cfg, seed):
    return rh.GenPicture(rh.RV_CODEC_HEVC, cfg, seed)


def refs_of(g, seed):
    c = g.pic.contents
    return [rh.gen_planes(c.width, c.height, c.bit_depth, cf.ref_seed(seed, k)) for k in range(2)]


def run(g, refs, stop=0):
    return oracle.reconstruct(rh.RV_CODEC_HEVC, g.pic, list(refs) + [None], stop)


def filters_off(g):
    c = g.pic.contents
    c.deblock = 0
    c.sao_luma = c.sao_chroma = 0


def no_residual(g):
    c = g.pic.contents
    for i in range(c.n_tus):
        c.tus[i].flags &= ~_abi.RV_TU_CODED


def clamp_window(plane, x0, y0, w, h):
    """plane[y][x] for y0 <= y < y0 + h, x0 <= x < x0 + w, coordinates
    clamped into the picture (the reference picture's padding)."""
    H, W = plane.shape
    ys = np.clip(np.arange(y0, y0 + h), 0, H - 1)
    xs = np.clip(np.arange(x0, x0 + w), 0, W - 1)
    return plane[ys[:, None], xs[None, :]].astype(np.int64)


def interp(plane, xi, yi, w, h, fx, fy, taps, bd):
    """8.5.3.3.3.1 / 8.5.3.3.3.2: the prediction sample array (14-bit) of a
    w x h block whose integer position in the reference is (xi, yi)."""
    n = len(taps[0])
    r = n // 2 - 1
    shift1, shift3 = min(4, bd - 8), 14 - bd
    win = clamp_window(plane, xi - r, yi - r, w + n - 1, h + n - 1)
    if fx == 0 and fy == 0:
        return win[r:r + h, r:r + w] << shift3
    if fy == 0:
        return sum(taps[fx][k] * win[r:r + h, k:k + w] for k in range(n)) >> shift1
    if fx == 0:
        return sum(taps[fy][k] * win[k:k + h, r:r + w] for k in range(n)) >> shift1
    rows = sum(taps[fx][k] * win[:, k:k + w] for k in range(n)) >> shift1
    return sum(taps[fy][k] * rows[k:k + h] for k in range(n)) >> 6


def combine(preds, bd, explicit=None):
    """8.5.3.3.4.2 (default) or 8.5.3.3.4.3 (explicit: (log2WD, [w], [o])
    per list used)."""
    mx = (1 << bd) - 1
    if explicit is None:
        if len(preds) == 1:
            s = 14 - bd
            v = (preds[0] + (1 << (s - 1))) >> s
        else:
            s = 15 - bd
            v = (preds[0] + preds[1] + (1 << (s - 1))) >> s
    else:
        log2wd, ws, os_ = explicit
        if len(preds) == 1:
            if log2wd >= 1:
                v = ((preds[0] * ws[0] + (1 << (log2wd - 1))) >> log2wd) + os_[0]
            else:
                v = preds[0] * ws[0] + os_[0]
        else:
            v = (preds[0] * ws[0] + preds[1] * ws[1] + ((os_[0] + os_[1] + 1) << log2wd)) >> (log2wd + 1)
    return np.clip(v, 0, mx)


def predict_picture(g, refs):
    """The prediction of every PU (and the host-supplied intra blocks) as
    the standard states it, in numpy."""
    c = g.pic.contents
    W, H, bd = c.width, c.height, c.bit_depth
    out = [np.zeros((H, W), np.int64), np.zeros((H // 2, W // 2), np.int64), np.zeros((H // 2, W // 2), np.int64)]
    for i in range(c.n_cus):
        cu = c.cus[i]
        if not cu.intra:
            continue
        s = 1 << cu.log2_size
        off = cu.pred_off
        for k in range(3):
            n = s if k == 0 else s // 2
            x, y = (cu.x, cu.y) if k == 0 else (cu.x // 2, cu.y // 2)
            out[k][y:y + n, x:x + n] = np.array(c.intra_pred[off:off + n * n], np.int64).reshape(n, n)
            off += n * n
    for i in range(c.n_pus):
        u = c.pus[i]
        lists = [l for l in range(2) if (u.dir >> l) & 1]
        for k in range(3):
            preds = []
            for l in lists:
                ref = refs[c.ref_slot[l][u.ref_idx[l]]][k]
                mvx, mvy = u.mv[l].x, u.mv[l].y
                if k == 0:
                    preds.append(interp(ref, u.x + (mvx >> 2), u.y + (mvy >> 2), u.w, u.h, mvx & 3, mvy & 3, QPEL,
                                        bd))
                else:  # 4:2:0: the vector in 1/8 chroma samples
                    preds.append(interp(ref, u.x // 2 + (mvx >> 3), u.y // 2 + (mvy >> 3), u.w // 2, u.h // 2,
                                        mvx & 7, mvy & 7, EPEL, bd))
            explicit = None
            if c.weighted:
                explicit = (c.log2_wd[k], [c.wp_w[l][u.ref_idx[l]][k] for l in lists],
                            [c.wp_o[l][u.ref_idx[l]][k] for l in lists])
            x, y, w, h = (u.x, u.y, u.w, u.h) if k == 0 else (u.x // 2, u.y // 2, u.w // 2, u.h // 2)
            out[k][y:y + h, x:x + w] = combine(preds, bd, explicit)
    return out


def assert_planes_equal(got, want, what):
    for k in range(3):
        assert np.array_equal(np.asarray(got[k], np.int64), np.asarray(want[k], np.int64)), f"{what}: plane {k}"


# Pictures of the configurations' distributions at 416x240 (partial CTBs on
# the right and bottom), seeds fixed.
PRED_CASES = {
    "8-bit P": (cf.HEVC_C1, 101),
    "8-bit B, vectors to +-300 px": (dict(cf.HEVC_C1, b_picture=1, bi_percent=60, poc=2, ref_poc=[0, 4],
                                          ref_slot=[0, 1], mv_range=1200), 102),
    "10-bit B explicit WP with intra CUs": (dict(cf.HEVC_C2, width=416, height=240), 103),
    "8-bit B explicit WP with intra CUs": (dict(cf.HEVC_C2, width=416, height=240, bit_depth=8), 104),
}


# ---------------------------------------------------------------- prediction

@pytest.mark.parametrize("bd,far", [(8, False), (10, False), (8, True)])
def test_whole_sample_vectors_give_the_clamped_reference_crop(bd, far):
    cfg = dict(cf.HEVC_C1, bit_depth=bd, mv_range=1200 if far else 256)
    g = gen(cfg, 201 + bd + far)
    refs = refs_of(g, 201)
    c = g.pic.contents
    no_residual(g)
    filters_off(g)
    for i in range(c.n_pus):
        u = c.pus[i]
        # Whole luma and whole chroma samples.
        u.mv[0].x = (u.mv[0].x >> 3) << 3
        u.mv[0].y = (u.mv[0].y >> 3) << 3
    got = run(g, refs)
    want = [np.zeros_like(p) for p in refs[0]]
    for i in range(c.n_pus):
        u = c.pus[i]
        dx, dy = u.mv[0].x >> 2, u.mv[0].y >> 2
        want[0][u.y:u.y + u.h, u.x:u.x + u.w] = clamp_window(refs[0][0], u.x + dx, u.y + dy, u.w, u.h)
        for k in (1, 2):
            x, y = u.x // 2, u.y // 2
            want[k][y:y + u.h // 2, x:x + u.w // 2] = clamp_window(refs[0][k], x + dx // 2, y + dy // 2, u.w // 2,
                                                                   u.h // 2)
    assert_planes_equal(got, want, "crop")
    if far:  # some windows lie wholly outside the padded reference
        assert max(abs(c.pus[i].mv[0].x) for i in range(c.n_pus)) > 4 * (80 + 64)


@pytest.mark.parametrize("name", sorted(PRED_CASES))
def test_prediction_is_the_standards_interpolation_and_weighting(name):
    cfg, seed = PRED_CASES[name]
    g = gen(cfg, seed)
    refs = refs_of(g, seed)
    got = run(g, refs, stop=1)
    want = predict_picture(g, refs)
    assert_planes_equal(got, want, name)


def test_bi_prediction_from_one_reference_and_vector_equals_uni():
    cfg = dict(cf.HEVC_C1, bit_depth=10)
    g = gen(cfg, 301)
    refs = refs_of(g, 301)
    filters_off(g)
    c = g.pic.contents
    uni = run(g, refs, stop=2)
    c.num_ref[1] = 1
    c.ref_slot[1][0] = c.ref_slot[0][0]
    c.ref_poc[1][0] = c.ref_poc[0][0]
    for i in range(c.n_pus):
        u = c.pus[i]
        u.dir = 3
        u.ref_idx[1] = 0
        u.mv[1].x, u.mv[1].y = u.mv[0].x, u.mv[0].y
    bi = run(g, refs, stop=2)
    assert_planes_equal(bi, uni, "bi(a, a) vs uni(a)")


@pytest.mark.parametrize("bd", [8, 10])
def test_explicit_weights_of_one_equal_the_default_weighting(bd):
    cfg = dict(cf.HEVC_C2, width=416, height=240, bit_depth=bd, explicit_wp=0)
    g = gen(cfg, 401 + bd)
    refs = refs_of(g, 401)
    c = g.pic.contents
    default = run(g, refs)  # deblocking and SAO on
    denom = 5
    c.weighted = 1
    for k in range(3):
        c.log2_wd[k] = denom + 14 - bd
        for l in range(2):
            for r in range(_abi.RV_MAX_REFS):
                c.wp_w[l][r][k] = 1 << denom
                c.wp_o[l][r][k] = 0
    unity = run(g, refs)
    assert_planes_equal(unity, default, "explicit w = 2^denom, o = 0")


# ---------------------------------------------------------------- residual

def test_residual_is_each_coded_tus_inverse_transform_in_place():
    cfg = dict(cf.HEVC_C2, width=416, height=240)  # DST 4x4 luma TUs in intra CUs, 10-bit
    g = gen(cfg, 501)
    refs = refs_of(g, 501)
    c = g.pic.contents
    bd = c.bit_depth
    pred = [p.astype(np.int64) for p in run(g, refs, stop=1)]
    got = run(g, refs, stop=2)
    n_dst = 0
    for i in range(c.n_tus):
        t = c.tus[i]
        if not t.flags & _abi.RV_TU_CODED:
            continue
        n = 1 << t.log2
        coefs = np.zeros(n * n, np.int16)
        for j in range(t.coef_off, t.coef_off + t.n_nz):
            coefs[c.nz[j].pos] = c.nz[j].val
        p = coefs.ctypes.data_as(C.POINTER(C.c_int16))
        if t.flags & _abi.RV_TU_DST:
            assert t.comp == 0 and n == 4
            oracle.lib.orc_hevc_idst4(p, 20 - bd)
            n_dst += 1
        else:
            oracle.lib.orc_hevc_idct(p, t.log2, 20 - bd, n - 1, n - 1)
        pl = pred[t.comp]
        pl[t.y:t.y + n, t.x:t.x + n] = np.clip(pl[t.y:t.y + n, t.x:t.x + n] + coefs.reshape(n, n), 0, (1 << bd) - 1)
    assert n_dst > 0
    assert_planes_equal(got, pred, "residual")


# ---------------------------------------------------------------- boundary strengths

class HandPicture:
    """A picture built by hand: CUs, PUs and luma TUs at given places. Two
    reference pictures per list: list 0 = (POC 0, POC 8), list 1 = (POC 8,
    POC 0); the current picture is POC 4."""

    def __init__(self, w, h):
        self.w, self.h = w, h
        self.cus, self.pus, self.tus = [], [], []
        self.long_term = [[0, 0], [0, 0]]

    def cu(self, x, y, log2, intra=0, qp=30):
        self.cus.append((x, y, log2, intra, qp))
        return self

    def pu(self, x, y, w, h, l0=None, l1=None):
        """l0 / l1 = (ref_idx, mvx, mvy) or None."""
        self.pus.append((x, y, w, h, l0, l1))
        return self

    def tu(self, x, y, log2, coded=False):
        self.tus.append((x, y, log2, coded))
        return self

    def build(self):
        c = _abi.HevcPic()
        c.width, c.height, c.bit_depth, c.log2_ctb, c.poc = self.w, self.h, 8, 6, 4
        for l, pocs in enumerate([(0, 8), (8, 0)]):
            c.num_ref[l] = 2
            for r, poc in enumerate(pocs):
                c.ref_slot[l][r] = 0 if poc == 0 else 1
                c.ref_poc[l][r] = poc
                c.ref_long_term[l][r] = self.long_term[l][r]
        c.deblock = 1
        cus = (_abi.HevcCu * max(1, len(self.cus)))()
        for i, (x, y, log2, intra, qp) in enumerate(self.cus):
            cus[i].x, cus[i].y, cus[i].log2_size, cus[i].intra, cus[i].qp_y = x, y, log2, intra, qp
        pus = (_abi.HevcPu * max(1, len(self.pus)))()
        for i, (x, y, w, h, l0, l1) in enumerate(self.pus):
            u = pus[i]
            u.x, u.y, u.w, u.h = x, y, w, h
            u.ref_idx[0] = u.ref_idx[1] = -1
            for l, m in enumerate((l0, l1)):
                if m is not None:
                    u.dir |= 1 << l
                    u.ref_idx[l] = m[0]
                    u.mv[l].x, u.mv[l].y = m[1], m[2]
        tus = (_abi.HevcTu * max(1, len(self.tus)))()
        for i, (x, y, log2, coded) in enumerate(self.tus):
            tus[i].x, tus[i].y, tus[i].comp, tus[i].log2 = x, y, 0, log2
            tus[i].flags = _abi.RV_TU_CODED if coded else 0
        c.n_cus, c.n_pus, c.n_tus = len(self.cus), len(self.pus), len(self.tus)
        c.cus, c.pus, c.tus = cus, pus, tus
        self._keep = (cus, pus, tus)
        self.c = c
        return C.pointer(c)

    def bs(self):
        return oracle.hevc_bs_map(self.build())


def side_by_side(a, b, coded=(False, False)):
    """Two 16x16 CUs, A at x = 0 and B at x = 16 (a 32x16 picture), each one
    PU and one luma TU. a / b = "intra" or (l0, l1)."""
    p = HandPicture(32, 16)
    for k, m in enumerate((a, b)):
        x = 16 * k
        if m == "intra":
            p.cu(x, 0, 4, intra=1)
        else:
            p.cu(x, 0, 4).pu(x, 0, 16, 16, *m)
        p.tu(x, 0, 4, coded[k])
    return p


def edge_at_x16(value):
    ver = np.zeros((4, 8), np.uint8)
    ver[:, 4] = value
    return ver


MV = (0, 12, -20)  # list-0 reference 0, vector (12, -20)
MV_L1 = (0, -8, 4)  # list-1 reference 0 (POC 8)


@pytest.mark.parametrize("a,b,coded,want", [
    ("intra", (MV, None), (False, False), 2),               # intra on the p side
    ((MV, None), "intra", (True, True), 2),                 # intra wins over coefficients
    ((MV, None), (MV, None), (False, True), 1),             # a transform edge with coefficients
    ((MV, None), (MV, None), (True, False), 1),
    ((MV, None), (MV, None), (False, False), 0),            # same motion, no coefficients
    ((MV, None), ((0, 16, -20), None), (False, False), 1),  # |dx| = 4 quarter samples
    ((MV, None), ((0, 15, -20), None), (False, False), 0),  # |dx| = 3
    ((MV, None), ((0, 12, -24), None), (False, False), 1),  # |dy| = 4
    ((MV, None), ((0, 12, -17), None), (False, False), 0),  # |dy| = 3
    ((MV, None), ((1, 12, -20), None), (False, False), 1),  # another reference picture
    ((MV, None), (MV, MV_L1), (False, False), 1),           # one vector against two
    ((MV, MV_L1), ((1, -8, 4), (1, 12, -20)), (False, False), 0),  # the same pair, lists swapped
    ((MV, MV_L1), ((1, -8, 4), (1, 16, -20)), (False, False), 1),  # swapped, one vector 4 apart
    ((MV, MV_L1), (MV, MV_L1), (False, False), 0),
])
def test_boundary_strength_across_a_coding_unit_edge(a, b, coded, want):
    ver, hor = side_by_side(a, b, coded).bs()
    assert np.array_equal(ver, edge_at_x16(want)), ver
    assert not hor.any()  # no horizontal edge inside; the top border is never filtered


def test_boundary_strength_with_both_vectors_on_one_reference():
    # Both lists point at POC 0: bS 0 when the vectors match straight or
    # crosswise, 1 only when both pairings differ by 4 or more.
    a = ((0, 0, 0), (1, 40, 0))
    for b, want in [(((0, 40, 0), (1, 0, 0)), 0), (((0, 0, 0), (1, 40, 0)), 0), (((0, 4, 0), (1, 40, 0)), 1),
                    (((0, 40, 0), (1, 4, 0)), 1)]:
        ver, _ = side_by_side(a, b).bs()
        assert np.array_equal(ver, edge_at_x16(want)), (b, ver)


def test_boundary_strength_long_term_against_short_term():
    p2 = HandPicture(32, 16)
    p2.long_term = [[0, 1], [0, 0]]
    p2.cu(0, 0, 4).pu(0, 0, 16, 16, (1, 0, 0)).tu(0, 0, 4)  # POC 8 long-term
    p2.cu(16, 0, 4).pu(16, 0, 16, 16, None, (0, 0, 0)).tu(16, 0, 4)  # POC 8 short-term
    assert np.array_equal(p2.bs()[0], edge_at_x16(1))
    # The same picture (POC 8) through both lists, both short-term: 0.
    p3 = HandPicture(32, 16)
    p3.cu(0, 0, 4).pu(0, 0, 16, 16, (1, 0, 0)).tu(0, 0, 4)
    p3.cu(16, 0, 4).pu(16, 0, 16, 16, None, (0, 0, 0)).tu(16, 0, 4)
    assert not p3.bs()[0].any()


def test_boundary_strength_inside_a_coding_unit():
    # Nx2N: the PU edge at x = 8 is on the 8x8 grid.
    p = HandPicture(16, 16).cu(0, 0, 4).tu(0, 0, 4)
    p.pu(0, 0, 8, 16, (0, 0, 0)).pu(8, 0, 8, 16, (0, 4, 0))
    ver, hor = p.bs()
    want = np.zeros((4, 4), np.uint8)
    want[:, 2] = 1
    assert np.array_equal(ver, want) and not hor.any()
    # nLx2N: the PU edge at x = 4 is off the grid.
    p = HandPicture(16, 16).cu(0, 0, 4).tu(0, 0, 4)
    p.pu(0, 0, 4, 16, (0, 0, 0)).pu(4, 0, 12, 16, (0, 40, 0))
    ver, hor = p.bs()
    assert not ver.any() and not hor.any()
    # 2NxN: the PU edge at y = 8.
    p = HandPicture(16, 16).cu(0, 0, 4).tu(0, 0, 4)
    p.pu(0, 0, 16, 8, (0, 0, 0)).pu(0, 8, 16, 8, (1, 0, 0))
    ver, hor = p.bs()
    want = np.zeros((4, 4), np.uint8)
    want[2, :] = 1
    assert np.array_equal(hor, want) and not ver.any()
    # Four 8x8 TUs, one coded, same motion: its two inner edges get 1.
    p = HandPicture(16, 16).cu(0, 0, 4).pu(0, 0, 16, 16, (0, 0, 0))
    p.tu(0, 0, 3).tu(8, 0, 3).tu(0, 8, 3, coded=True).tu(8, 8, 3)
    ver, hor = p.bs()
    wv = np.zeros((4, 4), np.uint8)
    wv[2:, 2] = 1  # x = 8, rows 8..15 (the coded TU's right edge)
    wh = np.zeros((4, 4), np.uint8)
    wh[2, :2] = 1  # y = 8, columns 0..7 (its top edge)
    assert np.array_equal(ver, wv) and np.array_equal(hor, wh)


def test_picture_borders_are_never_filtered():
    p = HandPicture(16, 32).cu(0, 0, 4, intra=1).tu(0, 0, 4).cu(0, 16, 4).pu(0, 16, 16, 16, (0, 0, 0)).tu(0, 16, 4)
    ver, hor = p.bs()
    assert not ver.any()  # x = 0 only
    want = np.zeros((8, 4), np.uint8)
    want[4, :] = 2  # y = 16: intra above
    assert np.array_equal(hor, want)


# ---------------------------------------------------------------- deblocking

DBK_CASES = {
    "8-bit P": (cf.HEVC_C1, 601),
    "10-bit B explicit WP with intra CUs": (dict(cf.HEVC_C2, width=416, height=240), 602),
    "8-bit B with intra CUs and PPS chroma offsets": (dict(cf.HEVC_C2, width=416, height=240, bit_depth=8), 603),
}


def _dbk_case(name, chroma_offsets=False):
    cfg, seed = DBK_CASES[name]
    g = gen(cfg, seed)
    if "PPS chroma" in name:
        c = g.pic.contents
        c.cb_qp_offset, c.cr_qp_offset = 5, -3
        c.beta_offset, c.tc_offset = 4, -2
    return g, refs_of(g, seed)


def test_zero_thresholds_leave_the_picture_unfiltered():
    g, refs = _dbk_case("10-bit B explicit WP with intra CUs")
    c = g.pic.contents
    before = run(g, refs, stop=2)
    assert not all(np.array_equal(a, b) for a, b in zip(run(g, refs, stop=3), before))  # not vacuous
    # Q <= 15 for beta, Q <= 17 for tC at every edge: QP 27 at most with
    # both offsets -12 (bS 2 adds 2 to tC's Q).
    for i in range(c.n_cus):
        c.cus[i].qp_y = min(c.cus[i].qp_y, 27)
    c.beta_offset = c.tc_offset = -12
    assert_planes_equal(run(g, refs, stop=3), before, "beta = tC = 0")


@pytest.mark.parametrize("name", sorted(DBK_CASES))
def test_deblocking_changes_only_samples_next_to_filtered_edges(name):
    g, refs = _dbk_case(name)
    c = g.pic.contents
    ver, hor = oracle.hevc_bs_map(g.pic)
    H, W = c.height, c.width
    ml = np.zeros((H, W), bool)
    mc = np.zeros((H // 2, W // 2), bool)
    for by, bx in zip(*np.nonzero(ver)):
        x, y = 4 * bx, 4 * by
        ml[y:y + 4, x - 3:x + 3] = True
        if ver[by, bx] == 2 and x % 16 == 0:
            mc[y // 2:y // 2 + 2, x // 2 - 1:x // 2 + 1] = True
    for by, bx in zip(*np.nonzero(hor)):
        x, y = 4 * bx, 4 * by
        ml[y - 3:y + 3, x:x + 4] = True
        if hor[by, bx] == 2 and y % 16 == 0:
            mc[y // 2 - 1:y // 2 + 1, x // 2:x // 2 + 2] = True
    before, after = run(g, refs, stop=2), run(g, refs, stop=3)
    changed = [a != b for a, b in zip(after, before)]
    assert changed[0].any() and not (changed[0] & ~ml).any()
    for k in (1, 2):
        assert not (changed[k] & ~mc).any()
    if (ver == 2).any():
        assert changed[1].any() or changed[2].any()


def _filter_picture(planes, c, ver, hor):
    """8.7.2: every vertical edge of the picture, then every horizontal one,
    each 4-line segment through the pinned edge filters, with beta / tC
    from Table 8-12 (and QpC from Table 8-10) as 8.7.2.5.3 / 8.7.2.5.5
    derive them."""
    bd = c.bit_depth
    mx = (1 << bd) - 1
    ss = 1 if bd == 8 else 2
    dt = np.uint8 if bd == 8 else np.uint16
    w4 = c.width // 4
    qp = np.zeros(ver.shape, np.int64)
    for i in range(c.n_cus):
        cu = c.cus[i]
        s = (1 << cu.log2_size) // 4
        qp[cu.y // 4:cu.y // 4 + s, cu.x // 4:cu.x // 4 + s] = cu.qp_y
    pad = 8
    P = [np.pad(np.asarray(p, dt), pad, mode="edge").copy() for p in planes]
    i32 = lambda *v: (C.c_int32 * len(v))(*v)  # noqa: E731
    u8 = lambda *v: (C.c_uint8 * len(v))(*v)  # noqa: E731
    for vertical, bsm in ((1, ver), (0, hor)):
        # Luma.
        Y = P[0]
        for by, bx in zip(*np.nonzero(bsm)):
            b = int(bsm[by, bx])
            qq = qp[by, bx]
            qpp = qp[by, bx - 1] if vertical else qp[by - 1, bx]
            qpl = (int(qq) + int(qpp) + 1) >> 1
            beta = BETA_P[min(max(qpl + c.beta_offset, 0), 51)] << (bd - 8)
            tc = TC_P[min(max(qpl + 2 * (b - 1) + c.tc_offset, 0), 53)] << (bd - 8)
            x, y = 4 * bx + pad, 4 * by + pad
            off = y * Y.shape[1] + x
            oracle.lib.orc_hevc_deblock_luma(ss, vertical, Y.ctypes.data, off, Y.shape[1], i32(beta, 0),
                                             i32(tc, 0), u8(0, 0), u8(0, 0), mx)
        # Chroma: bS 2 edges on the 16-luma (8-chroma) grid.
        for by, bx in zip(*np.nonzero(bsm == 2)):
            x, y = 4 * bx, 4 * by
            if (x if vertical else y) % 16:
                continue
            qq = qp[by, bx]
            qpp = qp[by, bx - 1] if vertical else qp[by - 1, bx]
            for k, off_c in ((1, c.cb_qp_offset), (2, c.cr_qp_offset)):
                qpi = ((int(qq) + int(qpp) + 1) >> 1) + off_c
                qpc = qpc_420(qpi)
                tc = TC_P[min(max(qpc + 2 + c.tc_offset, 0), 53)] << (bd - 8)
                Cp = P[k]
                off = (y // 2 + pad) * Cp.shape[1] + x // 2 + pad
                oracle.lib.orc_hevc_deblock_chroma(ss, vertical, Cp.ctypes.data, off, Cp.shape[1],
                                                   i32(tc, 0, 0, 0), u8(0, 0, 0, 0), u8(0, 0, 0, 0), mx)
    return [p[pad:-pad, pad:-pad] for p in P]


@pytest.mark.parametrize("name", sorted(DBK_CASES))
def test_deblocking_is_the_edge_filters_in_the_standards_order(name):
    g, refs = _dbk_case(name)
    c = g.pic.contents
    ver, hor = oracle.hevc_bs_map(g.pic)
    want = _filter_picture(run(g, refs, stop=2), c, ver, hor)
    assert_planes_equal(run(g, refs, stop=3), want, name)


# ---------------------------------------------------------------- SAO

HPOS = [(-1, 1), (0, 0), (-1, 1), (1, -1)]
VPOS = [(0, 0), (-1, 1), (-1, 1), (-1, 1)]


def sao_picture(deb, c):
    """8.7.3 over the whole picture from the deblocked planes: per CTB and
    component, band offsets (8.7.3.2 bandTable / bandShift) or edge offsets
    (edgeIdx from the two neighbours of the class, unchanged where a
    neighbour lies outside the picture)."""
    bd = c.bit_depth
    mx = (1 << bd) - 1
    ctb = 1 << c.log2_ctb
    wc, hc = -(-c.width // ctb), -(-c.height // ctb)
    out = [np.array(p, np.int64) for p in deb]
    for k in range(3):
        if not (c.sao_luma if k == 0 else c.sao_chroma):
            continue
        rec = np.asarray(deb[k], np.int64)
        Hk, Wk = rec.shape
        cs = ctb if k == 0 else ctb // 2
        for ry in range(hc):
            for rx in range(wc):
                p = c.sao[3 * (ry * wc + rx) + k]
                ys, xs = slice(ry * cs, min((ry + 1) * cs, Hk)), slice(rx * cs, min((rx + 1) * cs, Wk))
                v = rec[ys, xs]
                val = [0] + list(p.offsets)  # SaoOffsetVal[0..4]
                if p.type_idx == 1:
                    table = np.zeros(32, np.int64)
                    for j in range(4):
                        table[(j + p.band_or_class) & 31] = j + 1
                    idx = table[v >> (bd - 5)]
                    out[k][ys, xs] = np.clip(v + np.array(val)[idx], 0, mx)
                elif p.type_idx == 2:
                    (ha, hb), (va, vb) = HPOS[p.band_or_class], VPOS[p.band_or_class]
                    yy, xx = np.mgrid[ys, xs]
                    ya, xa, yb, xb = yy + va, xx + ha, yy + vb, xx + hb
                    inside = (ya >= 0) & (ya < Hk) & (xa >= 0) & (xa < Wk) & (yb >= 0) & (yb < Hk) & (xb >= 0) & (
                        xb < Wk)
                    a = rec[np.clip(ya, 0, Hk - 1), np.clip(xa, 0, Wk - 1)]
                    b = rec[np.clip(yb, 0, Hk - 1), np.clip(xb, 0, Wk - 1)]
                    e = 2 + np.sign(v - a) + np.sign(v - b)
                    e = np.where(e == 2, 0, np.where(e < 2, e + 1, e))
                    off = np.array(val)[e]
                    out[k][ys, xs] = np.where(inside, np.clip(v + off, 0, mx), v)
    return out


SAO_CASES = {
    "8-bit P": (cf.HEVC_C1, 701, None),
    "10-bit B with intra CUs": (dict(cf.HEVC_C2, width=416, height=240), 702, None),
    "8-bit, luma SAO only": (cf.HEVC_C1, 703, (1, 0)),
    "8-bit, chroma SAO only": (cf.HEVC_C1, 704, (0, 1)),
}


@pytest.mark.parametrize("name", sorted(SAO_CASES))
def test_sao_is_the_standards_offsets_over_the_deblocked_picture(name):
    cfg, seed, flags = SAO_CASES[name]
    g = gen(cfg, seed)
    refs = refs_of(g, seed)
    c = g.pic.contents
    if flags:
        c.sao_luma, c.sao_chroma = flags
    types = {c.sao[i].type_idx for i in range(3 * (-(-c.width // 64)) * (-(-c.height // 64)))}
    assert types == {0, 1, 2}
    deb = run(g, refs, stop=3)
    assert_planes_equal(run(g, refs), sao_picture(deb, c), name)


def test_sao_all_off_is_the_deblocked_picture_and_borders_stay():
    g = gen(cf.HEVC_C1, 711)
    refs = refs_of(g, 711)
    c = g.pic.contents
    n = 3 * (-(-c.width // 64)) * (-(-c.height // 64))
    deb = run(g, refs, stop=3)
    for i in range(n):
        c.sao[i].type_idx = 0
    assert_planes_equal(run(g, refs), deb, "SAO off")
    # Every CTB edge class 2 (135 degrees) with large offsets: the picture's
    # outer ring keeps its deblocked samples, the interior changes.
    for i in range(n):
        c.sao[i].type_idx, c.sao[i].band_or_class = 2, 2
        c.sao[i].offsets[:] = [7, 7, -7, -7]
    out = run(g, refs)
    for k in range(3):
        a, b = np.asarray(out[k]), np.asarray(deb[k])
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[-1], b[-1])
        assert np.array_equal(a[:, 0], b[:, 0]) and np.array_equal(a[:, -1], b[:, -1])
        assert not np.array_equal(a[1:-1, 1:-1], b[1:-1, 1:-1])


def test_sao_band_offsets_on_a_ramp():
    """A whole-sample copy of a horizontal ramp, filters off but SAO on:
    every CTB's band offsets land on the bands its samples fall in."""
    g = gen(cf.HEVC_C1, 721)
    c = g.pic.contents
    W, H = c.width, c.height
    ramp = np.tile((np.arange(W) * 255 // (W - 1)).astype(np.uint8), (H, 1))
    chroma = np.tile((np.arange(W // 2) * 255 // (W // 2 - 1)).astype(np.uint8), (H // 2, 1))
    refs = [[ramp, chroma, chroma], [ramp, chroma, chroma]]
    no_residual(g)
    c.deblock = 0
    for i in range(c.n_pus):
        c.pus[i].mv[0].x = c.pus[i].mv[0].y = 0
    wc, hc = -(-W // 64), -(-H // 64)
    for i in range(3 * wc * hc):
        c.sao[i].type_idx = 1
        c.sao[i].band_or_class = (5 * i) % 32
        c.sao[i].offsets[:] = [1 + i % 7, -3, 5, -(1 + i % 5)]
    out = run(g, refs)
    for k, base in enumerate((ramp, chroma, chroma)):
        cs = 64 if k == 0 else 32
        want = base.astype(np.int64).copy()
        for ry in range(hc):
            for rx in range(wc):
                p = c.sao[3 * (ry * wc + rx) + k]
                blk = want[ry * cs:(ry + 1) * cs, rx * cs:(rx + 1) * cs]
                band = blk >> 3
                for j in range(4):
                    blk[band == (p.band_or_class + j) % 32] += p.offsets[j]
                np.clip(blk, 0, 255, out=blk)
        assert np.array_equal(np.asarray(out[k], np.int64), want), k


# ---------------------------------------------------------------- transform matrix

# Column 0 of the 32-point matrix of 8.6.4.2 (transMatrix[k][0], k = 0..31).
COL0 = [64, 90, 90, 90, 89, 88, 87, 85, 83, 82, 80, 78, 75, 73, 70, 67, 64, 61, 57, 54, 50, 46, 43, 38, 36, 31, 25,
        22, 18, 13, 9, 4]
# Rows of the standard's listing, in full.
ROW1 = [90, 90, 88, 85, 82, 78, 73, 67, 61, 54, 46, 38, 31, 22, 13, 4, -4, -13, -22, -31, -38, -46, -54, -61, -67,
        -73, -78, -82, -85, -88, -90, -90]
ROW2 = [90, 87, 80, 70, 57, 43, 25, 9, -9, -25, -43, -57, -70, -80, -87, -90, -90, -87, -80, -70, -57, -43, -25, -9,
        9, 25, 43, 57, 70, 80, 87, 90]
ROW3 = [90, 82, 67, 46, 22, -4, -31, -54, -73, -85, -90, -88, -78, -61, -38, -13, 13, 38, 61, 78, 88, 90, 85, 73, 54,
        31, 4, -22, -46, -67, -82, -90]
ROW16 = [64, -64, -64, 64] * 8
ROW31 = [4, -13, 22, -31, 38, -46, 54, -61, 67, -73, 78, -82, 85, -88, 90, -90, 90, -90, 88, -85, 82, -78, 73, -67,
         61, -54, 46, -38, 31, -22, 13, -4]


def test_the_32_point_matrix_is_the_standards():
    """Every entry of the oracle's 32-point matrix (read back from impulse
    responses) is the standard's: entry (k, n) is +-COL0[j] where
    cos((2n+1) k pi / 64) = +-cos(j pi / 64), with the cosine's sign; and
    five rows equal the listing verbatim."""
    from tests.test_oracle_hevc import _matrix_rows_from_impulses

    m = _matrix_rows_from_impulses(32)
    assert m[:, 0].tolist() == COL0
    for k in range(32):
        for n in range(32):
            j = ((2 * n + 1) * k) % 128
            if j <= 32:
                want = COL0[j] if j < 32 else 0
            elif j <= 64:
                want = -COL0[64 - j]
            elif j <= 96:
                want = -COL0[j - 64]
            else:
                want = COL0[128 - j]
            assert m[k, n] == want, (k, n)
    for k, row in ((1, ROW1), (2, ROW2), (3, ROW3), (16, ROW16), (31, ROW31)):
        assert m[k].tolist() == row, k


# ---------------------------------------------------------------- the threaded driver

@pytest.mark.parametrize("threads", [2, 3, 7, 16])
def test_threaded_driver_equals_the_serial_one(threads):
    """orc_hevc_reconstruct_mt (the CPU baseline's row- and block-parallel
    driver) gives the serial driver's bytes at every cut point."""
    for cfg, seed in ((cf.HEVC_C1, 901), (dict(cf.HEVC_C2, width=832, height=480), 902)):
        g = gen(cfg, seed)
        refs = refs_of(g, seed)
        for stop in (1, 2, 3, 0):
            want = run(g, refs, stop)
            got = oracle.reconstruct(rh.RV_CODEC_HEVC, g.pic, list(refs) + [None], stop, threads=threads)
            assert_planes_equal(got, want, f"threads {threads} stop {stop}")
