"""Pins of the oracle's H.264 picture drivers (CPU, no GPU).

The kernel-level tests pin the oracle's dsp entries; these pin what the
whole-picture drivers add (predict_partition, src/h264/inter.rs:172-332;
the weighting glue, src/h264/recon.rs:142-233; the residual dispatch,
:870-926; the strength derivation and the macroblock loop of
deblock_mb_rows, src/h264/deblock.rs:97-815) against statements written here
from H.264 itself, never from the oracle's code:

- prediction: whole-sample vectors give the clamped reference crop; every
  partition equals 8.4.2.2.1 (the six-tap b, h, j and the quarter-sample
  means of Table 8-12) and 8.4.2.2.2 (eighth-sample chroma) over clamped
  reference coordinates, combined by 8.4.2.3 (default, explicit, or
  implicit weights from 8.4.2.3.1's DistScaleFactor), written here in numpy;
- residual: the reconstruction minus the prediction is each coded block's
  kernel-level-pinned residual at its place;
- boundary strengths of hand-built pictures against 8.7.2.1 by hand;
- deblocking: alpha forced to zero leaves the picture alone; samples
  change only next to filtered edges; the whole picture equals the pinned
  edge filters applied macroblock by macroblock in raster order (luma
  vertical edges left to right, then horizontal top to bottom; chroma
  likewise), odd luma edges skipped under the 8x8 transform, with alpha,
  beta and tC0 from Tables 8-16 / 8-17, typed here.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import _abi
from rivet_hip import configs as cf

# ---------------------------------------------------------------- the standard

# Table 8-16: alpha' and beta' by indexA / indexB = 0..51.
ALPHA_P = [0] * 16 + [4, 4, 5, 6, 7, 8, 9, 10, 12, 13, 15, 17, 20, 22, 25, 28, 32, 36, 40, 45, 50, 56, 63, 71, 80,
                      90, 101, 113, 127, 144, 162, 182, 203, 226, 255, 255]
BETA_P = [0] * 16 + [2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15,
                     15, 16, 16, 17, 17, 18, 18]
# Table 8-17: tC0' by indexA = 0..51 for bS = 1, 2, 3.
TC0_P = [(0, 0, 0)] * 17 + [(0, 0, 1)] * 4 + [(0, 1, 1)] * 2 + [(1, 1, 1)] * 4 + [(1, 1, 2)] * 4 + [
    (1, 2, 3), (1, 2, 3), (2, 2, 3), (2, 2, 4), (2, 3, 4), (2, 3, 4), (3, 3, 5), (3, 4, 6), (3, 4, 6), (4, 5, 7),
    (4, 5, 8), (4, 6, 9), (5, 7, 10), (6, 8, 11), (6, 8, 13), (7, 10, 14), (8, 11, 16), (9, 12, 18), (10, 13, 20),
    (11, 15, 23), (13, 17, 25)]
TAP6 = (1, -5, 20, 20, -5, 1)


def test_typed_tables_have_the_standards_shape():
    assert len(ALPHA_P) == len(BETA_P) == len(TC0_P) == 52
    assert ALPHA_P[16] == 4 and ALPHA_P[51] == 255 and BETA_P[51] == 18 and TC0_P[51] == (13, 17, 25)
    assert TC0_P[17] == (0, 0, 1) and TC0_P[23] == (1, 1, 1) and TC0_P[31] == (1, 2, 3)


# ---------------------------------------------------------------- helpers

def gen(cfg, seed):
    return rh.GenPicture(rh.RV_CODEC_H264, cfg, seed)


def refs_of(g, seed):
    c = g.pic.contents
    return [rh.gen_planes(c.width, c.height, 8, cf.ref_seed(seed, k)) for k in range(2)]


def run(g, refs, stop=0):
    return oracle.reconstruct(rh.RV_CODEC_H264, g.pic, list(refs) + [None], stop)


def assert_planes_equal(got, want, what):
    for k in range(3):
        assert np.array_equal(np.asarray(got[k], np.int64), np.asarray(want[k], np.int64)), f"{what}: plane {k}"


def clamp_window(plane, x0, y0, w, h):
    H, W = plane.shape
    ys = np.clip(np.arange(y0, y0 + h), 0, H - 1)
    xs = np.clip(np.arange(x0, x0 + w), 0, W - 1)
    return plane[ys[:, None], xs[None, :]].astype(np.int64)


def clip1(v):
    return np.clip(v, 0, 255)


def luma_interp(plane, xi, yi, w, h, fx, fy):
    """8.4.2.2.1 for a w x h block at integer position (xi, yi)."""
    g = clamp_window(plane, xi - 2, yi - 2, w + 6, h + 6)  # G at [2 + y][2 + x]
    b1 = sum(t * g[:, k:k + w + 1] for k, t in enumerate(TAP6))  # rows -2..h+3, columns 0..w
    h1 = sum(t * g[k:k + h + 1, :] for k, t in enumerate(TAP6))  # rows 0..h, columns -2..w+3
    j1 = sum(t * b1[k:k + h + 1, :] for k, t in enumerate(TAP6))  # rows 0..h, columns 0..w
    G = g[2:2 + h + 1, 2:2 + w + 1]
    B = clip1((b1[2:2 + h + 1] + 16) >> 5)
    Hh = clip1((h1[:, 2:2 + w + 1] + 16) >> 5)
    J = clip1((j1 + 512) >> 10)
    G0, Gr, Gd = G[:h, :w], G[:h, 1:], G[1:, :w]
    b, s = B[:h, :w], B[1:, :w]
    hh, m = Hh[:h, :w], Hh[:h, 1:]
    j = J[:h, :w]
    avg = lambda a, c: (a + c + 1) >> 1  # noqa: E731
    table = {(0, 0): G0, (0, 1): avg(G0, hh), (0, 2): hh, (0, 3): avg(Gd, hh),
             (1, 0): avg(G0, b), (1, 1): avg(b, hh), (1, 2): avg(hh, j), (1, 3): avg(hh, s),
             (2, 0): b, (2, 1): avg(b, j), (2, 2): j, (2, 3): avg(j, s),
             (3, 0): avg(Gr, b), (3, 1): avg(b, m), (3, 2): avg(j, m), (3, 3): avg(m, s)}
    return table[(fx, fy)]


def chroma_interp(plane, xi, yi, w, h, fx, fy):
    """8.4.2.2.2: eighth-sample bilinear."""
    g = clamp_window(plane, xi, yi, w + 1, h + 1)
    a, b, c, d = g[:h, :w], g[:h, 1:], g[1:, :w], g[1:, 1:]
    return ((8 - fx) * (8 - fy) * a + fx * (8 - fy) * b + (8 - fx) * fy * c + fx * fy * d + 32) >> 6


def cdiv(a, b):
    """C's integer division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def implicit_weights(cur, poc0, poc1, long0, long1):
    """8.4.2.3.1, implicit mode: (w0, w1) from DistScaleFactor (8.4.1.2.3)."""
    if poc1 == poc0 or long0 or long1:
        return 32, 32
    tb = max(-128, min(127, cur - poc0))
    td = max(-128, min(127, poc1 - poc0))
    tx = cdiv(16384 + abs(cdiv(td, 2)), td)
    dsf = max(-1024, min(1023, (tb * tx + 32) >> 6))
    if (dsf >> 2) < -64 or (dsf >> 2) > 128:
        return 32, 32
    return 64 - (dsf >> 2), dsf >> 2


def predict_picture(g, refs):
    c = g.pic.contents
    W, H = c.width, c.height
    out = [np.zeros((H, W), np.int64), np.zeros((H // 2, W // 2), np.int64), np.zeros((H // 2, W // 2), np.int64)]
    mbw = W // 16
    for a in range(c.n_mbs):
        m = c.mbs[a]
        if not m.intra:
            continue
        x, y = (a % mbw) * 16, (a // mbw) * 16
        off = m.pred_off
        for k in range(3):
            n = 16 if k == 0 else 8
            px, py = (x, y) if k == 0 else (x // 2, y // 2)
            out[k][py:py + n, px:px + n] = np.array(c.intra_pred[off:off + n * n], np.int64).reshape(n, n)
            off += n * n
    for i in range(c.n_parts):
        u = c.parts[i]
        lists = [l for l in range(2) if (u.dir >> l) & 1]
        for k in range(3):
            preds = []
            for l in lists:
                ref = refs[c.ref_slot[l][u.ref_idx[l]]][k]
                mvx, mvy = u.mv[l].x, u.mv[l].y
                if k == 0:
                    preds.append(luma_interp(ref, u.x + (mvx >> 2), u.y + (mvy >> 2), u.w, u.h, mvx & 3, mvy & 3))
                else:
                    preds.append(chroma_interp(ref, u.x // 2 + (mvx >> 3), u.y // 2 + (mvy >> 3), u.w // 2, u.h // 2,
                                               mvx & 7, mvy & 7))
            if c.weighting == _abi.RV_H264_EXPLICIT:
                logwd = c.log_wd[k]
                ws = [c.wp_w[l][u.ref_idx[l]][k] for l in lists]
                os_ = [c.wp_o[l][u.ref_idx[l]][k] for l in lists]
            elif c.weighting == _abi.RV_H264_IMPLICIT and len(lists) == 2:
                r0, r1 = u.ref_idx[0], u.ref_idx[1]
                logwd, os_ = 5, [0, 0]
                ws = list(implicit_weights(c.poc, c.ref_poc[0][r0], c.ref_poc[1][r1], c.ref_long_term[0][r0],
                                           c.ref_long_term[1][r1]))
            else:
                logwd = None
            if logwd is None:
                v = preds[0] if len(preds) == 1 else (preds[0] + preds[1] + 1) >> 1
            elif len(preds) == 1:
                if logwd >= 1:
                    v = clip1(((preds[0] * ws[0] + (1 << (logwd - 1))) >> logwd) + os_[0])
                else:
                    v = clip1(preds[0] * ws[0] + os_[0])
            else:
                v = clip1(((preds[0] * ws[0] + preds[1] * ws[1] + (1 << logwd)) >> (logwd + 1))
                          + ((os_[0] + os_[1] + 1) >> 1))
            x, y, w, h = (u.x, u.y, u.w, u.h) if k == 0 else (u.x // 2, u.y // 2, u.w // 2, u.h // 2)
            out[k][y:y + h, x:x + w] = v
    return out


SMALL = dict(cf.H264_C4, width=352, height=288)

PRED_CASES = {
    "B implicit, current POC 2 between 0 and 4": (SMALL, 801, None),
    "B implicit, current POC 3 (unequal weights)": (SMALL, 802, 3),
    "B default, half uni": (dict(SMALL, weighting=0, bi_percent=50), 803, None),
    "P default": (dict(SMALL, b_picture=0, weighting=0, poc=1, ref_poc=[0, 0]), 804, None),
    "B explicit": (dict(SMALL, weighting=_abi.RV_H264_EXPLICIT, bi_percent=50), 805, None),
}


def _pred_case(name):
    cfg, seed, poc = PRED_CASES[name]
    g = gen(cfg, seed)
    c = g.pic.contents
    if poc is not None:
        c.poc = poc
    if c.weighting == _abi.RV_H264_EXPLICIT:
        rng = np.random.default_rng(seed)
        for k in range(3):
            c.log_wd[k] = int(rng.integers(0, 8))
            for l in range(2):
                c.wp_w[l][0][k] = int(rng.integers(-128, 128))
                c.wp_o[l][0][k] = int(rng.integers(-128, 128))
    return g, refs_of(g, seed)


# ---------------------------------------------------------------- prediction

@pytest.mark.parametrize("far", [False, True])
def test_whole_sample_vectors_give_the_clamped_reference_crop(far):
    g = gen(dict(SMALL, b_picture=0, weighting=0, poc=1, ref_poc=[0, 0], intra_percent=0,
                 mv_range=1400 if far else 512), 811 + far)
    refs = refs_of(g, 811)
    c = g.pic.contents
    c.n_blks = 0
    c.deblock = 0
    for i in range(c.n_parts):
        u = c.parts[i]
        u.mv[0].x = (u.mv[0].x >> 3) << 3
        u.mv[0].y = (u.mv[0].y >> 3) << 3
    got = run(g, refs)
    want = [np.zeros_like(p) for p in refs[0]]
    for i in range(c.n_parts):
        u = c.parts[i]
        dx, dy = u.mv[0].x >> 2, u.mv[0].y >> 2
        want[0][u.y:u.y + u.h, u.x:u.x + u.w] = clamp_window(refs[0][0], u.x + dx, u.y + dy, u.w, u.h)
        for k in (1, 2):
            x, y = u.x // 2, u.y // 2
            want[k][y:y + u.h // 2, x:x + u.w // 2] = clamp_window(refs[0][k], x + dx // 2, y + dy // 2, u.w // 2,
                                                                   u.h // 2)
    assert_planes_equal(got, want, "crop")
    if far:  # windows wholly outside the 32-sample padded reference
        assert max(abs(c.parts[i].mv[0].x) for i in range(c.n_parts)) > 4 * (32 + 16)


@pytest.mark.parametrize("name", sorted(PRED_CASES))
def test_prediction_is_the_standards_interpolation_and_weighting(name):
    g, refs = _pred_case(name)
    assert_planes_equal(run(g, refs, stop=1), predict_picture(g, refs), name)


def test_implicit_weights_hand_values():
    assert implicit_weights(2, 0, 4, 0, 0) == (32, 32)
    assert implicit_weights(3, 0, 4, 0, 0) == (16, 48)
    assert implicit_weights(1, 0, 4, 0, 0) == (48, 16)
    assert implicit_weights(3, 0, 4, 1, 0) == (32, 32)
    assert implicit_weights(3, 4, 4, 0, 0) == (32, 32)


# ---------------------------------------------------------------- residual

def test_residual_is_each_coded_blocks_residual_in_place():
    g = gen(dict(SMALL, bi_percent=50, intra_percent=30), 821)
    refs = refs_of(g, 821)
    c = g.pic.contents
    pred = [np.ascontiguousarray(p) for p in run(g, refs, stop=1)]
    got = run(g, refs, stop=2)
    sizes = set()
    for i in range(c.n_blks):
        bk = c.blks[i]
        n = bk.size
        sizes.add((bk.comp > 0, n, bk.has_dc))
        coefs = np.array(c.coefs[bk.coef_off:bk.coef_off + n * n], np.int32)
        pl = pred[bk.comp]
        dst = pl[bk.y:, bk.x:]
        ptr = dst.ctypes.data_as(C.POINTER(C.c_uint8))
        cp = coefs.ctypes.data_as(C.POINTER(C.c_int32))
        if n == 8:
            oracle.lib.orc_h264_residual8(ptr, pl.shape[1], cp, 255)
        else:
            oracle.lib.orc_h264_residual4(ptr, pl.shape[1], cp, bk.dc, bk.has_dc, 255)
    assert {(False, 4, 0), (False, 8, 0), (True, 4, 1)} <= sizes
    assert_planes_equal(got, pred, "residual")


# ---------------------------------------------------------------- boundary strengths

class HandPicture:
    """Macroblocks and partitions by hand. List 0 = (POC 0, POC 8), list 1 =
    (POC 8, POC 0), slots 0 (POC 0) and 1 (POC 8); the picture is POC 4."""

    def __init__(self, mbw, mbh):
        self.mbw, self.mbh = mbw, mbh
        self.mbs = [dict(intra=0, t8=0, qp=30, nz=0) for _ in range(mbw * mbh)]
        self.parts = []

    def mb(self, a, **kw):
        self.mbs[a].update(kw)
        return self

    def part(self, x, y, w, h, l0=None, l1=None):
        self.parts.append((x, y, w, h, l0, l1))
        return self

    def whole(self, a, l0=None, l1=None):
        """One 16x16 partition for (inter) macroblock a."""
        return self.part((a % self.mbw) * 16, (a // self.mbw) * 16, 16, 16, l0, l1)

    def build(self):
        c = _abi.H264Pic()
        c.width, c.height, c.bit_depth, c.poc = self.mbw * 16, self.mbh * 16, 8, 4
        for l, pocs in enumerate([(0, 8), (8, 0)]):
            c.num_ref[l] = 2
            for r, poc in enumerate(pocs):
                c.ref_slot[l][r] = 0 if poc == 0 else 1
                c.ref_poc[l][r] = poc
        c.deblock = 1
        mbs = (_abi.H264Mb * len(self.mbs))()
        for a, m in enumerate(self.mbs):
            mbs[a].intra, mbs[a].transform_8x8, mbs[a].qp, mbs[a].nz_mask = m["intra"], m["t8"], m["qp"], m["nz"]
        parts = (_abi.H264Part * max(1, len(self.parts)))()
        for i, (x, y, w, h, l0, l1) in enumerate(self.parts):
            u = parts[i]
            u.x, u.y, u.w, u.h = x, y, w, h
            u.ref_idx[0] = u.ref_idx[1] = -1
            for l, mv in enumerate((l0, l1)):
                if mv is not None:
                    u.dir |= 1 << l
                    u.ref_idx[l] = mv[0]
                    u.mv[l].x, u.mv[l].y = mv[1], mv[2]
        c.n_mbs, c.n_parts = len(self.mbs), len(self.parts)
        c.mbs, c.parts = mbs, parts
        self._keep = (mbs, parts)
        return C.pointer(c)

    def bs(self):
        return oracle.h264_bs_map(self.build())


V, HZ = 0, 1
MV = (0, 12, -20)


def two_side_by_side(m0, m1, **kw):
    """Macroblocks 0 (x 0..15) and 1 (x 16..31); m = "intra" or (l0, l1)."""
    p = HandPicture(2, 1)
    for a, m in enumerate((m0, m1)):
        if m == "intra":
            p.mb(a, intra=1)
        else:
            p.whole(a, *m)
    for a, extra in kw.items():
        p.mb(int(a[-1]), **extra)
    return p


def test_strength_at_an_intra_macroblock():
    b = two_side_by_side("intra", (MV, None)).bs()
    assert (b[0, 1, V, 0] == 4).all()           # macroblock edge next to intra: 4
    assert (b[0, 0, V, 1:] == 3).all() and (b[0, 0, HZ, 1:] == 3).all()  # inside intra: 3
    assert not b[0, 0, V, 0].any() and not b[0, :, HZ, 0].any()  # picture borders
    assert not b[0, 1, :, 1:].any()             # the inter macroblock's inside: same motion, no coefficients
    b = two_side_by_side((MV, None), "intra").bs()
    assert (b[0, 1, V, 0] == 4).all() and (b[0, 1, V, 1:] == 3).all()


def test_strength_from_coefficients():
    # Block 0 of macroblock 1 (its top-left 4x4) has coefficients.
    b = two_side_by_side((MV, None), (MV, None), mb1=dict(nz=1)).bs()
    want = np.zeros((2, 4, 4), np.uint8)
    want[V, 0, 0] = 2   # the macroblock edge, segment 0
    want[V, 1, 0] = 2   # between blocks 0 and 1
    want[HZ, 1, 0] = 2  # between blocks 0 and 4
    assert np.array_equal(b[0, 1], want) and not b[0, 0].any()
    # Block 3 of macroblock 0 (its top-right 4x4).
    b = two_side_by_side((MV, None), (MV, None), mb0=dict(nz=1 << 3)).bs()
    assert b[0, 1, V, 0].tolist() == [2, 0, 0, 0]
    w0 = np.zeros((2, 4, 4), np.uint8)
    w0[V, 3, 0] = 2
    w0[HZ, 1, 3] = 2
    assert np.array_equal(b[0, 0], w0)


@pytest.mark.parametrize("mv1,want", [((0, 16, -20), 1), ((0, 15, -20), 0), ((0, 12, -24), 1), ((0, 12, -17), 0),
                                      ((1, 12, -20), 1), (MV, 0)])
def test_strength_from_motion(mv1, want):
    b = two_side_by_side((MV, None), (mv1, None)).bs()
    assert (b[0, 1, V, 0] == want).all() and not b[0, :, :, 1:].any()


def test_strength_from_the_number_and_pairing_of_vectors():
    l1 = (0, -8, 4)  # list 1 reference 0 = POC 8
    assert (two_side_by_side((MV, None), (MV, l1)).bs()[0, 1, V, 0] == 1).all()
    # The same two pictures with the lists swapped: 0; one vector 4 apart: 1.
    assert not two_side_by_side((MV, l1), ((1, -8, 4), (1, 12, -20))).bs()[0, 1, V, 0].any()
    assert (two_side_by_side((MV, l1), ((1, -8, 4), (1, 16, -20))).bs()[0, 1, V, 0] == 1).all()


def test_internal_partition_edges_and_the_8x8_transform():
    # Two 8x16 partitions with different vectors: internal vertical edge 2.
    p = HandPicture(1, 1).part(0, 0, 8, 16, (0, 0, 0)).part(8, 0, 8, 16, (0, 8, 0))
    b = p.bs()[0, 0]
    assert (b[V, 2] == 1).all() and not b[V, [0, 1, 3]].any() and not b[HZ].any()
    # 4x4 transform, coefficients in block 5: its four edges get 2.
    p = HandPicture(1, 1).mb(0, nz=1 << 5).part(0, 0, 16, 16, (0, 0, 0))
    b = p.bs()[0, 0]
    assert b[V, 1].tolist() == [0, 2, 0, 0] and b[V, 2].tolist() == [0, 2, 0, 0]
    assert b[HZ, 1].tolist() == [0, 2, 0, 0] and b[HZ, 2].tolist() == [0, 2, 0, 0]
    # 8x8 transform, coefficients in the top-left 8x8 (bits 0, 1, 4, 5), four
    # 8x8 partitions with different vectors: odd edges carry nothing, edge 2
    # gets 2 next to the coded block and 1 elsewhere.
    p = HandPicture(1, 1).mb(0, t8=1, nz=0x33)
    for k, (x, y) in enumerate([(0, 0), (8, 0), (0, 8), (8, 8)]):
        p.part(x, y, 8, 8, (0, 4 * k, 0))
    b = p.bs()[0, 0]
    assert not b[:, [1, 3]].any()
    assert b[V, 2].tolist() == [2, 2, 1, 1] and b[HZ, 2].tolist() == [2, 2, 1, 1]


# ---------------------------------------------------------------- deblocking

DBK_CASES = {
    "C4 distributions": (dict(SMALL), 831, (0, 0)),
    "30 % intra, 8x8 transforms, offsets A / B": (dict(SMALL, intra_percent=30, bi_percent=50), 832, (4, -6)),
    "P default": (dict(SMALL, b_picture=0, weighting=0, poc=1, ref_poc=[0, 0]), 833, (-2, 2)),
}


def _dbk_case(name):
    cfg, seed, (oa, ob) = DBK_CASES[name]
    g = gen(cfg, seed)
    c = g.pic.contents
    c.offset_a, c.offset_b = oa, ob
    return g, refs_of(g, seed)


def test_zero_alpha_leaves_the_picture_unfiltered():
    g, refs = _dbk_case("30 % intra, 8x8 transforms, offsets A / B")
    c = g.pic.contents
    before = run(g, refs, stop=2)
    assert not all(np.array_equal(a, b) for a, b in zip(run(g, refs, stop=3), before))
    for a in range(c.n_mbs):
        c.mbs[a].qp = min(c.mbs[a].qp, 27)
        c.mbs[a].qpc[0] = min(c.mbs[a].qpc[0], 27)
        c.mbs[a].qpc[1] = min(c.mbs[a].qpc[1], 27)
    c.offset_a = -12  # indexA <= 15: alpha' = 0
    assert_planes_equal(run(g, refs, stop=3), before, "alpha = 0")


def _edges(c, b):
    """(plane, vertical, x, y, bS bytes) of every edge segment group the
    standard filters, macroblock by macroblock in raster order."""
    mbw = c.width // 16
    out = []
    for a in range(c.n_mbs):
        m = c.mbs[a]
        mx, my = (a % mbw) * 16, (a // mbw) * 16
        luma_edges = [0, 2] if m.transform_8x8 else [0, 1, 2, 3]
        for vertical in (1, 0):
            for e in luma_edges:
                if e == 0 and (mx if vertical else my) == 0:
                    continue
                seg = b[my // 16, mx // 16, 0 if vertical else 1, e]
                if seg.any():
                    x, y = (mx + 4 * e, my) if vertical else (mx, my + 4 * e)
                    out.append((0, vertical, x, y, seg, a, e))
        for vertical in (1, 0):
            for e in (0, 2):
                if e == 0 and (mx if vertical else my) == 0:
                    continue
                seg = b[my // 16, mx // 16, 0 if vertical else 1, e]
                if seg.any():
                    x, y = (mx // 2 + 2 * e, my // 2) if vertical else (mx // 2, my // 2 + 2 * e)
                    out.append((1, vertical, x, y, seg, a, e))
    return out


def _filter_picture(planes, c, b):
    mbw = c.width // 16
    P = [np.ascontiguousarray(np.array(p, np.uint8)) for p in planes]
    for plane, vertical, x, y, seg, a, e in _edges(c, b):
        m = c.mbs[a]
        nb = (a - 1 if vertical else a - mbw) if e == 0 else a
        for k, arr in ([(0, P[0])] if plane == 0 else [(1, P[1]), (2, P[2])]):
            if k == 0:
                qp_p, qp_q = c.mbs[nb].qp, m.qp
            else:
                qp_p, qp_q = c.mbs[nb].qpc[k - 1], m.qpc[k - 1]
            qpav = (qp_p + qp_q + 1) >> 1
            ia = min(max(qpav + c.offset_a, 0), 51)
            ib = min(max(qpav + c.offset_b, 0), 51)
            if seg[0] == 4:
                tc0 = None
            else:
                tc0 = (C.c_int16 * 4)(*[-1 if s == 0 else TC0_P[ia][s - 1] for s in seg])
            off = y * arr.shape[1] + x
            oracle.lib.orc_h264_deblock(int(k > 0), vertical, arr.ctypes.data_as(C.POINTER(C.c_uint8)), off,
                                        arr.shape[1], ALPHA_P[ia], BETA_P[ib], tc0, 255)
    return P


@pytest.mark.parametrize("name", sorted(DBK_CASES))
def test_deblocking_is_the_edge_filters_in_macroblock_order(name):
    g, refs = _dbk_case(name)
    c = g.pic.contents
    b = oracle.h264_bs_map(g.pic)
    want = _filter_picture(run(g, refs, stop=2), c, b)
    assert_planes_equal(run(g, refs, stop=3), want, name)


@pytest.mark.parametrize("name", sorted(DBK_CASES))
def test_deblocking_changes_only_samples_next_to_filtered_edges(name):
    g, refs = _dbk_case(name)
    c = g.pic.contents
    b = oracle.h264_bs_map(g.pic)
    H, W = c.height, c.width
    ml = np.zeros((H, W), bool)
    mc = np.zeros((H // 2, W // 2), bool)
    for plane, vertical, x, y, seg, a, e in _edges(c, b):
        n = 16 if plane == 0 else 8
        r = 3 if plane == 0 else 1
        mask = ml if plane == 0 else mc
        if vertical:
            mask[y:y + n, x - r:x + r] = True
        else:
            mask[y - r:y + r, x:x + n] = True
    before, after = run(g, refs, stop=2), run(g, refs, stop=3)
    for k, mask in ((0, ml), (1, mc), (2, mc)):
        changed = np.asarray(after[k]) != np.asarray(before[k])
        assert changed.any() and not (changed & ~mask).any(), k
