"""The oracle's scaling of levels, pinned by closed forms.

HEVC scale_coefficients (src/hevc/residual.rs:1223-1283): levelScale from
the quantiser step 2^((qP - 4) / 6), the doubling every 6 qP, the step of
qP 4 + 6k at every size and depth, clipping, scaling-list replication
against an np.kron upsampling, and the nonzero bounds.

H.264 (src/h264/transform.rs:9-84, src/h264/mb.rs:1363-1428): the
reference's own known answers for dequant4x4 (transform.rs:278-295);
normAdjust4x4 / 8x8 against 64 * Qstep * the inverse transform's basis
norms (the tables' derivation), both 8.5.12.1 / 8.5.13.1 branch forms
against the folded multiply-shift, the DC transforms against numpy
Hadamard products with the pre-FRExt (flat-list) scaling formulas, and
wrapping arithmetic.

The drivers: a picture of levels reconstructs exactly like the same
picture with its TUs / blocks scaled here through the pinned functions
(list and matrix selection restated from src/hevc/ctu.rs:2016-2029 and
src/h264/mb.rs:1384-1414, recon.rs:700-716, 1200-1226).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import _abi
from rivet_hip import configs as cf

L = oracle.lib
I16 = C.POINTER(C.c_int16)
I32 = C.POINTER(C.c_int32)
U8 = C.POINTER(C.c_uint8)


# ------------------------------------------------------------------ HEVC

def scale(coeffs, log2, qp, bd, lst=None, dc=16, max_x=None, max_y=None):
    n = 1 << log2
    c = np.ascontiguousarray(np.asarray(coeffs, np.int16).reshape(n, n).copy())
    lp = None
    if lst is not None:
        la = np.ascontiguousarray(np.asarray(lst, np.uint8))
        lp = la.ctypes.data_as(U8)
    L.orc_hevc_scale_coefficients(c.ctypes.data_as(I16), log2, qp, bd, lp, dc,
                                  n - 1 if max_x is None else max_x, n - 1 if max_y is None else max_y)
    return c.astype(np.int64)


def test_level_scale_follows_the_quantiser_step():
    """c = 512 at 8 bits, 4x4, qP = m: d = (512 * 16 * ls + 16) >> 5 =
    256 ls exactly, and ls = round(64 * 2^((m - 4) / 6))."""
    for m in range(6):
        c = np.zeros(16, int)
        c[0] = 512
        d = scale(c, 2, m, 8)[0, 0]
        assert d % 256 == 0
        assert d // 256 == round(64 * 2 ** ((m - 4) / 6)), m


@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("log2", [2, 3, 4, 5])
def test_step_of_qp_4_plus_6k(bd, log2):
    """Flat lists, qP = 4 + 6k: the step is 2^k and the scaled value c *
    2^(15 - bd - log2 + k), rounded half up where the exponent is
    negative (the transform's normalisation, 8.6.2 / 8.6.4)."""
    rng = np.random.default_rng(bd * 10 + log2)
    n = 1 << log2
    for k in range(0, (51 + 6 * (bd - 8) - 4) // 6 + 1):
        c = rng.integers(-60, 61, n * n)
        e = 15 - bd - log2 + k
        if e >= 0:
            want = c * 2 ** e
        else:
            want = np.floor(c * 2.0 ** e + 0.5).astype(np.int64)
        want = np.clip(want, -32768, 32767)
        assert np.array_equal(scale(c, log2, 4 + 6 * k, bd).reshape(-1), want), k


def test_doubling_every_six():
    """Where no rounding happens (c a multiple of 2^bdShift), qP + 6
    doubles every output."""
    rng = np.random.default_rng(7)
    c = rng.integers(-8, 9, 64) * 64
    for qp in range(0, 40):
        a = scale(c, 3, qp, 8)
        b = scale(c, 3, qp + 6, 8)
        assert np.array_equal(b, np.clip(2 * a, -32768, 32767)), qp


def test_clipping():
    c = np.zeros(16, int)
    c[0], c[5] = 32767, -32768
    d = scale(c, 2, 51, 8).reshape(-1)
    assert d[0] == 32767 and d[5] == -32768


@pytest.mark.parametrize("log2, qp, mult", [(2, 4, 2), (3, 4, 1), (4, 10, 1), (5, 16, 1)])
def test_scaling_list_replication(log2, qp, mult):
    """At the qP where a factor of m gives exactly mult * m for a level of
    1 (8 bits), an all-ones TU returns the list upsampled: 4x4 / 8x8 raster,
    16x16 / 32x32 the 8x8 grid repeated over 2x2 / 4x4 squares with the dc
    at (0, 0). The list has distinct, asymmetric entries, so a transposed
    or mis-strided index fails."""
    rng = np.random.default_rng(log2)
    n = 1 << log2
    size = 16 if log2 == 2 else 64
    lst = rng.permutation(np.arange(5, 5 + size * 3, 3))[:size].astype(np.uint8)
    dc = 200
    got = scale(np.ones(n * n, int), log2, qp, 8, lst, dc)
    if log2 <= 3:
        want = lst.reshape(n, n).astype(np.int64)
    else:
        want = np.kron(lst.reshape(8, 8).astype(np.int64), np.ones((n // 8, n // 8), np.int64))
        want[0, 0] = dc
    assert np.array_equal(got, mult * want)


def test_flat_list_of_16_is_flat():
    rng = np.random.default_rng(3)
    for log2 in (2, 3, 4, 5):
        n = 1 << log2
        c = rng.integers(-100, 101, n * n)
        lst = np.full(16 if log2 == 2 else 64, 16, np.uint8)
        assert np.array_equal(scale(c, log2, 29, 10, lst, 16), scale(c, log2, 29, 10))


def test_only_inside_the_bounds():
    c = np.full(64, 5)
    d = scale(c, 3, 30, 8, max_x=2, max_y=4)
    assert (d[:5, :3] != 5).all()
    assert (d[5:, :] == 5).all() and (d[:, 3:] == 5).all()


# ----------------------------------------------------------------- H.264

# Qstep for qP % 6 = 0..5 (the step doubles every 6; H.264's design table).
QSTEP = [0.625, 0.6875, 0.8125, 0.875, 1.0, 1.125]
# Inverse transform bases (8.5.12.2 / 8.5.13.2), as in test_oracle_h264.py:
# B4 at twice scale, B8 at eight times.
B4 = np.array([[2, 2, 2, 2], [2, 1, -1, -2], [2, -2, -2, 2], [1, -2, 2, -1]]) / 2
B8 = np.array([
    [8, 8, 8, 8, 8, 8, 8, 8],
    [12, 10, 6, 3, -3, -6, -10, -12],
    [8, 4, -4, -8, -8, -4, 4, 8],
    [10, -3, -12, -6, 6, 12, 3, -10],
    [8, -8, -8, 8, 8, -8, -8, 8],
    [6, -12, 3, 10, -10, -3, 12, -6],
    [4, -8, 8, -4, -4, 8, -8, 4],
    [3, -6, 10, -12, 12, -10, 6, -3],
]) / 8


def ls4(sl, lst, m, pos):
    return L.orc_h264_level_scale4(sl, lst, m, pos)


def ls8(sl, lst, m, pos):
    return L.orc_h264_level_scale8(sl, lst, m, pos)


def test_reference_known_answers():
    """transform.rs:278-295: level 3 at position 0 of a flat 4x4 scales to
    768 at qP 28 and 96 at qP 10."""
    for qp, want in ((28, 768), (10, 96)):
        got = L.orc_h264_dequant_level(3, ls4(None, 0, qp % 6, 0), qp // 6 + 2)
        assert got == want


def test_norm_adjust_4x4_from_the_basis():
    """Flat LevelScale4x4 / 16 = normAdjust4x4(m, i, j), within 1 of 64 *
    Qstep(m) / (|b_i| |b_j|) for the 4x4 inverse basis rows b — which puts
    each value in its (even, odd, mixed) class."""
    norms = np.sqrt((B4 ** 2).sum(1))
    for m in range(6):
        for pos in range(16):
            i, j = divmod(pos, 4)
            v = ls4(None, 0, m, pos)
            assert v % 16 == 0
            pred = 64 * QSTEP[m] / (norms[i] * norms[j])
            assert abs(v // 16 - pred) <= 1, (m, i, j, v // 16, pred)


def test_norm_adjust_8x8_from_the_basis():
    """The same for 8x8 with 4 * 64 * Qstep(m) / (|b_i| |b_j|): the 8x8
    scale carries two more bits, taken back by its shift qP / 6 instead of
    qP / 6 + 2."""
    norms = np.sqrt((B8 ** 2).sum(1))
    for m in range(6):
        for pos in range(64):
            i, j = divmod(pos, 8)
            v = ls8(None, 0, m, pos)
            assert v % 16 == 0
            pred = 4 * 64 * QSTEP[m] / (norms[i] * norms[j])
            assert abs(v // 16 - pred) <= 1, (m, i, j, v // 16, pred)


def test_norm_classes_are_exact_classes():
    """Positions of one class share one value (normAdjust depends on the
    class only): the 4x4 classes by parity, the 8x8 ones by (i % 4, j % 4)
    up to symmetry."""
    for m in range(6):
        g = {}
        for pos in range(64):
            i, j = divmod(pos, 8)
            key = tuple(sorted((i % 4 if i % 2 == 0 else 1, j % 4 if j % 2 == 0 else 1)))
            g.setdefault(key, set()).add(ls8(None, 0, m, pos))
        assert all(len(v) == 1 for v in g.values()) and len(g) == 6


def _hs(lists4, lists8):
    s = _abi.H264Scaling()
    for k in range(6):
        s.list4[k][:] = [int(x) for x in lists4[k]]
    for k in range(2):
        s.list8[k][:] = [int(x) for x in lists8[k]]
    return s


def test_weights_multiply_by_raster_position():
    rng = np.random.default_rng(11)
    l4 = rng.integers(4, 65, (6, 16))
    l8 = rng.integers(4, 65, (2, 64))
    s = _hs(l4, l8)
    for m in range(6):
        for lst in range(6):
            for pos in range(16):
                assert ls4(C.byref(s), lst, m, pos) == l4[lst][pos] * ls4(None, 0, m, pos) // 16
        for lst in range(2):
            for pos in range(64):
                assert ls8(C.byref(s), lst, m, pos) == l8[lst][pos] * ls8(None, 0, m, pos) // 16


@pytest.mark.parametrize("qp", range(52))
def test_folded_shift_equals_both_branches(qp):
    """8.5.12.1 (4x4): qP >= 24: (c * LS) << (qP/6 - 4), else (c * LS +
    2^(3 - qP/6)) >> (4 - qP/6); 8.5.13.1 (8x8): thresholds 36 and 6 —
    against the reference's (c * LS << shift + 32) >> 6."""
    rng = np.random.default_rng(qp)
    q6 = qp // 6
    for c in rng.integers(-300, 301, 40):
        c = int(c)
        for pos in (0, 1, 5, 7):
            t = ls4(None, 0, qp % 6, pos)
            want = (c * t) << (q6 - 4) if qp >= 24 else (c * t + (1 << (3 - q6))) >> (4 - q6)
            assert L.orc_h264_dequant_level(c, t, q6 + 2) == want
        for pos in (0, 9, 18, 27, 63):
            t = ls8(None, 1, qp % 6, pos)
            want = (c * t) << (q6 - 6) if qp >= 36 else (c * t + (1 << (5 - q6))) >> (6 - q6)
            assert L.orc_h264_dequant_level(c, t, q6) == want


def test_dequant_level_wraps_in_32_bits():
    big, t, sh = 1 << 20, 4000, 10
    v = (big * t << sh) & 0xFFFFFFFF
    v = (v + 32) & 0xFFFFFFFF
    v = v - (1 << 32) if v >= 1 << 31 else v
    assert L.orc_h264_dequant_level(big, t, sh) == v >> 6


# Hadamard matrices of the DC transforms (8.5.10, 8.5.11.1).
H4 = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]])
H2 = np.array([[1, 1], [1, -1]])
V0 = [10, 11, 13, 14, 16, 18]  # LevelScale(m, 0, 0) before FRExt (flat)


@pytest.mark.parametrize("qp", range(52))
def test_luma_dc_against_matrix_and_pre_frext_form(qp):
    """f = H c H (c the 4x4 raster of DC levels); with flat lists
    (LevelScale4x4(m, 0, 0) = 16 v0(m)) the result equals the pre-FRExt
    8.5.6: QP >= 12: (f v0) << (QP/6 - 2), else (f v0 + 2^(1 - QP/6)) >>
    (2 - QP/6)."""
    rng = np.random.default_rng(100 + qp)
    c = rng.integers(-40, 41, 16).astype(np.int32)
    f = H4 @ c.reshape(4, 4).astype(np.int64) @ H4.T
    v0 = V0[qp % 6]
    q6 = qp // 6
    want = (f * v0) << (q6 - 2) if qp >= 12 else (f * v0 + (1 << (1 - q6))) >> (2 - q6)
    got = c.copy()
    L.orc_h264_luma_dc(got.ctypes.data_as(I32), ls4(None, 0, qp % 6, 0), qp)
    assert np.array_equal(got.reshape(4, 4), want)


@pytest.mark.parametrize("qp", range(52))
def test_chroma_dc_against_matrix_and_pre_frext_form(qp):
    """f = H2 c H2 (2x2 raster); flat: ((f v0) << (QP/6)) >> 1."""
    rng = np.random.default_rng(200 + qp)
    c = rng.integers(-60, 61, 4).astype(np.int32)
    f = H2 @ c.reshape(2, 2).astype(np.int64) @ H2.T
    want = ((f * V0[qp % 6]) << (qp // 6)) >> 1
    got = c.copy()
    L.orc_h264_chroma_dc(got.ctypes.data_as(I32), ls4(None, 1, qp % 6, 0), qp)
    assert np.array_equal(got.reshape(2, 2), want)


# --------------------------------------------------------------- drivers

HEVC_LEVELS = dict(cf.HEVC_C1, width=128, height=64, intra_percent=50, levels=1, lap_b=8)


def _hevc_scaled_copy(pic):
    """The picture with every levels TU scaled here, flags cleared: the list
    and DC chosen as src/hevc/ctu.rs:2016-2029 (matrixId = cIdx for intra,
    3 + cIdx otherwise; sizeId = log2 - 2)."""
    c = pic.contents
    tus = [_abi.HevcTu.from_buffer_copy(c.tus[i]) for i in range(c.n_tus)]
    nz = [(c.nz[i].pos, c.nz[i].val) for i in range(c.n_nz)]
    sl = c.scaling.contents if c.scaling else None
    bd = c.bit_depth
    for t in tus:
        if not (t.flags & _abi.RV_TU_LEVELS):
            continue
        n = 1 << t.log2
        raster = np.zeros(n * n, np.int16)
        for k in range(t.n_nz):
            p, v = nz[t.coef_off + k]
            raster[p] = v
        lst, dc = None, 16
        if sl is not None:
            mid = t.comp + (0 if t.flags & _abi.RV_TU_INTRA else 3)
            lst = list(sl.list4[mid]) if t.log2 == 2 else list(sl.list8[t.log2 - 3][mid])
            if t.log2 >= 4:
                dc = sl.dc[t.log2 - 4][mid]
        d = scale(raster, t.log2, t.qp, bd, lst, dc, t.max_x, t.max_y).reshape(-1)
        for k in range(t.n_nz):
            p, _ = nz[t.coef_off + k]
            nz[t.coef_off + k] = (p, int(d[p]))
        t.flags &= ~(_abi.RV_TU_LEVELS | _abi.RV_TU_INTRA)
    new = _abi.HevcPic.from_buffer_copy(c)
    tu_arr = (_abi.HevcTu * len(tus))(*tus)
    nz_arr = (_abi.Nz * len(nz))(*[_abi.Nz(p, v) for p, v in nz])
    new.tus = C.cast(tu_arr, C.POINTER(_abi.HevcTu))
    new.nz = C.cast(nz_arr, C.POINTER(_abi.Nz))
    new.scaling = None
    return C.pointer(new), (tu_arr, nz_arr)


@pytest.mark.parametrize("bd", [8, 10])
@pytest.mark.parametrize("scaling", [0, 1])
def test_hevc_driver_levels_equal_prescaled(bd, scaling):
    cfg = dict(HEVC_LEVELS, bit_depth=bd, scaling=scaling)
    g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, 0x1234 + bd + scaling)
    c = g.pic.contents
    assert any(c.tus[i].flags & _abi.RV_TU_LEVELS for i in range(c.n_tus))
    assert any(c.tus[i].flags & _abi.RV_TU_INTRA for i in range(c.n_tus))
    refs = [rh.gen_planes(c.width, c.height, bd, 77 + k) for k in range(2)]
    want_pic, keep = _hevc_scaled_copy(g.pic)
    a = oracle.reconstruct(0, g.pic, refs + [None], 2)
    b = oracle.reconstruct(0, want_pic, refs + [None], 2)
    for k in range(3):
        assert np.array_equal(a[k], b[k])
    # And the levels do change the picture (not scaled as if flat).
    if scaling:
        flat = _abi.HevcPic.from_buffer_copy(c)
        flat.scaling = None
        f = oracle.reconstruct(0, C.pointer(flat), refs + [None], 2)
        assert not np.array_equal(f[0], a[0])


H264_LEVELS = dict(cf.H264_C4, width=96, height=64, intra_percent=40, levels=1, lap_b=8)


def _h264_scaled_copy(pic):
    """Every levels block scaled here (MbDequant::for_mb, src/h264/mb.rs:
    1384-1414: lists Y/Cb/Cr intra 0..2, inter 3..5, 8x8 Y intra 0, inter
    1; DC groups: Intra_16x16 luma with list 0, recon.rs:700-716, chroma
    with the macroblock's list, recon.rs:1200-1226)."""
    c = pic.contents
    mbw = c.width // 16
    coefs = np.ctypeslib.as_array(c.coefs, (c.n_coefs,)).astype(np.int32).copy()
    blks = [_abi.H264Blk.from_buffer_copy(c.blks[i]) for i in range(c.n_blks)]
    sl = C.byref(c.scaling.contents) if c.scaling else None
    for b in blks:
        if not (b.flags & _abi.RV_BLK_LEVELS):
            continue
        mb = c.mbs[(b.y // 8) * mbw + b.x // 8] if b.comp else c.mbs[(b.y // 16) * mbw + b.x // 16]
        qp = mb.qpc[b.comp - 1] if b.comp else mb.qp
        inter = not mb.intra
        n = b.size * b.size
        lv = coefs[b.coef_off:b.coef_off + n].copy()
        if b.size == 8:
            out = [L.orc_h264_dequant_level(int(lv[i]), ls8(sl, int(inter), qp % 6, i), qp // 6) for i in range(64)]
        else:
            lst = (3 if inter else 0) + b.comp
            out = [L.orc_h264_dequant_level(int(lv[i]), ls4(sl, lst, qp % 6, i), qp // 6 + 2) for i in range(16)]
        if b.has_dc:
            if b.comp:
                g = coefs[b.dc:b.dc + 4].copy()
                L.orc_h264_chroma_dc(g.ctypes.data_as(I32), ls4(sl, (3 if inter else 0) + b.comp, qp % 6, 0), qp)
                b.dc = int(g[((b.y % 8) // 4) * 2 + (b.x % 8) // 4])
            else:
                g = coefs[b.dc:b.dc + 16].copy()
                L.orc_h264_luma_dc(g.ctypes.data_as(I32), ls4(sl, 0, qp % 6, 0), qp)
                b.dc = int(g[((b.y % 16) // 4) * 4 + (b.x % 16) // 4])
        coefs[b.coef_off:b.coef_off + n] = out
        b.flags = 0
    new = _abi.H264Pic.from_buffer_copy(c)
    arr = (_abi.H264Blk * len(blks))(*blks)
    new.blks = C.cast(arr, C.POINTER(_abi.H264Blk))
    new.coefs = coefs.ctypes.data_as(I32)
    new.scaling = None
    return C.pointer(new), (arr, coefs)


@pytest.mark.parametrize("scaling", [0, 1])
def test_h264_driver_levels_equal_prescaled(scaling):
    g = rh.GenPicture(rh.RV_CODEC_H264, dict(H264_LEVELS, scaling=scaling), 0x4321 + scaling)
    c = g.pic.contents
    blks = [c.blks[i] for i in range(c.n_blks)]
    assert any(b.has_dc and b.comp == 0 for b in blks)
    assert any(b.has_dc and b.comp != 0 for b in blks)
    assert any(b.size == 8 for b in blks)
    refs = [rh.gen_planes(c.width, c.height, 8, 91 + k) for k in range(2)]
    want_pic, keep = _h264_scaled_copy(g.pic)
    a = oracle.reconstruct(1, g.pic, refs + [None], 2)
    b = oracle.reconstruct(1, want_pic, refs + [None], 2)
    for k in range(3):
        assert np.array_equal(a[k], b[k])
