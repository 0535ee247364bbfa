"""Pins of the HEVC oracle (CPU, no GPU).

Each check is something the oracle's formulas do not restate: the
reference's own closed-form tests (butterfly == matrix product,
src/dsp/hevc.rs:1511-1536; the add_residual known answer :1495-1503),
matrix entries read back from impulse responses and compared with the
standard's published 4/8-point matrices and with its cosine / sine
definitions, filter invariants (constant planes, the symmetric half-sample
filter on ramps, mirror symmetry of the quarter filters, the two-pass
filter against a direct 2-D sum), combiner identities, hand-computed
deblocking and SAO results.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle

L = oracle.lib
I16 = C.POINTER(C.c_int16)
I32 = C.POINTER(C.c_int32)
U8 = C.POINTER(C.c_uint8)


def lcg(seed):
    s = seed
    while True:
        s = (s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        yield s >> 33


def idct(c, n, bd_shift, mx, my):
    a = np.ascontiguousarray(c, dtype=np.int16).copy()
    L.orc_hevc_idct(a.ctypes.data_as(I16), int(math.log2(n)), bd_shift, mx, my)
    return a


def idct_matrix(c, n, bd_shift):
    a = np.ascontiguousarray(c, dtype=np.int16)
    out = np.zeros(n * n, np.int16)
    L.orc_hevc_idct_matrix(a.ctypes.data_as(I16), n, bd_shift, out.ctypes.data_as(I16))
    return out


# ---------------------------------------------------------------- transforms

@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_butterfly_equals_matrix(n):
    """src/dsp/hevc.rs:1511-1536: sparse bounded and dense blocks, 200 trials."""
    r = lcg(7 + n)
    for trial in range(200):
        c = np.zeros(n * n, np.int16)
        if trial % 3 == 0:
            mx, my = n - 1, n - 1
        else:
            mx, my = next(r) % n, next(r) % n
        for y in range(my + 1):
            for x in range(mx + 1):
                if next(r) % 3 == 0:
                    c[y * n + x] = (next(r) % 65536) - 32768
        bd_shift = 20 - 8 - (trial % 3) * 2
        assert np.array_equal(idct(c, n, bd_shift, mx, my), idct_matrix(c, n, bd_shift)), (n, trial)


def _matrix_rows_from_impulses(n):
    """Row k of the n-point matrix, read from the inverse transform of an
    impulse of 128 at (row k, column 0) with bdShift 6: the first stage
    gives M[k][y] exactly and the second multiplies by M[0][0] = 64."""
    rows = []
    for k in range(n):
        c = np.zeros(n * n, np.int16)
        c[k * n] = 128
        out = idct(c, n, 6, 0, k).reshape(n, n)
        assert (out == out[:, :1]).all()  # constant along x (DC row)
        rows.append(out[:, 0].astype(int))
    return np.array(rows)


def test_matrix_4_and_8_point_published_values():
    m4 = _matrix_rows_from_impulses(4)
    assert m4.tolist() == [[64, 64, 64, 64], [83, 36, -36, -83], [64, -64, -64, 64], [36, -83, 83, -36]]
    m8 = _matrix_rows_from_impulses(8)
    assert m8[1].tolist() == [89, 75, 50, 18, -18, -50, -75, -89]
    assert m8[3].tolist() == [75, -18, -89, -50, 50, 89, 18, -75]
    assert m8[5].tolist() == [50, -89, 18, 75, -75, -18, 89, -50]
    assert m8[7].tolist() == [18, -50, 75, -89, 89, -75, 50, -18]


@pytest.mark.parametrize("n", [16, 32])
def test_matrix_follows_cosines_and_is_nearly_orthogonal(n):
    m = _matrix_rows_from_impulses(n)
    for k in range(n):
        for i in range(n):
            ideal = 64.0 if k == 0 else 64 * math.sqrt(2) * math.cos((2 * i + 1) * k * math.pi / (2 * n))
            assert abs(m[k, i] - ideal) < 2.6, (k, i)
    g = m @ m.T
    d = np.diag(g)
    assert np.all(np.abs(d - n * 64 * 64) < n * 64 * 64 * 0.01)
    off = g - np.diag(d)
    assert np.max(np.abs(off)) < n * 64 * 64 * 0.005


def test_dst_matrix_is_the_sine_basis():
    """idst4: an impulse of 128 at (row k, column 0) with bdShift 1 gives
    out[y][i] = (M[0][i] * M[k][y] + 1) >> 1; M[k][n] = round(256/3 sin((2k+1)(n+1) pi / 9))."""
    M = [[round(256 / 3 * math.sin((2 * k + 1) * (n + 1) * math.pi / 9)) for n in range(4)] for k in range(4)]
    assert M == [[29, 55, 74, 84], [74, 74, 0, -74], [84, -29, -74, 55], [55, -84, 74, -29]]
    for k in range(4):
        c = np.zeros(16, np.int16)
        c[k * 4] = 128
        L.orc_hevc_idst4(c.ctypes.data_as(I16), 1)
        want = np.array([[(M[0][i] * M[k][y] + 1) >> 1 for i in range(4)] for y in range(4)])
        assert np.array_equal(c.reshape(4, 4), want), k


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_dc_only_shortcut_equals_full_transform(n):
    for dc in (-32768, -1000, -1, 1, 77, 32767):
        for bd_shift in (12, 10, 8):
            c = np.zeros(n * n, np.int16)
            c[0] = dc
            assert np.array_equal(idct(c, n, bd_shift, 0, 0), idct(c, n, bd_shift, n - 1, n - 1))


def test_add_residual_known_answer():
    """src/dsp/hevc.rs:1495-1503: clipping at both ends, 16-bit samples."""
    a = np.array([100, 65535, 0, 30000], np.uint16)
    res = np.array([-200, 5, -1, 32767], np.int16)
    L.orc_hevc_add_residual(2, C.c_void_p(a.ctypes.data), 2, res.ctypes.data_as(I16), 2, 65535)
    assert a.tolist() == [0, 65535, 0, 62767]


# ------------------------------------------------------------ interpolation

def _filt(fn, ss, src, stride, w, h, frac, shift):
    out = np.zeros(w * h, np.int16)
    fn(ss, out.ctypes.data_as(I16), C.c_void_p(src.ctypes.data), stride, w, h, frac, shift)
    return out.reshape(h, w)


@pytest.mark.parametrize("bd", [8, 10])
def test_filters_on_constant_planes(bd):
    dt = np.uint8 if bd == 8 else np.uint16
    ss = 1 if bd == 8 else 2
    c = (1 << bd) - 1
    src = np.full((24, 24), c, dt)
    shift1 = bd - 8
    for fn, fr in ((L.orc_hevc_qpel_h, range(1, 4)), (L.orc_hevc_qpel_v, range(1, 4)),
                   (L.orc_hevc_epel_h, range(1, 8)), (L.orc_hevc_epel_v, range(1, 8))):
        for f in fr:
            assert (_filt(fn, ss, src, 24, 8, 8, f, shift1) == (c * 64) >> shift1).all()


def test_half_sample_filters_are_exact_on_ramps():
    """The symmetric half-sample taps interpolate a linear ramp exactly:
    s(x) = 2x + 3 gives 64 * (2 (x + 1/2) + 3) at the half position."""
    src = np.tile((2 * np.arange(40) + 3).astype(np.uint8), (16, 1))
    got = _filt(L.orc_hevc_qpel_h, 1, src, 40, 16, 4, 2, 0)
    want = np.array([64 * (2 * (x + 3) + 1 + 3) for x in range(16)])  # window origin 3 left
    assert (got == want).all()
    got = _filt(L.orc_hevc_epel_h, 1, src, 40, 16, 4, 4, 0)
    want = np.array([64 * (2 * (x + 1) + 1 + 3) for x in range(16)])
    assert (got == want).all()


def test_quarter_filters_mirror_each_other():
    r = np.random.default_rng(1)
    src = r.integers(0, 256, (12, 40), dtype=np.uint8)
    mir = np.ascontiguousarray(src[:, ::-1])
    for f in (1, 2, 3):
        # Over the full width, output x at phase f mirrors output 32 - x at
        # phase 4 - f on the mirrored rows.
        full_a = _filt(L.orc_hevc_qpel_h, 1, src, 40, 33, 4, f, 0)
        full_b = _filt(L.orc_hevc_qpel_h, 1, mir, 40, 33, 4, 4 - f, 0)
        assert np.array_equal(full_a, full_b[:, ::-1])
    for f in range(1, 8):
        full_a = _filt(L.orc_hevc_epel_h, 1, src, 40, 37, 4, f, 0)
        full_b = _filt(L.orc_hevc_epel_h, 1, mir, 40, 37, 4, 8 - f, 0)
        assert np.array_equal(full_a, full_b[:, ::-1])


QPEL = {1: [-1, 4, -10, 58, 17, -5, 1, 0], 2: [-1, 4, -11, 40, 40, -11, 4, -1], 3: [0, 1, -5, 17, 58, -10, 4, -1]}


@pytest.mark.parametrize("luma", [True, False])
def test_two_pass_equals_direct_2d_sum_at_8_bits(luma):
    """At 8 bits the first stage is exact (shift1 = 0, no i16 wrap), so the
    fused uni prediction is the rounded 2-D sum over the window."""
    r = np.random.default_rng(5)
    taps = 8 if luma else 4
    w, h = 16, 8
    src = r.integers(0, 256, (h + taps - 1, w + taps - 1), dtype=np.uint8)
    fr = range(1, 4) if luma else range(1, 8)
    for fx in fr:
        for fy in fr:
            if luma:
                fxt, fyt = QPEL[fx], QPEL[fy]
            else:  # Table 8-13 rows are the 4-tap phases
                e = {1: [-2, 58, 10, -2], 2: [-4, 54, 16, -2], 3: [-6, 46, 28, -4], 4: [-4, 36, 36, -4],
                     5: [-4, 28, 46, -6], 6: [-2, 16, 54, -4], 7: [-2, 10, 58, -2]}
                fxt, fyt = e[fx], e[fy]
            k2 = np.outer(fyt, fxt)
            acc = np.zeros((h, w), np.int64)
            for i in range(taps):
                for j in range(taps):
                    acc += k2[i, j] * src[i:i + h, j:j + w].astype(np.int64)
            want = np.clip((acc // 64 + 32) >> 6, 0, 255)  # pred >> 6 (14-bit), then uni rounding
            dst = np.zeros((h, w), np.uint8)
            L.orc_hevc_mc_fused(1, int(luma), C.c_void_p(dst.ctypes.data), w, C.c_void_p(src.ctypes.data),
                                w + taps - 1, w, h, fx, fy, None, 8)
            assert np.array_equal(dst, want), (fx, fy)


@pytest.mark.parametrize("bd", [8, 10])
def test_fraction_zero_is_a_copy(bd):
    dt = np.uint8 if bd == 8 else np.uint16
    ss = 1 if bd == 8 else 2
    r = np.random.default_rng(2)
    src = r.integers(0, 1 << bd, (23, 23)).astype(dt)
    dst = np.zeros((16, 16), dt)
    L.orc_hevc_mc_fused(ss, 1, C.c_void_p(dst.ctypes.data), 16, C.c_void_p(src.ctypes.data), 23, 16, 16, 0, 0,
                        None, bd)
    assert np.array_equal(dst, src[3:19, 3:19])


@pytest.mark.parametrize("bd", [8, 10])
def test_combiner_identities(bd):
    dt = np.uint8 if bd == 8 else np.uint16
    ss = 1 if bd == 8 else 2
    mx = (1 << bd) - 1
    r = np.random.default_rng(3)
    a = r.integers(-2000, 20000, 64).astype(np.int16)
    uni = np.zeros(64, dt)
    L.orc_hevc_uni(ss, C.c_void_p(uni.ctypes.data), 8, a.ctypes.data_as(I16), 8, 8, 14 - bd, mx)
    bi = np.zeros(64, dt)
    L.orc_hevc_bi(ss, C.c_void_p(bi.ctypes.data), 8, a.ctypes.data_as(I16), a.ctypes.data_as(I16), 8, 8, 15 - bd, mx)
    assert np.array_equal(uni, bi)
    for d in range(0, 8):
        lwd = d + 14 - bd
        wu = np.zeros(64, dt)
        L.orc_hevc_weighted_uni(ss, C.c_void_p(wu.ctypes.data), 8, a.ctypes.data_as(I16), 8, 8, lwd, 1 << d, 0, mx)
        assert np.array_equal(wu, uni), d
        b = r.integers(-2000, 20000, 64).astype(np.int16)
        wb = np.zeros(64, dt)
        L.orc_hevc_weighted_bi(ss, C.c_void_p(wb.ctypes.data), 8, a.ctypes.data_as(I16), b.ctypes.data_as(I16),
                               8, 8, lwd, 1 << d, 1 << d, 0, 0, mx)
        db = np.zeros(64, dt)
        L.orc_hevc_bi(ss, C.c_void_p(db.ctypes.data), 8, a.ctypes.data_as(I16), b.ctypes.data_as(I16), 8, 8,
                      15 - bd, mx)
        assert np.array_equal(wb, db), d


# ---------------------------------------------------------------- deblocking

def _luma_line(vals, beta, tc, vertical=True):
    """One 8-sample line p3..q3 replicated over 8 lines; returns line 0."""
    blk = np.tile(np.array(vals, np.uint8), (8, 1)) if vertical else np.tile(np.array(vals, np.uint8), (8, 1)).T.copy()
    stride = 8
    off = 4 if vertical else 4 * stride
    b = (C.c_int32 * 2)(beta, beta)
    t = (C.c_int32 * 2)(tc, tc)
    z = (C.c_uint8 * 2)(0, 0)
    L.orc_hevc_deblock_luma(1, int(vertical), C.c_void_p(blk.ctypes.data), off, stride, b, t, z, z, 255)
    return (blk[0] if vertical else blk[:, 0]).tolist(), blk


def test_luma_deblock_hand_computed():
    step = [100, 100, 100, 100, 110, 110, 110, 110]
    # Normal filter (|p0 - q0| = 10 is not < (5 tc + 1) >> 1 = 10).
    assert _luma_line(step, 64, 4)[0] == [100, 100, 102, 104, 106, 108, 110, 110]
    # Strong filter (10 < 13).
    assert _luma_line(step, 64, 5)[0] == [100, 101, 103, 104, 106, 108, 109, 110]
    # The same through the horizontal-edge kernel.
    assert _luma_line(step, 64, 5, vertical=False)[0] == [100, 101, 103, 104, 106, 108, 109, 110]


def test_luma_deblock_identities():
    r = np.random.default_rng(9)
    for _ in range(50):
        vals = r.integers(0, 256, 8).tolist()
        assert _luma_line(vals, 0, 0)[0] == vals  # bS 0: beta = tc = 0
        flat = [vals[0]] * 8
        assert _luma_line(flat, 64, 24)[0] == flat


def test_chroma_deblock_hand_computed():
    blk = np.tile(np.array([100, 100, 110, 110], np.uint8), (8, 1))
    tc = (C.c_int32 * 4)(3, 3, 0, 3)
    z = (C.c_uint8 * 4)(0, 0, 0, 0)
    L.orc_hevc_deblock_chroma(1, 1, C.c_void_p(blk.ctypes.data), 2, 4, tc, z, z, 255)
    for line in range(8):
        want = [100, 100, 110, 110] if line in (4, 5) else [100, 103, 107, 110]
        assert blk[line].tolist() == want, line


# ---------------------------------------------------------------------- SAO

def test_sao_band_on_a_ramp():
    src = np.arange(256, dtype=np.uint8).reshape(8, 32)
    table = np.zeros(32, np.int16)
    band, offs = 30, [3, -2, 5, 7]  # bands 30, 31, 0, 1 (wraps)
    for k in range(4):
        table[(k + band) & 31] = offs[k]
    dst = np.zeros_like(src)
    L.orc_hevc_sao_band(1, C.c_void_p(dst.ctypes.data), 32, C.c_void_p(src.ctypes.data), 32, 32, 8,
                        table.ctypes.data_as(I16), 3, 255)
    for v in range(256):
        k = {30: 0, 31: 1, 0: 2, 1: 3}.get(v >> 3)
        want = v if k is None else min(255, max(0, v + offs[k]))
        assert dst.flat[v] == want, v


def test_sao_edge_categories():
    # One row, horizontal class: local min, half-min, flat, half-max, local max.
    row = np.array([50, 40, 50, 50, 60, 60, 50, 60, 60, 60, 70, 60, 50], np.uint8)
    src = np.tile(row, (3, 1))
    dst = src.copy()
    off = np.array([4, 2, 0, -2, -4], np.int16)  # by raw edgeIdx 0..4
    stride = row.size
    L.orc_hevc_sao_edge(1, C.c_void_p(dst.ctypes.data), C.c_void_p(src.ctypes.data), stride + 1, stride,
                        row.size - 2, 1, -1, 1, off.ctypes.data_as(I16), 255)
    got = dst[1].tolist()
    # x=1: 40 < both -> edgeIdx 0 (+4); x=2: 50 > 40, == 50 -> 3 (-2);
    # x=3: 50 == 50, < 60 -> 1 (+2); x=4: 60 > 50, == 60 -> 3; x=5: 60 == 60, > 50 -> 3;
    # x=6: 50 < 60, < 60 -> 0 (+4); x=7: 60 > 50, == 60 -> 3; x=8: flat -> 2;
    # x=9: 60 == 60, < 70 -> 1; x=10: 70 > both -> 4 (-4); x=11: 60 < 70, > 50 -> 2.
    assert got == [50, 44, 48, 52, 58, 58, 54, 58, 60, 62, 66, 60, 50]
