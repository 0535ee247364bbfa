"""Pins of the H.264 oracle (CPU, no GPU).

Each check comes from outside the oracle's formulas:
- The six-tap and bilinear filters are exact on linear ramps (8.4.2.2),
  so every one of the 16 luma positions and every chroma eighth has a
  closed form there.
- The centre position j equals a direct 6x6 double sum.
- The 4x4 / 8x8 inverse transforms reproduce the standard's integer basis
  (8.5.12, 8.5.13) on impulses. The 8-bit path wraps in i16.
- A DC of 64 adds 1 (src/h264/transform.rs:263-276).
- implicit_pair gives 64 * tb / td on exact ratios and the (32, 32)
  default in each of its fallback cases (8.4.2.3.1).
- Deblocking lines are computed by hand from 8.7.2.3 / 8.7.2.4.
"""
import ctypes as C

import numpy as np
import pytest

import oracle

L = oracle.lib
I16 = C.POINTER(C.c_int16)
I32 = C.POINTER(C.c_int32)
U8 = C.POINTER(C.c_uint8)
PS = 16  # PRED_STRIDE (src/dsp/h264.rs:22-24)


def u8(a):
    return a.ctypes.data_as(U8)


def qpel(pos, win, w, h, max_=255):
    """win: the (h+5) x (w+5) window, origin two left / two above."""
    win = np.ascontiguousarray(win, np.uint8)
    dst = np.zeros((16, PS), np.uint8)
    L.orc_h264_qpel(pos, u8(dst), u8(win), win.shape[1], w, h, max_)
    return dst[:h, :w].astype(int)


# ------------------------------------------------------------ interpolation

@pytest.mark.parametrize("pos", range(16))
def test_qpel_exact_on_a_ramp(pos):
    """On v = 4x + 4y + c the half-sample filter gives v + 2 per half step
    (taps sum to 32, first moment 16), so position (xf, yf) is v + xf + yf
    exactly, including the rounding of every averaged quarter position."""
    w = h = 16
    yy, xx = np.mgrid[0:h + 5, 0:w + 5]
    win = 4 * xx + 4 * yy + 3
    assert win.max() <= 255
    xf, yf = pos & 3, pos >> 2
    out = qpel(pos, win, w, h)
    want = win[2:2 + h, 2:2 + w] + xf + yf
    assert np.array_equal(out, want), pos


def test_qpel_centre_is_a_direct_6x6_sum():
    """j = clip((sum_k sum_l t_k t_l G(x-2+l, y-2+k) + 512) >> 10): one
    double sum over the window, with no intermediate rounding or clipping
    (the standard's j1 from unclipped b1 values)."""
    t = np.array([1, -5, 20, 20, -5, 1])
    k2 = np.outer(t, t)
    r = np.random.default_rng(5)
    for w, h in [(4, 4), (8, 4), (16, 16)]:
        win = r.integers(0, 256, (h + 5, w + 5))
        out = qpel(10, win, w, h)
        want = np.zeros((h, w), int)
        for y in range(h):
            for x in range(w):
                s = int((k2 * win[y:y + 6, x:x + 6]).sum())
                want[y, x] = min(max((s + 512) >> 10, 0), 255)
        assert np.array_equal(out, want)


def test_qpel_half_positions_clip():
    """b and h clip at both ends: a 0 0 255 255 0 0 step overshoots."""
    line = [0, 0, 0, 255, 255, 0, 0, 0, 0]
    win = np.tile(np.array(line), (9, 1))
    out = qpel(2, win, 4, 4)
    # Output x is the half position between window columns x+2 and x+3:
    # x=0: 15*255 = 3825 -> 120; x=1: 40*255 = 10200 -> 319 -> 255;
    # x=2: 3825 -> 120; x=3: 255 - 5*255 = -1020 -> -32 -> 0.
    assert (out == np.array([120, 255, 120, 0])).all()


def test_qpel_position_zero_is_a_copy():
    r = np.random.default_rng(6)
    win = r.integers(0, 256, (21, 21))
    assert np.array_equal(qpel(0, win, 16, 16), win[2:18, 2:18])


@pytest.mark.parametrize("xf,yf", [(a, b) for a in range(8) for b in range(8)])
def test_chroma_bilinear_exact_on_a_ramp(xf, yf):
    """Bilinear weights sum to 64 with first moments 8 xf and 8 yf, so on
    v = 8x + 8y + c the result is v + xf + yf exactly."""
    yy, xx = np.mgrid[0:9, 0:9]
    src = np.ascontiguousarray(8 * xx + 8 * yy + 5, np.uint8)
    dst = np.zeros((8, PS), np.uint8)
    L.orc_h264_chroma(u8(dst), u8(src), 9, 8, 8, xf, yf)
    assert np.array_equal(dst[:8, :8].astype(int), src[:8, :8].astype(int) + xf + yf)


# ---------------------------------------------------------------- transforms

# The inverse transforms' integer basis (8.5.12.2 / 8.5.13.2): row k is
# the contribution of coefficient k to the eight outputs, times 8.
B4 = np.array([[2, 2, 2, 2], [2, 1, -1, -2], [2, -2, -2, 2], [1, -2, 2, -1]])  # times 2
B8 = np.array([
    [8, 8, 8, 8, 8, 8, 8, 8],
    [12, 10, 6, 3, -3, -6, -10, -12],
    [8, 4, -4, -8, -8, -4, 4, 8],
    [10, -3, -12, -6, 6, 12, 3, -10],
    [8, -8, -8, 8, 8, -8, -8, 8],
    [6, -12, 3, 10, -10, -3, 12, -6],
    [4, -8, 8, -4, -4, 8, -8, 4],
    [3, -6, 10, -12, 12, -10, 6, -3],
])


def _add(n, coeffs, base=128):
    dst = np.full((n, n), base, np.uint8)
    c = np.ascontiguousarray(coeffs, np.int16).reshape(-1)
    f = L.orc_h264_idct4_add if n == 4 else L.orc_h264_idct8_add
    f(u8(dst), n, c.ctypes.data_as(I16), 255)
    return dst.astype(int)


@pytest.mark.parametrize("u", range(4))
@pytest.mark.parametrize("v", range(4))
def test_idct4_impulses_give_the_basis(u, v):
    """Coefficient (row u, column v) = 256 adds 256 * b_u(y) * b_v(x) / 64
    at (y, x), exact in integers (b = B4 / 2)."""
    c = np.zeros((4, 4), int)
    c[u, v] = 256
    got = _add(4, c) - 128
    want = 256 * np.outer(B4[u], B4[v]) // 4 // 64
    assert np.array_equal(got, want)


@pytest.mark.parametrize("u", range(8))
@pytest.mark.parametrize("v", range(8))
def test_idct8_impulses_give_the_basis(u, v):
    """Coefficient (u, v) = 64: the row pass gives 8 * B8[v] (exact: every
    shifted value is a multiple of 8), the column pass B8[u][y] * B8[v][x],
    then (x + 32) >> 6."""
    c = np.zeros((8, 8), int)
    c[u, v] = 64
    got = _add(8, c) - 128
    want = (np.outer(B8[u], B8[v]) + 32) >> 6
    assert np.array_equal(got, want)


def test_idct4_wraps_in_i16_at_8_bits():
    """Row 0 = [30000, 0, 30000, 0]: d0 + d2 = 60000 wraps to -5536, so
    columns 0 and 3 get (-5536 + 32) >> 6 = -86 (i32 arithmetic would give
    +938 and clip to 255)."""
    c = np.zeros((4, 4), int)
    c[0, 0] = c[0, 2] = 30000
    got = _add(4, c)
    assert (got == np.array([42, 128, 128, 42])).all()


def test_dc_64_adds_one():
    """src/h264/transform.rs:263-276: DC 64 alone adds 1 everywhere."""
    for n in (4, 8):
        c = np.zeros((n, n), int)
        c[0, 0] = 64
        assert (_add(n, c) == 129).all()
        dst = np.full((n, n), 128, np.uint8)
        L.orc_h264_dc_add(u8(dst), n, 64, n, 255)
        assert (dst == 129).all()


def test_residual_dc_only_path_equals_transform():
    """residual4/8 with no AC take the DC shortcut; on DC values that stay
    inside i16 it equals the full transform of the same block."""
    r = np.random.default_rng(8)
    for n in (4, 8):
        for dc in r.integers(-3000, 3000, 20):
            co = np.zeros(n * n, np.int32)
            co[0] = dc
            a = np.full((n, n), 100, np.uint8)
            if n == 4:
                L.orc_h264_residual4(u8(a), n, co.ctypes.data_as(I32), 0, 0, 255)
            else:
                L.orc_h264_residual8(u8(a), n, co.ctypes.data_as(I32), 255)
            c = np.zeros((n, n), int)
            c[0, 0] = dc
            assert np.array_equal(a.astype(int), _add(n, c, 100)), (n, dc)


def test_residual4_chroma_dc_replaces_position_0():
    co = np.zeros(16, np.int32)
    co[0] = 9999  # replaced
    co[5] = 64
    a = np.full((4, 4), 100, np.uint8)
    L.orc_h264_residual4(u8(a), 4, co.ctypes.data_as(I32), 128, 1, 255)
    c = np.zeros((4, 4), int)
    c[0, 0] = 128
    c[1, 1] = 64
    assert np.array_equal(a.astype(int), _add(4, c, 100))


# ----------------------------------------------------------------- weights

def implicit(cur, p0, p1, l0=0, l1=0):
    out = (C.c_int32 * 2)()
    L.orc_h264_implicit_pair(cur, p0, p1, l0, l1, out)
    return tuple(out)


def test_implicit_pair():
    # Exact ratios: w1 = 64 * tb / td, w0 = 64 - w1.
    assert implicit(2, 0, 4) == (32, 32)
    assert implicit(1, 0, 4) == (48, 16)
    assert implicit(3, 0, 4) == (16, 48)
    assert implicit(6, 4, 12) == (48, 16)
    assert implicit(-4, 0, 4) == (128, -64)  # w1 = -64 is still allowed
    # Fallbacks to (32, 32): equal POCs, a long-term reference, and a
    # DistScaleFactor >> 2 outside [-64, 128].
    assert implicit(5, 4, 4) == (32, 32)
    assert implicit(1, 0, 4, l0=1) == (32, 32)
    assert implicit(1, 0, 4, l1=1) == (32, 32)
    assert implicit(20, 0, 2) == (32, 32)


# -------------------------------------------------------------- deblocking

def _edge(line, alpha, beta, tc0, chroma=False, vertical=True):
    """The same line p3..p0 | q0..q3 (p1 p0 | q0 q1 for chroma) on every
    line of a 16-line luma / 8-line chroma edge; returns all lines."""
    n = 8 if chroma else 16
    blk = np.tile(np.array(line, np.uint8), (n, 1))
    if not vertical:
        blk = blk.T.copy()
    half = len(line) // 2
    stride = blk.shape[1]
    off = half if vertical else half * stride
    t = None if tc0 is None else (C.c_int16 * 4)(*tc0)
    L.orc_h264_deblock(int(chroma), int(vertical), u8(blk), off, stride, alpha, beta, t, 255)
    return blk if vertical else blk.T


def test_luma_bs_below_4_hand_computed():
    # tc = tc0 + (ap < beta) + (aq < beta) = 3; delta = (16 - 8 + 4) >> 3 = 1;
    # p1 += clip((100 + 105 - 202) >> 1, -1, 1) = 1; q1 += clip(-2, -1, 1).
    line = [100, 100, 101, 103, 107, 109, 110, 110]
    want = [100, 100, 102, 104, 106, 108, 110, 110]
    for vertical in (True, False):
        out = _edge(line, 15, 4, [1, 1, 1, 1], vertical=vertical)
        assert all(r.tolist() == want for r in out)
    # tC0 = -1 on one segment leaves its four lines alone.
    out = _edge(line, 15, 4, [1, -1, 1, 1])
    for i, r in enumerate(out):
        assert r.tolist() == (line if 4 <= i < 8 else want), i
    # The alpha gate: |p0 - q0| = 4 is not < 4.
    assert all(r.tolist() == line for r in _edge(line, 4, 4, [1, 1, 1, 1]))


def test_luma_bs4_hand_computed():
    # Strong on both sides: ap = aq = 3 < beta, |p0 - q0| = 2 < 40/4 + 2.
    line = [100, 101, 103, 104, 106, 108, 109, 110]
    want = [100, 102, 104, 104, 106, 107, 108, 110]
    for vertical in (True, False):
        out = _edge(line, 40, 4, None, vertical=vertical)
        assert all(r.tolist() == want for r in out)
    # alpha 3: the gate passes (2 < 3) but not the strong test (2 < 3/4 + 2
    # fails): p0' = (2 p1 + p0 + q1 + 2) >> 2 = 105, q0' = (2 q1 + q0 + p1 + 2)
    # >> 2 = 106, the rest unchanged.
    out = _edge(line, 3, 4, None)
    assert all(r.tolist() == [100, 101, 103, 105, 106, 108, 109, 110] for r in out)
    # The alpha gate: |p0 - q0| = 2 is not < alpha = 2, so nothing changes.
    assert all(r.tolist() == line for r in _edge(line, 2, 4, None))


def test_chroma_hand_computed():
    # bS < 4: tc = tc0 + 1 = 2; delta = clip((4*6 + (-2) + 4) >> 3, -2, 2) = 2.
    line = [100, 101, 107, 103]
    out = _edge(line, 15, 8, [1, 1, 1, 1], chroma=True)
    assert all(r.tolist() == [100, 103, 105, 103] for r in out)
    # bS 4: p0' = (2 p1 + p0 + q1 + 2) >> 2, q0' = (2 q1 + q0 + p1 + 2) >> 2.
    out = _edge(line, 15, 8, None, chroma=True)
    assert all(r.tolist() == [100, (200 + 101 + 103 + 2) >> 2, (206 + 107 + 100 + 2) >> 2, 103] for r in out)


def test_flat_stays_flat():
    for tc0 in ([2, 2, 2, 2], None):
        assert (_edge([90] * 8, 50, 10, tc0) == 90).all()
        assert (_edge([90] * 4, 50, 10, tc0, chroma=True) == 90).all()
