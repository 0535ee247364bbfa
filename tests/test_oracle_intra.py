"""The oracle's HEVC intra prediction (src/hevc/intra.rs with the caller's
availability, src/hevc/ctu.rs:1727-1833), pinned by closed forms:

- the angle tables: symmetric about mode 18, 0 at modes 10 / 26, +-32 on
  the diagonals, and invAngle = round(8192 / intraPredAngle);
- constant references predict the constant in every mode, size and filter
  setting;
- positive angles on a linear ramp give ramp(x + (y + 1) * angle / 32)
  exactly; mode 18 is the diagonal; modes 10 / 26 copy (chroma) or add the
  boundary filter (luma below 32);
- transposing the references maps mode m to mode 36 - m;
- DC and planar closed forms with constant sides;
- the smoothing decision per (mode, size) against numpy's [1 2 1] filter,
  and strong smoothing against the bilinear filter;
- substitution against a numpy forward fill;
- availability against decoding order taken from the TU list itself;
- the driver: an all-intra picture equals the TUs predicted and
  reconstructed one by one here, in list order, through those pieces.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import _abi
from rivet_hip import configs as cf

L = oracle.lib


def tables():
    a = np.zeros(33, np.int32)
    inv = np.zeros(15, np.int32)
    L.orc_hevc_intra_tables(a.ctypes.data_as(C.POINTER(C.c_int32)), inv.ctypes.data_as(C.POINTER(C.c_int32)))
    return a, inv


def scan_of(left, top, corner):
    """Scan order: p[-1][2n-1] .. p[-1][0], p[-1][-1], p[0][-1] .. p[2n-1][-1]."""
    return np.concatenate([np.asarray(left)[::-1], [corner], np.asarray(top)]).astype(np.uint16)


def pred(scan, n, mode, luma=0, strong=0, bd=8):
    s = np.ascontiguousarray(scan, np.uint16)
    out = np.zeros(n * n, np.int32)
    L.orc_hevc_intra_pred(out.ctypes.data, s.ctypes.data, n, mode, luma, strong, bd)
    return out.reshape(n, n)


def parts(scan, n):
    s = np.asarray(scan, np.int64)
    return s[:2 * n][::-1], s[2 * n + 1:], int(s[2 * n])  # left, top, corner


SIZES = [4, 8, 16, 32]


def test_angle_tables():
    a, inv = tables()
    ang = {m: int(a[m - 2]) for m in range(2, 35)}
    assert ang[10] == ang[26] == 0 and ang[2] == ang[34] == 32 and ang[18] == -32
    assert all(ang[m] == ang[36 - m] for m in range(2, 35))
    assert all(ang[m] > ang[m + 1] for m in range(2, 18)) and all(ang[m] < ang[m + 1] for m in range(18, 34))
    for m in range(11, 26):
        assert inv[m - 11] == round(8192 / ang[m]), m


@pytest.mark.parametrize("n", SIZES)
def test_constant_references_predict_the_constant(n):
    scan = np.full(4 * n + 1, 77, np.uint16)
    for mode in range(35):
        for luma in (0, 1):
            for strong in (0, 1):
                assert (pred(scan, n, mode, luma, strong) == 77).all(), (mode, luma, strong)


@pytest.mark.parametrize("n", SIZES)
def test_positive_angles_on_a_ramp(n):
    """Vertical modes with angle A >= 0 over top[x] = C + 32 x (extended to
    2n) predict C + 32 x + (y + 1) A; the horizontal ones mirror it."""
    a, _ = tables()
    C0 = 100
    rng = np.random.default_rng(n)
    ramp = C0 + 32 * np.arange(2 * n)
    other = rng.integers(0, 1024, 2 * n)
    x = np.arange(n)[None, :]
    y = np.arange(n)[:, None]
    for mode in range(26, 35):
        A = int(a[mode - 2])
        got = pred(scan_of(other, ramp, 55), n, mode, bd=12)
        assert np.array_equal(got, C0 + 32 * x + (y + 1) * A), mode
        got = pred(scan_of(ramp, other, 55), n, 36 - mode, bd=12)
        assert np.array_equal(got, (C0 + 32 * x + (y + 1) * A).T), 36 - mode


@pytest.mark.parametrize("n", SIZES)
def test_transposed_references_mirror_the_mode(n):
    rng = np.random.default_rng(10 + n)
    for trial in range(3):
        left, top, corner = rng.integers(0, 1024, 2 * n), rng.integers(0, 1024, 2 * n), int(rng.integers(0, 1024))
        for luma in (0, 1):
            for mode in range(35):
                m2 = mode if mode < 2 else 36 - mode
                p1 = pred(scan_of(left, top, corner), n, mode, luma, bd=10)
                p2 = pred(scan_of(top, left, corner), n, m2, luma, bd=10)
                assert np.array_equal(p1, p2.T), (mode, luma)


@pytest.mark.parametrize("n", SIZES)
def test_diagonal_vertical_horizontal(n):
    rng = np.random.default_rng(20 + n)
    left, top, corner = rng.integers(0, 256, 2 * n), rng.integers(0, 256, 2 * n), int(rng.integers(0, 256))
    s = scan_of(left, top, corner)
    p18 = pred(s, n, 18)
    for yy in range(n):
        for xx in range(n):
            want = top[xx - yy - 1] if xx > yy else corner if xx == yy else left[yy - xx - 1]
            assert p18[yy, xx] == want
    assert np.array_equal(pred(s, n, 26), np.tile(top[:n], (n, 1)))
    assert np.array_equal(pred(s, n, 10), np.tile(left[:n, None], (1, n)))
    if n < 32:
        # Luma: the boundary filter on the first column (26) / row (10).
        p = pred(s, n, 26, luma=1)
        assert np.array_equal(p[:, 1:], np.tile(top[1:n], (n, 1)))
        assert np.array_equal(p[:, 0], np.clip(top[0] + ((left[:n] - corner) >> 1), 0, 255))
        p = pred(s, n, 10, luma=1)
        assert np.array_equal(p[0, :], np.clip(left[0] + ((top[:n] - corner) >> 1), 0, 255))
    else:
        assert np.array_equal(pred(s, n, 26, luma=1), np.tile(top[:n], (n, 1)))


@pytest.mark.parametrize("n", SIZES)
def test_dc_and_planar_with_constant_sides(n):
    log2n = n.bit_length() - 1
    for Lv, Tv in ((10, 200), (255, 0), (37, 38)):
        s = scan_of(np.full(2 * n, Lv), np.full(2 * n, Tv), 99)
        dc = (Lv + Tv + 1) >> 1
        assert (pred(s, n, 1) == dc).all()
        p = pred(s, n, 1, luma=1)
        if n < 32:
            assert p[0, 0] == (Lv + 2 * dc + Tv + 2) >> 2
            assert (p[0, 1:] == (Tv + 3 * dc + 2) >> 2).all() and (p[1:, 0] == (Lv + 3 * dc + 2) >> 2).all()
            assert (p[1:, 1:] == dc).all()
        else:
            assert (p == dc).all()
        x = np.arange(n)[None, :]
        y = np.arange(n)[:, None]
        want = ((n - x + y) * Lv + (n + x - y) * Tv + n) >> (log2n + 1)
        assert np.array_equal(pred(s, n, 0), want)


def smooth121(scan):
    s = np.asarray(scan, np.int64)
    f = s.copy()
    f[1:-1] = (s[:-2] + 2 * s[1:-1] + s[2:] + 2) >> 2
    return f.astype(np.uint16)


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_smoothing_decision_and_filter(n):
    """Luma planar / angular modes other than 10 and 26 are predicted from
    the [1 2 1]-smoothed references exactly when min(|m - 26|, |m - 10|)
    exceeds 7 (8x8), 1 (16x16) or 0 (32x32), never at 4x4 (8.4.4.2.3)."""
    thres = {4: 99, 8: 7, 16: 1, 32: 0}[n]
    rng = np.random.default_rng(30 + n)
    s = rng.integers(0, 1024, 4 * n + 1).astype(np.uint16)
    for mode in [0] + [m for m in range(2, 35) if m not in (10, 26)]:
        filtered = min(abs(mode - 26), abs(mode - 10)) > thres
        want = pred(smooth121(s) if filtered else s, n, mode, luma=0, bd=10)
        assert np.array_equal(pred(s, n, mode, luma=1, bd=10), want), mode


def bilinear(scan, n=32):
    left, top, c = parts(scan, n)
    i = np.arange(63)
    fl, ft = left.copy(), top.copy()
    fl[:63] = ((63 - i) * c + (i + 1) * left[63] + 32) >> 6
    ft[:63] = ((63 - i) * c + (i + 1) * top[63] + 32) >> 6
    return scan_of(fl, ft, c)


def test_strong_smoothing():
    """32x32 luma planar under the SPS flag: the bilinear filter when both
    sides are flat (|c + far - 2 mid| < 2^(bd - 5)), else [1 2 1]."""
    rng = np.random.default_rng(40)
    n, bd = 32, 8
    c = 120
    i = np.arange(64)
    left = (c + (i * 3) // 4 + rng.integers(-1, 2, 64)).clip(0, 255)
    top = (c - i // 2 + rng.integers(-1, 2, 64)).clip(0, 255)
    s = scan_of(left, top, c)
    assert abs(c + top[63] - 2 * top[31]) < 8 and abs(c + left[63] - 2 * left[31]) < 8
    assert np.array_equal(pred(s, n, 0, luma=1, strong=1, bd=bd), pred(bilinear(s), n, 0, bd=bd))
    assert np.array_equal(pred(s, n, 0, luma=1, strong=0, bd=bd), pred(smooth121(s), n, 0, bd=bd))
    bumpy = top.copy()
    bumpy[31] += 20
    s2 = scan_of(left, bumpy, c)
    assert np.array_equal(pred(s2, n, 0, luma=1, strong=1, bd=bd), pred(smooth121(s2), n, 0, bd=bd))
    # Ramps of slope 2 and 3 are flat (second difference 2 and 3); the far
    # ends matter to every filtered sample. Modes 2 and 34 read them all.
    ramp_l = c + 2 * i + rng.integers(0, 2, 64)
    ramp_t = c - 3 * (i // 2)
    s3 = scan_of(ramp_l.clip(0, 255), ramp_t.clip(0, 255), c)
    for mode in (0, 2, 34):
        assert np.array_equal(pred(s3, n, mode, luma=1, strong=1, bd=bd), pred(bilinear(s3), n, mode, bd=bd)), mode
    # |c + far - 2 mid| = 8 = 2^(bd - 5) on one side: not flat.
    edge_t = top.copy()
    edge_t[63] = 2 * edge_t[31] - c + 8
    s4 = scan_of(left, edge_t, c)
    assert abs(c + int(edge_t[63]) - 2 * int(edge_t[31])) == 8
    assert np.array_equal(pred(s4, n, 0, luma=1, strong=1, bd=bd), pred(smooth121(s4), n, 0, bd=bd))


def test_dc_with_random_sides():
    """8.4.4.2.5: dcVal = (sum of top[0..n) and left[0..n) + n) >> (log2 n
    + 1); luma below 32 filters the first row and column against it."""
    rng = np.random.default_rng(45)
    for n in SIZES:
        log2n = n.bit_length() - 1
        for _ in range(10):
            left, top, corner = rng.integers(0, 1024, 2 * n), rng.integers(0, 1024, 2 * n), int(rng.integers(0, 1024))
            s = scan_of(left, top, corner)
            dc = (int(left[:n].sum() + top[:n].sum()) + n) >> (log2n + 1)
            assert (pred(s, n, 1, bd=10) == dc).all()
            p = pred(s, n, 1, luma=1, bd=10)
            if n < 32:
                assert p[0, 0] == (left[0] + 2 * dc + top[0] + 2) >> 2
                assert np.array_equal(p[0, 1:], (top[1:n] + 3 * dc + 2) >> 2)
                assert np.array_equal(p[1:, 0], (left[1:n] + 3 * dc + 2) >> 2)


def angular_numpy(left, top, corner, n, mode, angle, inv):
    """8.4.4.2.6 without filters, written from the equations: ref[x] for
    x = 0..n is p[-1 + x][-1] (vertical modes); for intraPredAngle < 0 and
    (n * angle) >> 5 < -1, ref[x] = p[-1][-1 + ((x * invAngle + 128) >> 8)]
    for x = (n * angle) >> 5 .. -1, else ref[x] = p[-1 + x][-1] up to 2n;
    predSamples[x][y] from iIdx = ((y + 1) angle) >> 5, iFact = ((y + 1)
    angle) & 31. Horizontal modes swap the roles of x and y."""
    vert = mode >= 18
    main = np.concatenate([[corner], top if vert else left])   # main[k] = ref[k], k >= 0
    side = np.concatenate([[corner], left if vert else top])   # side[k] = p[-1][k - 1]
    ref = {k: int(main[k]) for k in range(0, 2 * n + 1)}
    if angle < 0 and (n * angle) >> 5 < -1:
        for x in range((n * angle) >> 5, 0):
            ref[x] = int(side[(x * inv + 128) >> 8])
    out = np.zeros((n, n), np.int64)
    for y in range(n):
        i_idx, i_fact = ((y + 1) * angle) >> 5, ((y + 1) * angle) & 31
        for x in range(n):
            a = ref[x + i_idx + 1]
            v = ((32 - i_fact) * a + i_fact * ref[x + i_idx + 2] + 16) >> 5 if i_fact else a
            if vert:
                out[y, x] = v
            else:
                out[x, y] = v
    return out


@pytest.mark.parametrize("n", SIZES)
def test_angular_against_the_equations(n):
    a, inv = tables()
    rng = np.random.default_rng(60 + n)
    for _ in range(2):
        left, top, corner = rng.integers(0, 1024, 2 * n), rng.integers(0, 1024, 2 * n), int(rng.integers(0, 1024))
        s = scan_of(left, top, corner)
        for mode in range(2, 35):
            iv = int(inv[mode - 11]) if 11 <= mode <= 25 else 0
            want = angular_numpy(left, top, corner, n, mode, int(a[mode - 2]), iv)
            assert np.array_equal(pred(s, n, mode, bd=10), want), mode


def substitute(scan, avail, n, bd):
    s = np.ascontiguousarray(scan, np.uint16).copy()
    a = np.ascontiguousarray(avail, np.uint8)
    L.orc_hevc_intra_substitute(s.ctypes.data, a.ctypes.data, n, bd)
    return s


@pytest.mark.parametrize("n", SIZES)
def test_substitution(n):
    rng = np.random.default_rng(50 + n)
    m = 4 * n + 1
    assert (substitute(np.zeros(m), np.zeros(m), n, 10) == 512).all()
    for p in (0.1, 0.5, 0.9):
        for _ in range(20):
            a = rng.random(m) < p
            if not a.any():
                continue
            raw = rng.integers(0, 256, m)
            idx = np.where(a, np.arange(m), -1)
            idx = np.maximum.accumulate(idx)       # last available at or before j
            idx[idx < 0] = np.argmax(a)            # leading run: the first available
            want = raw[idx]
            assert np.array_equal(substitute(np.where(a, raw, 0), a, n, 8), want)


INTRA = dict(cf.HEVC_C1, width=128, height=64, intra_percent=40, intra_device=1, b_picture=1, bi_percent=50,
             poc=2, ref_poc=[0, 4], ref_slot=[0, 1])


def _avail(pic, i, n):
    a = np.zeros(4 * n + 1, np.uint8)
    assert L.orc_hevc_intra_avail(C.cast(pic, C.c_void_p), i, a.ctypes.data) == 0
    return a.astype(bool)


@pytest.mark.parametrize("cip", [0, 1])
def test_availability_is_decoding_order(cip):
    """A neighbouring 4x4 luma block is available when it lies in the
    picture, its luma TU comes earlier in the TU list than the block at the
    current TU's top-left (decoding order), and, with constrained intra
    prediction, it belongs to an intra CU."""
    g = rh.GenPicture(rh.RV_CODEC_HEVC, dict(INTRA, constrained_intra_pred=cip), 0x77 + cip)
    c = g.pic.contents
    W, H = c.width, c.height
    order = np.full((H // 4, W // 4), -1)
    intra = np.zeros((H // 4, W // 4), bool)
    for i in range(c.n_cus):
        cu = c.cus[i]
        s = 1 << cu.log2_size
        intra[cu.y // 4:(cu.y + s) // 4, cu.x // 4:(cu.x + s) // 4] = cu.intra != 0
    k = 0
    for i in range(c.n_tus):
        t = c.tus[i]
        if t.comp == 0:
            n = 1 << t.log2
            order[t.y // 4:(t.y + n) // 4, t.x // 4:(t.x + n) // 4] = k
            k += 1
    assert (order >= 0).all()
    checked = 0
    for i in range(c.n_tus):
        t = c.tus[i]
        sc = 1 if t.comp else 0
        n = 1 << t.log2
        cur = order[(t.y << sc) // 4, (t.x << sc) // 4]
        pos = [(t.x - 1, t.y + 2 * n - 1 - j) for j in range(2 * n)] + [(t.x - 1, t.y - 1)] + \
              [(t.x + j, t.y - 1) for j in range(2 * n)]
        want = []
        for xs, ys in pos:
            xl, yl = xs << sc, ys << sc
            ok = xs >= 0 and ys >= 0 and xl < W and yl < H
            ok = ok and order[yl // 4, xl // 4] < cur and (not cip or intra[yl // 4, xl // 4])
            want.append(ok)
        assert np.array_equal(_avail(g.pic, i, n), np.array(want)), i
        checked += 1
    assert checked > 100


def _tu_mode(cu, t):
    if t.comp:
        return cu.chroma_mode
    if cu.part_mode != 3:
        return cu.luma_mode[0]
    h = 1 << (cu.log2_size - 1)
    return cu.luma_mode[(2 if t.y >= cu.y + h else 0) + (1 if t.x >= cu.x + h else 0)]


@pytest.mark.parametrize("bd, strong", [(8, 0), (10, 1)])
def test_driver_on_an_all_intra_picture(bd, strong):
    """Every CU predicted on the device (no PUs): the oracle's picture equals
    the TUs reconstructed here in list order — the reference samples read
    from the picture so far, substituted, predicted with the mode of the
    TU's PU (or the chroma mode), then the inverse transform added."""
    cfg = dict(INTRA, bit_depth=bd, intra_percent=100, strong_intra_smoothing=strong, nz_permille=80)
    g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, 0x99 + bd)
    c = g.pic.contents
    assert c.n_pus == 0
    W, H = c.width, c.height
    planes = [np.zeros((H, W), np.int64), np.zeros((H // 2, W // 2), np.int64), np.zeros((H // 2, W // 2), np.int64)]
    cu_at = np.zeros((H // 4, W // 4), np.int64)
    for i in range(c.n_cus):
        cu = c.cus[i]
        s = 1 << cu.log2_size
        cu_at[cu.y // 4:(cu.y + s) // 4, cu.x // 4:(cu.x + s) // 4] = i
    mx = (1 << bd) - 1
    for i in range(c.n_tus):
        t = c.tus[i]
        sc = 1 if t.comp else 0
        n = 1 << t.log2
        pl = planes[t.comp]
        a = _avail(g.pic, i, n)
        pos = [(t.x - 1, t.y + 2 * n - 1 - j) for j in range(2 * n)] + [(t.x - 1, t.y - 1)] + \
              [(t.x + j, t.y - 1) for j in range(2 * n)]
        raw = np.array([pl[ys, xs] if ok else 0 for (xs, ys), ok in zip(pos, a)])
        s = substitute(raw, a, n, bd)
        cu = c.cus[cu_at[(t.y << sc) // 4, (t.x << sc) // 4]]
        p = pred(s, n, _tu_mode(cu, t), luma=int(t.comp == 0), strong=strong, bd=bd)
        if t.flags & _abi.RV_TU_CODED:
            co = np.zeros(n * n, np.int16)
            for k in range(t.n_nz):
                e = c.nz[t.coef_off + k]
                co[e.pos] = e.val
            if t.flags & _abi.RV_TU_DST:
                L.orc_hevc_idst4(co.ctypes.data_as(C.POINTER(C.c_int16)), 20 - bd)
            else:
                L.orc_hevc_idct(co.ctypes.data_as(C.POINTER(C.c_int16)), t.log2, 20 - bd, t.max_x, t.max_y)
            p = np.clip(p + co.reshape(n, n), 0, mx)
        pl[t.y:t.y + n, t.x:t.x + n] = p
    got = oracle.reconstruct(0, g.pic, [None, None, None], 2)
    for k in range(3):
        assert np.array_equal(got[k].astype(np.int64), planes[k]), k
