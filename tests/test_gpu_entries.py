"""Per-entry parity: every HevcDsp / H264Dsp entry of the HIP tier against
the oracle's restatement of the same scalar kernel, bit for bit, on seeded
inputs. Sizes cover the reference's block shapes (luma 4..64 with the AMP
widths 12/24/48, chroma 2..32), every fraction, 8/10/12-bit samples, and
the edge cases the reference tests (DC-only transforms, zero bounds,
exempt deblocking sides, bS 0 segments).
"""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

L = oracle.lib
I16 = C.POINTER(C.c_int16)
I32 = C.POINTER(C.c_int32)
U8 = C.POINTER(C.c_uint8)
VP = C.c_void_p


def vp(a):
    return VP(a.ctypes.data)


def i16(a):
    return a.ctypes.data_as(I16)


@pytest.fixture(scope="module")
def dsp(device):
    from rivet_hip.dsp import H264Dsp, HevcDsp

    return {8: HevcDsp(np.uint8), 10: HevcDsp(np.uint16), 12: HevcDsp(np.uint16), "h264": H264Dsp()}


def samples(r, shape, bd):
    return r.integers(0, 1 << bd, shape).astype(np.uint8 if bd == 8 else np.uint16)


def inter14(r, n):
    """14-bit intermediates as the first filter stage produces them."""
    return r.integers(-8000, 24000, n).astype(np.int16)


LUMA_SIZES = [(4, 4), (8, 4), (4, 8), (8, 8), (12, 16), (16, 16), (24, 32), (32, 32), (48, 64), (64, 64)]
CHROMA_SIZES = [(2, 2), (4, 2), (2, 4), (4, 4), (6, 8), (8, 8), (12, 16), (16, 16), (32, 32)]
BDS = [8, 10, 12]


# ------------------------------------------------------------- transforms

@pytest.mark.parametrize("log2", [2, 3, 4, 5])
@pytest.mark.parametrize("bd", [8, 10])
def test_idct(dsp, log2, bd):
    r = np.random.default_rng(100 + log2 * 10 + bd)
    n = 1 << log2
    for trial in range(24):
        mx, my = (0, 0) if trial < 3 else (int(r.integers(0, n)), int(r.integers(0, n)))
        c = np.zeros((n, n), np.int16)
        dens = [0.1, 0.5, 1.0][trial % 3]
        m = r.random((my + 1, mx + 1)) < dens
        c[:my + 1, :mx + 1] = np.where(m, r.integers(-3000, 3000, (my + 1, mx + 1)), 0)
        if trial % 5 == 4:
            c[:my + 1, :mx + 1] = r.integers(-32768, 32768, (my + 1, mx + 1))  # clipping
        # Garbage beyond the bounds is ignored by both (idct1 never reads it).
        if mx + 1 < n:
            c[:, mx + 1:] = r.integers(-100, 100, (n, n - mx - 1))
        got = c.reshape(-1).copy()
        dsp[8].idct[log2 - 2](got, 20 - bd, mx, my)
        want = c.reshape(-1).copy()
        L.orc_hevc_idct(i16(want), log2, 20 - bd, mx, my)
        assert np.array_equal(got, want), (log2, bd, trial, mx, my)


def test_idst4(dsp):
    r = np.random.default_rng(7)
    for trial in range(64):
        c = r.integers(-5000, 5000, 16).astype(np.int16)
        if trial % 8 == 0:
            c = r.integers(-32768, 32768, 16).astype(np.int16)
        got = c.copy()
        dsp[8].idst4(got, 12 - (trial % 3) * 2)
        want = c.copy()
        L.orc_hevc_idst4(i16(want), 12 - (trial % 3) * 2)
        assert np.array_equal(got, want), trial


@pytest.mark.parametrize("bd", BDS)
def test_add_residual(dsp, bd):
    r = np.random.default_rng(11 + bd)
    d = dsp[bd]
    ss = 1 if bd == 8 else 2
    mx = (1 << bd) - 1
    for n in (4, 8, 16, 32):
        stride = n + 13
        dst = samples(r, n * stride, bd)
        res = r.integers(-(1 << bd), 1 << bd, n * n).astype(np.int16)
        got = dst.copy()
        d.add_residual(got, stride, res, n, mx)
        want = dst.copy()
        L.orc_hevc_add_residual(ss, vp(want), stride, i16(res), n, mx)
        assert np.array_equal(got, want), n


# ----------------------------------------------------------- interpolation

@pytest.mark.parametrize("bd", BDS)
@pytest.mark.parametrize("luma", [True, False])
def test_one_stage_filters(dsp, bd, luma):
    r = np.random.default_rng(21 + bd + luma)
    d = dsp[bd]
    ss = 1 if bd == 8 else 2
    taps = 8 if luma else 4
    shift1 = min(bd, 12) - 8
    pre = "qpel" if luma else "epel"
    for (w, h) in (LUMA_SIZES if luma else CHROMA_SIZES):
        stride = w + taps + 5
        src = samples(r, stride * (h + taps), bd)
        for frac in range(1, 4 if luma else 8):
            for kind in ("h", "v"):
                got = np.zeros(w * h, np.int16)
                getattr(d, f"{pre}_{kind}")(got, src, stride, w, h, frac, shift1)
                want = np.zeros(w * h, np.int16)
                getattr(L, f"orc_hevc_{pre}_{kind}")(ss, i16(want), vp(src), stride, w, h, frac, shift1)
                assert np.array_equal(got, want), (w, h, frac, kind)
        got = np.zeros(w * h, np.int16)
        getattr(d, f"{pre}_copy")(got, src, stride, w, h, 14 - bd)
        want = np.zeros(w * h, np.int16)
        L.orc_hevc_copy(ss, i16(want), vp(src), stride, w, h, 14 - bd)
        assert np.array_equal(got, want), (w, h, "copy")


@pytest.mark.parametrize("luma", [True, False])
def test_second_stage_filters(dsp, luma):
    r = np.random.default_rng(31 + luma)
    taps = 8 if luma else 4
    fn = dsp[8].qpel_v2 if luma else dsp[8].epel_v2
    ofn = L.orc_hevc_qpel_v2 if luma else L.orc_hevc_epel_v2
    for (w, h) in (LUMA_SIZES if luma else CHROMA_SIZES):
        src = inter14(r, w * (h + taps - 1))
        for frac in range(1, 4 if luma else 8):
            got = np.zeros(w * h, np.int16)
            fn(got, src, w, w, h, frac)
            want = np.zeros(w * h, np.int16)
            ofn(i16(want), i16(src), w, w, h, frac)
            assert np.array_equal(got, want), (w, h, frac)


@pytest.mark.parametrize("bd", BDS)
def test_combiners(dsp, bd):
    r = np.random.default_rng(41 + bd)
    d = dsp[bd]
    ss = 1 if bd == 8 else 2
    mx = (1 << bd) - 1
    for (w, h) in [(4, 4), (8, 4), (12, 16), (16, 16), (64, 64)]:
        stride = w + 9
        a, b = inter14(r, w * h), inter14(r, w * h)
        base = samples(r, stride * h, bd)
        cases = [
            ("uni", (a, w, h, 14 - bd, mx), lambda o: L.orc_hevc_uni(ss, vp(o), stride, i16(a), w, h, 14 - bd, mx)),
            ("bi", (a, b, w, h, 15 - bd, mx),
             lambda o: L.orc_hevc_bi(ss, vp(o), stride, i16(a), i16(b), w, h, 15 - bd, mx)),
        ]
        for log2_wd in (0, 1, 6 + min(bd, 12) - 8, 14):
            wt, o0 = int(r.integers(-128, 128)), int(r.integers(-512, 512))
            w1, o1 = int(r.integers(-128, 128)), int(r.integers(-512, 512))
            cases.append(("weighted_uni", (a, w, h, log2_wd, wt, o0, mx),
                          lambda o, l=log2_wd, wt=wt, o0=o0: L.orc_hevc_weighted_uni(ss, vp(o), stride, i16(a), w, h,
                                                                                     l, wt, o0, mx)))
            cases.append(("weighted_bi", (a, b, w, h, log2_wd, wt, w1, o0, o1, mx),
                          lambda o, l=log2_wd, wt=wt, w1=w1, o0=o0, o1=o1: L.orc_hevc_weighted_bi(
                              ss, vp(o), stride, i16(a), i16(b), w, h, l, wt, w1, o0, o1, mx)))
        for name, args, ref in cases:
            got = base.copy()
            getattr(d, name)(got, stride, *args)
            want = base.copy()
            ref(want)
            assert np.array_equal(got, want), (name, w, h, args[-4:])


@pytest.mark.parametrize("bd", BDS)
@pytest.mark.parametrize("luma", [True, False])
def test_fused_mc(dsp, bd, luma):
    """qpel_uni / epel_uni / qpel_bi / epel_bi at every (fx, fy)."""
    r = np.random.default_rng(51 + bd + 7 * luma)
    d = dsp[bd]
    ss = 1 if bd == 8 else 2
    taps = 8 if luma else 4
    pre = "qpel" if luma else "epel"
    nf = 4 if luma else 8
    sizes = LUMA_SIZES if luma else CHROMA_SIZES
    for k, (w, h) in enumerate(sizes):
        sstride, dstride = w + taps + 3, w + 5
        src = samples(r, sstride * (h + taps - 1), bd)
        base = samples(r, dstride * h, bd)
        other = inter14(r, w * h)
        fracs = [(fx, fy) for fx in range(nf) for fy in range(nf)]
        if not luma and k % 2:
            fracs = fracs[::3]  # chroma: every third pair on half the sizes
        for fx, fy in fracs:
            for bi in (False, True):
                got = base.copy()
                if bi:
                    getattr(d, f"{pre}_bi")(got, dstride, src, sstride, w, h, fx, fy, None, other, bd)
                else:
                    getattr(d, f"{pre}_uni")(got, dstride, src, sstride, w, h, fx, fy, None, bd)
                want = base.copy()
                L.orc_hevc_mc_fused(ss, int(luma), vp(want), dstride, vp(src), sstride, w, h, fx, fy,
                                    i16(other) if bi else None, bd)
                assert np.array_equal(got, want), (w, h, fx, fy, bi)


# -------------------------------------------------------------------- SAO

@pytest.mark.parametrize("bd", BDS)
def test_sao(dsp, bd):
    r = np.random.default_rng(61 + bd)
    d = dsp[bd]
    ss = 1 if bd == 8 else 2
    mx = (1 << bd) - 1
    for (w, h) in [(64, 64), (32, 32), (16, 8), (5, 3)]:
        stride = w + 6
        src = samples(r, stride * (h + 2), bd)
        base = samples(r, stride * (h + 2), bd)
        for trial in range(6):
            table = np.zeros(32, np.int16)
            pos = int(r.integers(0, 32))
            for k in range(4):
                table[(pos + k) & 31] = int(r.integers(-7, 8)) << (bd - min(bd, 10))
            got, want = base.copy(), base.copy()
            d.sao_band(got, stride, src, stride, w, h, table, bd - 5, mx)
            L.orc_hevc_sao_band(ss, vp(want), stride, vp(src), stride, w, h, i16(table), bd - 5, mx)
            assert np.array_equal(got, want), ("band", w, h, trial)
        for na, nb in [(-1, 1), (-stride, stride), (-stride - 1, stride + 1), (-stride + 1, stride - 1)]:
            off = np.array([int(r.integers(0, 8)), int(r.integers(0, 8)), 0, -int(r.integers(0, 8)),
                            -int(r.integers(0, 8))], np.int16)
            origin = stride + 1
            hh, ww = h, min(w, stride - 2)
            got, want = base.copy(), base.copy()
            d.sao_edge(got, src, origin, stride, ww, hh, na, nb, off, mx)
            L.orc_hevc_sao_edge(ss, vp(want), vp(src), origin, stride, ww, hh, na, nb, i16(off), mx)
            assert np.array_equal(got, want), ("edge", w, h, na, nb)


# ------------------------------------------------------------- deblocking

def _edge_block(r, bd, vertical, luma, smooth):
    """A 16 x 16 block around an edge at column / row 8, smooth enough for
    the filters to fire (random steps on a ramp)."""
    n = 16
    base = r.integers(0, (1 << bd) - 64)
    blk = np.full((n, n), base, np.int64)
    ramp = np.arange(n) * int(r.integers(0, 3))
    blk += ramp[None, :] if vertical else ramp[:, None]
    noise = r.integers(0, smooth + 1, (n, n))
    step = int(r.integers(0, 24)) << max(bd - 8, 0)
    if vertical:
        blk[:, 8:] += step
    else:
        blk[8:, :] += step
    blk = np.clip(blk + noise, 0, (1 << bd) - 1)
    return blk.astype(np.uint8 if bd == 8 else np.uint16)


@pytest.mark.parametrize("bd", BDS)
@pytest.mark.parametrize("vertical", [True, False])
def test_hevc_deblock(dsp, bd, vertical):
    r = np.random.default_rng(71 + bd + 3 * vertical)
    d = dsp[bd]
    ss = 1 if bd == 8 else 2
    mx = (1 << bd) - 1
    sh = bd - 8
    name = "v" if vertical else "h"
    off = 8 if vertical else 8 * 16  # q0 of line 0 at (y 0, x 8) / (y 8, x 0)
    for trial in range(60):
        blk = _edge_block(r, bd, vertical, True, [0, 2, 6][trial % 3])
        beta = [int(r.integers(0, 65)) << sh, int(r.integers(0, 65)) << sh]
        tc = [int(r.integers(0, 25)) << sh, int(r.integers(0, 25)) << sh]
        if trial % 7 == 0:
            beta[1], tc[1] = 0, 0  # bS 0 segment
        no_p = [int(r.random() < 0.2), int(r.random() < 0.2)]
        no_q = [int(r.random() < 0.2), int(r.random() < 0.2)]
        got = blk.reshape(-1).copy()
        getattr(d, f"deblock_luma_{name}")(got, off, 16, beta, tc, no_p, no_q, mx)
        want = blk.reshape(-1).copy()
        L.orc_hevc_deblock_luma(ss, int(vertical), vp(want), off, 16, (C.c_int32 * 2)(*beta),
                                (C.c_int32 * 2)(*tc), (C.c_uint8 * 2)(*no_p), (C.c_uint8 * 2)(*no_q), mx)
        assert np.array_equal(got, want), ("luma", trial)
        tcc = [int(r.integers(0, 12)) << sh for _ in range(4)]
        if trial % 5 == 0:
            tcc[2] = 0
        no_p4 = [int(r.random() < 0.2) for _ in range(4)]
        no_q4 = [int(r.random() < 0.2) for _ in range(4)]
        got = blk.reshape(-1).copy()
        getattr(d, f"deblock_chroma_{name}")(got, off, 16, tcc, no_p4, no_q4, mx)
        want = blk.reshape(-1).copy()
        L.orc_hevc_deblock_chroma(ss, int(vertical), vp(want), off, 16, (C.c_int32 * 4)(*tcc),
                                  (C.c_uint8 * 4)(*no_p4), (C.c_uint8 * 4)(*no_q4), mx)
        assert np.array_equal(got, want), ("chroma", trial)


# ------------------------------------------------------------------ H.264

def test_h264_qpel(dsp):
    r = np.random.default_rng(81)
    d = dsp["h264"]
    for (w, h) in [(4, 4), (8, 4), (4, 8), (8, 8), (16, 8), (8, 16), (16, 16)]:
        stride = w + 5 + 3
        src = r.integers(0, 256, stride * (h + 5)).astype(np.uint8)
        for pos in range(16):
            got = np.zeros(16 * 16, np.uint8)
            d.qpel[pos](got, src, stride, w, h, 255)
            want = np.zeros(16 * 16, np.uint8)
            L.orc_h264_qpel(pos, want.ctypes.data_as(U8), src.ctypes.data_as(U8), stride, w, h, 255)
            assert np.array_equal(got, want), (w, h, pos)


def test_h264_chroma_and_combiners(dsp):
    r = np.random.default_rng(82)
    d = dsp["h264"]
    u8 = lambda a: a.ctypes.data_as(U8)  # noqa: E731
    for (w, h) in [(2, 2), (4, 2), (2, 4), (4, 4), (8, 8), (16, 16)]:
        stride = w + 4
        src = r.integers(0, 256, stride * (h + 1)).astype(np.uint8)
        for xf in range(8):
            for yf in range(8):
                got = np.zeros(16 * 16, np.uint8)
                d.chroma(got, src, stride, w, h, xf, yf)
                want = np.zeros(16 * 16, np.uint8)
                L.orc_h264_chroma(u8(want), u8(src), stride, w, h, xf, yf)
                assert np.array_equal(got, want), (w, h, xf, yf)
        a = r.integers(0, 256, 256).astype(np.uint8)
        b = r.integers(0, 256, 256).astype(np.uint8)
        base = r.integers(0, 256, stride * h).astype(np.uint8)
        got, want = base.copy(), base.copy()
        d.copy(got, stride, a, w, h)
        L.orc_h264_copy(u8(want), stride, u8(a), w, h)
        assert np.array_equal(got, want)
        got, want = base.copy(), base.copy()
        d.avg(got, stride, a, b, w, h)
        L.orc_h264_avg(u8(want), stride, u8(a), u8(b), w, h)
        assert np.array_equal(got, want)
        for log_wd in (0, 1, 5, 7):
            w0, w1 = int(r.integers(-128, 128)), int(r.integers(-128, 128))
            o0, o1 = int(r.integers(-128, 128)), int(r.integers(-128, 128))
            got, want = base.copy(), base.copy()
            d.weighted_uni(got, stride, a, w, h, log_wd, w0, o0, 255)
            L.orc_h264_weighted_uni(u8(want), stride, u8(a), w, h, log_wd, w0, o0, 255)
            assert np.array_equal(got, want), ("wuni", log_wd)
            got, want = base.copy(), base.copy()
            d.weighted_bi(got, stride, a, b, w, h, log_wd, w0, w1, o0, o1, 255)
            L.orc_h264_weighted_bi(u8(want), stride, u8(a), u8(b), w, h, log_wd, w0, w1, o0, o1, 255)
            assert np.array_equal(got, want), ("wbi", log_wd)


@pytest.mark.parametrize("vertical", [True, False])
@pytest.mark.parametrize("chroma", [False, True])
def test_h264_deblock(dsp, vertical, chroma):
    r = np.random.default_rng(91 + 2 * vertical + chroma)
    d = dsp["h264"]
    kind = ("chroma" if chroma else "luma") + ("_v" if vertical else "_h")
    off = 8 if vertical else 8 * 16
    for trial in range(60):
        blk = _edge_block(r, 8, vertical, not chroma, [0, 2, 6][trial % 3]).reshape(-1)
        alpha, beta = int(r.integers(0, 60)), int(r.integers(0, 18))
        tc0 = [int(r.integers(-1, 14)) for _ in range(4)]
        got, want = blk.copy(), blk.copy()
        getattr(d, f"deblock_{kind}")(got, off, 16, alpha, beta, tc0, 255)
        L.orc_h264_deblock(int(chroma), int(vertical), want.ctypes.data_as(U8), off, 16, alpha, beta,
                           (C.c_int16 * 4)(*tc0), 255)
        assert np.array_equal(got, want), (kind, trial, tc0)
        got, want = blk.copy(), blk.copy()
        getattr(d, f"deblock_{kind}_intra")(got, off, 16, alpha, beta, 255)
        L.orc_h264_deblock(int(chroma), int(vertical), want.ctypes.data_as(U8), off, 16, alpha, beta, None, 255)
        assert np.array_equal(got, want), (kind, "intra", trial)


def test_h264_transforms(dsp):
    r = np.random.default_rng(99)
    d = dsp["h264"]
    u8 = lambda a: a.ctypes.data_as(U8)  # noqa: E731
    for trial in range(80):
        stride = 11
        base = r.integers(0, 256, stride * 8).astype(np.uint8)
        big = trial % 10 == 9  # values that wrap in i16
        c4 = r.integers(-30000 if big else -600, 30000 if big else 600, 16).astype(np.int16)
        c8 = r.integers(-30000 if big else -600, 30000 if big else 600, 64).astype(np.int16)
        got, want = base.copy(), base.copy()
        d.idct4_add(got, stride, c4, 255)
        L.orc_h264_idct4_add(u8(want), stride, i16(c4), 255)
        assert np.array_equal(got, want), ("idct4", trial)
        got, want = base.copy(), base.copy()
        d.idct8_add(got, stride, c8, 255)
        L.orc_h264_idct8_add(u8(want), stride, i16(c8), 255)
        assert np.array_equal(got, want), ("idct8", trial)
        dc = int(r.integers(-40000, 40000))
        for n in (4, 8):
            got, want = base.copy(), base.copy()
            (d.idct4_dc_add if n == 4 else d.idct8_dc_add)(got, stride, dc, 255)
            L.orc_h264_dc_add(u8(want), stride, dc, n, 255)
            assert np.array_equal(got, want), ("dc", n, dc)
        # residual4 / residual8: sparse i32 levels (some out of i16 range),
        # with and without a replacing DC, and DC-only blocks.
        co4 = np.where(r.random(16) < 0.3, r.integers(-2000, 2000, 16), 0).astype(np.int32)
        co8 = np.where(r.random(64) < 0.2, r.integers(-2000, 2000, 64), 0).astype(np.int32)
        if trial % 4 == 0:
            co4[1:] = 0
            co8[1:] = 0
        if trial % 9 == 0:
            co4[3] = 70000
            co8[5] = -70000
        for dcv in (None, int(r.integers(-3000, 3000))):
            got, want = base.copy(), base.copy()
            from rivet_hip.dsp import NO_DC
            d.residual4(got, stride, co4, NO_DC if dcv is None else dcv, 255)
            L.orc_h264_residual4(u8(want), stride, co4.ctypes.data_as(I32), 0 if dcv is None else dcv,
                                 0 if dcv is None else 1, 255)
            assert np.array_equal(got, want), ("res4", trial, dcv)
        got, want = base.copy(), base.copy()
        d.residual8(got, stride, co8, 255)
        L.orc_h264_residual8(u8(want), stride, co8.ctypes.data_as(I32), 255)
        assert np.array_equal(got, want), ("res8", trial)


# ------------------------------------------------------------ error paths

def test_out_of_range_calls_change_nothing(dsp):
    """Calls the reference would panic on are refused before any copy."""
    from rivet_hip import RivetError

    d = dsp[8]
    src = np.zeros(84, np.uint8)
    dst = np.full(64, 7, np.int16)
    with pytest.raises(RivetError) as e:
        d.qpel_h(dst, src, 10, 8, 8, 1, 0)  # the window's last index is 7 * 10 + 14 = 84
    assert e.value.code == -5
    assert (dst == 7).all()
    with pytest.raises(RivetError):
        d.idct[2](np.zeros(255, np.int16), 12, 0, 0)  # 16x16 needs 256
    with pytest.raises(RivetError):
        d.deblock_luma_v(np.zeros(64, np.uint8), 2, 8, [10, 10], [2, 2], [0, 0], [0, 0], 255)  # p3 at -2
