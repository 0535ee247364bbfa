"""On-device scaling of levels: HEVC scale_coefficients and H.264 Dequant
(with the Intra_16x16 luma and chroma DC transforms) inside stage 1,
against the oracle byte for byte, with flat and seeded scaling lists, at
8, 10 and 12 bits; and the work-list checks that refuse malformed levels.
Run through the C ABI (rivet_hip)."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import _abi
from rivet_hip import configs as cf

pytestmark = pytest.mark.gpu

RV_E_RANGE = -5


def _run(codec, cfg, seed, mask, refs, pic=None):
    bd = cfg.get("bit_depth", 8)
    g = rh.GenPicture(codec, cfg, seed)
    s = rh.Session(codec, cfg["width"], cfg["height"], bd, 3)
    for k in range(2):
        s.upload(k, refs[k])
    rh.set_stage_mask(mask)
    try:
        j = s.prepare(g if pic is None else pic, 2)
    finally:
        rh.set_stage_mask(0)
    b = rh.Batch([j])
    b.launch(rh.RV_LAUNCH_PLAIN)
    rh.sync()
    out = s.download(2)
    b.close()
    j.close()
    s.close()
    return g, out


def _check(codec, cfg, seed, label):
    bd = cfg.get("bit_depth", 8)
    print(f"levels parity: {label} {cfg['width']}x{cfg['height']} seed {seed:#x}", flush=True)
    refs = [rh.gen_planes(cfg["width"], cfg["height"], bd, cf.ref_seed(seed, k)) for k in range(2)]
    for mask, stop in ((3, 2), (0, 0)):
        g, got = _run(codec, cfg, seed, mask, refs)
        if codec == rh.RV_CODEC_HEVC:
            c = g.pic.contents
            assert any(c.tus[i].flags & _abi.RV_TU_LEVELS for i in range(c.n_tus))
        want = oracle.reconstruct(codec, g.pic, refs + [None], stop)
        for k in range(3):
            if not np.array_equal(got[k], want[k]):
                bad = np.argwhere(got[k] != want[k])
                y, x = bad[0]
                raise AssertionError(f"{label} stop={stop} plane {k}: {len(bad)} samples differ, first at "
                                     f"(x={x}, y={y}): got {got[k][y, x]} want {want[k][y, x]}")


HEVC_SMALL = dict(cf.HEVC_C1, intra_percent=30, levels=1)


@pytest.mark.parametrize("scaling", [0, 1])
@pytest.mark.parametrize("bd", [8, 10, 12])
def test_hevc_levels_416x240(device, bd, scaling):
    cfg = dict(HEVC_SMALL, bit_depth=bd, scaling=scaling, explicit_wp=1 if bd == 10 else 0)
    _check(rh.RV_CODEC_HEVC, cfg, rh.subseed(cf.BASE_SEED, 21, bd, scaling), f"HEVC {bd}-bit lists={scaling}")


def test_hevc_levels_main10_b_1080p(device):
    _, cfg, seed = cf.single(2)
    _check(rh.RV_CODEC_HEVC, dict(cfg, levels=1, scaling=1), seed, "HEVC config 2 with levels")


@pytest.mark.parametrize("scaling", [0, 1])
def test_h264_levels_352x288(device, scaling):
    cfg = dict(cf.H264_C4, width=352, height=288, intra_percent=30, levels=1, scaling=scaling)
    _check(rh.RV_CODEC_H264, cfg, rh.subseed(cf.BASE_SEED, 22, scaling, 0), f"H.264 lists={scaling}")


def test_h264_levels_1080p(device):
    _, cfg, seed = cf.single(4)
    _check(rh.RV_CODEC_H264, dict(cfg, levels=1, scaling=1), seed, "H.264 config 4 with levels")


# ------------------------------------------------------------ refusals

def _prepare_rc(codec, w, h, bd, pic, words=""):
    """The code prepare fails with; its message must contain `words`."""
    s = rh.Session(codec, w, h, bd, 3)
    try:
        with pytest.raises(rh.RivetError) as e:
            s.prepare(pic, 2)
        assert words in str(e.value), str(e.value)
        return e.value.code
    finally:
        s.close()


def _hevc_with(edit):
    cfg = dict(HEVC_SMALL, width=64, height=64)
    g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, 5)
    c = g.pic.contents
    tus = (_abi.HevcTu * c.n_tus)(*[_abi.HevcTu.from_buffer_copy(c.tus[i]) for i in range(c.n_tus)])
    k = next(i for i in range(c.n_tus) if tus[i].flags & _abi.RV_TU_LEVELS)
    edit(tus[k])
    p = _abi.HevcPic.from_buffer_copy(c)
    p.tus = C.cast(tus, C.POINTER(_abi.HevcTu))
    return g, tus, C.pointer(p)


def test_hevc_refuses_qp_beyond_the_depth(device):
    def edit(t):
        t.qp = 52
    keep = _hevc_with(edit)
    assert _prepare_rc(rh.RV_CODEC_HEVC, 64, 64, 8, keep[2], "qP") == RV_E_RANGE


def test_hevc_refuses_unknown_flag(device):
    def edit(t):
        t.flags |= 16
    keep = _hevc_with(edit)
    assert _prepare_rc(rh.RV_CODEC_HEVC, 64, 64, 8, keep[2], "flag") == RV_E_RANGE


def test_hevc_refuses_dst_levels_outside_the_bounds(device):
    """A scaled DST TU may carry coefficients outside its bounds (the DST
    transforms the whole 4x4); levels may not, since scale_coefficients
    leaves them unscaled."""
    cfg = dict(HEVC_SMALL, width=64, height=64, intra_percent=100)
    g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, 3)
    c = g.pic.contents
    tus = (_abi.HevcTu * c.n_tus)(*[_abi.HevcTu.from_buffer_copy(c.tus[i]) for i in range(c.n_tus)])
    k = next(i for i in range(c.n_tus) if (tus[i].flags & _abi.RV_TU_DST) and (tus[i].flags & _abi.RV_TU_CODED))
    tus[k].max_x = tus[k].max_y = 0
    nz = (_abi.Nz * c.n_nz)(*[_abi.Nz(c.nz[i].pos, c.nz[i].val) for i in range(c.n_nz)])
    nz[tus[k].coef_off].pos = 15
    p = _abi.HevcPic.from_buffer_copy(c)
    p.tus = C.cast(tus, C.POINTER(_abi.HevcTu))
    p.nz = C.cast(nz, C.POINTER(_abi.Nz))
    assert _prepare_rc(rh.RV_CODEC_HEVC, 64, 64, 8, C.pointer(p), "bounds") == RV_E_RANGE


def _h264_with(edit, intra_percent=30):
    cfg = dict(cf.H264_C4, width=64, height=64, intra_percent=intra_percent, levels=1)
    g = rh.GenPicture(rh.RV_CODEC_H264, cfg, 11)
    c = g.pic.contents
    blks = (_abi.H264Blk * c.n_blks)(*[_abi.H264Blk.from_buffer_copy(c.blks[i]) for i in range(c.n_blks)])
    mbs = (_abi.H264Mb * c.n_mbs)(*[_abi.H264Mb.from_buffer_copy(c.mbs[i]) for i in range(c.n_mbs)])
    edit(c, blks, mbs)
    p = _abi.H264Pic.from_buffer_copy(c)
    p.blks = C.cast(blks, C.POINTER(_abi.H264Blk))
    p.mbs = C.cast(mbs, C.POINTER(_abi.H264Mb))
    return g, blks, mbs, C.pointer(p)


def test_h264_refuses_dc_levels_outside_coefs(device):
    def edit(c, blks, mbs):
        b = next(b for b in blks if b.has_dc)
        b.dc = c.n_coefs - 2
    keep = _h264_with(edit)
    assert _prepare_rc(rh.RV_CODEC_H264, 64, 64, 8, keep[3], "DC levels") == RV_E_RANGE


def test_h264_refuses_luma_dc_in_an_inter_macroblock(device):
    """An Intra_16x16 luma block moved onto an inter macroblock."""
    def edit(c, blks, mbs):
        mbw = c.width // 16
        b = next(b for b in blks if b.has_dc and b.comp == 0)
        a = next(a for a in range(c.n_mbs) if not mbs[a].intra)
        b.x = (a % mbw) * 16 + b.x % 16
        b.y = (a // mbw) * 16 + b.y % 16
    keep = _h264_with(edit, intra_percent=50)
    assert _prepare_rc(rh.RV_CODEC_H264, 64, 64, 8, keep[3], "intra") == RV_E_RANGE


def test_h264_refuses_dc_on_8x8_and_unknown_flags(device):
    def dc8(c, blks, mbs):
        b = next(b for b in blks if b.size == 8)
        b.has_dc = 1
    keep = _h264_with(dc8)
    assert _prepare_rc(rh.RV_CODEC_H264, 64, 64, 8, keep[3], "8x8") == RV_E_RANGE

    def flag(c, blks, mbs):
        blks[0].flags |= 2
    keep = _h264_with(flag)
    assert _prepare_rc(rh.RV_CODEC_H264, 64, 64, 8, keep[3], "flag") == RV_E_RANGE
