"""Intra prediction on the device (RV_INTRA_DEVICE CUs): pictures mixing
inter and device-intra CUs of every size, all-intra pictures, 8 and 10
bits, strong intra smoothing, constrained intra prediction, levels with
scaling lists — against the oracle byte for byte at two cut points; and
the work-list checks that refuse malformed device-intra CUs. Run through
the C ABI (rivet_hip)."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import _abi
from rivet_hip import configs as cf

pytestmark = pytest.mark.gpu

RV_E_RANGE = -5


def _check(cfg, seed, label, batched=False):
    bd = cfg.get("bit_depth", 8)
    print(f"device intra: {label} {cfg['width']}x{cfg['height']} seed {seed:#x}", flush=True)
    refs = [rh.gen_planes(cfg["width"], cfg["height"], bd, cf.ref_seed(seed, k)) for k in range(2)]
    for mask, stop in ((3, 2), (0, 0)):
        g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, seed)
        c = g.pic.contents
        assert any(c.cus[i].intra == _abi.RV_INTRA_DEVICE for i in range(c.n_cus))
        s = rh.Session(rh.RV_CODEC_HEVC, cfg["width"], cfg["height"], bd, 3)
        for k in range(2):
            s.upload(k, refs[k])
        rh.set_stage_mask(mask)
        try:
            j = s.prepare(g, 2)
        finally:
            rh.set_stage_mask(0)
        b = rh.Batch([j])
        b.launch(rh.RV_LAUNCH_BATCHED if batched else rh.RV_LAUNCH_PLAIN)
        rh.sync()
        got = s.download(2)
        b.close()
        j.close()
        s.close()
        want = oracle.reconstruct(0, g.pic, refs + [None], stop)
        for k in range(3):
            if not np.array_equal(got[k], want[k]):
                bad = np.argwhere(got[k] != want[k])
                y, x = bad[0]
                raise AssertionError(f"{label} stop={stop} plane {k}: {len(bad)} samples differ, first at "
                                     f"(x={x}, y={y}): got {got[k][y, x]} want {want[k][y, x]}")


MIXED = dict(cf.HEVC_C1, intra_percent=30, intra_device=1)


@pytest.mark.parametrize("cip", [0, 1])
def test_mixed_inter_and_device_intra_8bit(device, cip):
    _check(dict(MIXED, constrained_intra_pred=cip), rh.subseed(cf.BASE_SEED, 31, cip, 0), f"8-bit cip={cip}")


def test_main10_b_with_strong_smoothing_and_levels(device):
    cfg = dict(cf.HEVC_C2, width=416, height=240, intra_percent=25, intra_device=1, strong_intra_smoothing=1,
               levels=1, scaling=1)
    _check(cfg, rh.subseed(cf.BASE_SEED, 32, 0, 0), "10-bit strong, levels")


@pytest.mark.parametrize("bd", [8, 10])
def test_all_intra_picture(device, bd):
    cfg = dict(cf.HEVC_C1, width=192, height=128, bit_depth=bd, intra_percent=100, intra_device=1,
               strong_intra_smoothing=1)
    _check(cfg, rh.subseed(cf.BASE_SEED, 33, bd, 0), f"all intra {bd}-bit")


def test_device_intra_1080p_batched(device):
    _, cfg, seed = cf.single(2)
    _check(dict(cfg, intra_percent=15, intra_device=1), seed, "config 2 with device intra", batched=True)


# ------------------------------------------------------------ refusals

def _prepare_rc(pic, words):
    s = rh.Session(rh.RV_CODEC_HEVC, 64, 64, 8, 3)
    try:
        with pytest.raises(rh.RivetError) as e:
            s.prepare(pic, 2)
        assert words in str(e.value), str(e.value)
        return e.value.code
    finally:
        s.close()


def _edited(edit_cus=None, edit_tus=None):
    cfg = dict(cf.HEVC_C1, width=64, height=64, intra_percent=100, intra_device=1)
    g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, 7)
    c = g.pic.contents
    cus = (_abi.HevcCu * c.n_cus)(*[_abi.HevcCu.from_buffer_copy(c.cus[i]) for i in range(c.n_cus)])
    tus = [_abi.HevcTu.from_buffer_copy(c.tus[i]) for i in range(c.n_tus)]
    if edit_cus:
        edit_cus(cus)
    if edit_tus:
        tus = edit_tus(tus, cus)
    ta = (_abi.HevcTu * len(tus))(*tus)
    p = _abi.HevcPic.from_buffer_copy(c)
    p.cus = C.cast(cus, C.POINTER(_abi.HevcCu))
    p.tus = C.cast(ta, C.POINTER(_abi.HevcTu))
    p.n_tus = len(tus)
    return (g, cus, ta), C.pointer(p)


def test_refuses_bad_modes(device):
    def edit(cus):
        cus[0].luma_mode[0] = 35
    keep, pic = _edited(edit)
    assert _prepare_rc(pic, "mode") == RV_E_RANGE

    def edit2(cus):
        cus[0].part_mode = 1
    keep, pic = _edited(edit2)
    assert _prepare_rc(pic, "2Nx2N") == RV_E_RANGE


def test_refuses_a_missing_chroma_tu(device):
    def edit(tus, cus):
        k = next(i for i, t in enumerate(tus) if t.comp == 2)
        return tus[:k] + tus[k + 1:]
    keep, pic = _edited(edit_tus=edit)
    assert _prepare_rc(pic, "chroma") == RV_E_RANGE
