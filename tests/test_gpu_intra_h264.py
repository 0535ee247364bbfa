"""H.264 intra prediction on the device (RV_INTRA_DEVICE macroblocks):
pictures mixing inter and device-intra macroblocks of every kind
(Intra_4x4, Intra_8x8, Intra_16x16), all-intra pictures whose
dependency levels run the width of the picture, one-row and one-column
pictures, constrained intra prediction, levels with scaling lists, plain,
batched and graph launches over pictures with different level counts —
against the oracle byte for byte, deblocked; and the work-list checks that
refuse modes without their neighbours. Run through the C ABI (rivet_hip)."""
import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import _abi
from rivet_hip import configs as cf

pytestmark = pytest.mark.gpu

RV_E_INVALID, RV_E_RANGE = -1, -5


def _pictures(specs):
    out = []
    for cfg, seed in specs:
        g = rh.GenPicture(rh.RV_CODEC_H264, cfg, seed)
        refs = [rh.gen_planes(cfg["width"], cfg["height"], 8, cf.ref_seed(seed, k)) for k in range(2)]
        out.append((cfg, seed, g, refs))
    return out


def _compare(label, got, want):
    for k in range(3):
        if not np.array_equal(got[k], want[k]):
            bad = np.argwhere(got[k] != want[k])
            y, x = bad[0]
            raise AssertionError(f"{label} plane {k}: {len(bad)} samples differ, first at (x={x}, y={y}): "
                                 f"got {got[k][y, x]} want {want[k][y, x]}")


def _run(specs, mode, label):
    pics = _pictures(specs)
    for cfg, seed, g, _ in pics:
        c = g.pic.contents
        n_dev = sum(c.mbs[i].intra == _abi.RV_INTRA_DEVICE for i in range(c.n_mbs))
        print(f"h264 device intra: {label} {cfg['width']}x{cfg['height']} seed {seed:#x}, {n_dev} device-intra "
              f"macroblocks, launch {mode}", flush=True)
        assert n_dev > 0 or cfg.get("intra_percent", 0) == 0 or not cfg.get("intra_device")
    sessions, jobs = [], []
    for cfg, seed, g, refs in pics:
        s = rh.Session(rh.RV_CODEC_H264, cfg["width"], cfg["height"], 8, 3)
        for k in range(2):
            s.upload(k, refs[k])
        sessions.append(s)
        jobs.append(s.prepare(g, 2))
    b = rh.Batch(jobs)
    for _ in range(2 if mode == rh.RV_LAUNCH_GRAPH else 1):
        b.launch(mode)
    rh.sync()
    for (cfg, seed, g, refs), s in zip(pics, sessions):
        want = oracle.reconstruct(rh.RV_CODEC_H264, g.pic, refs + [None], 0)
        _compare(f"{label} seed {seed:#x}", s.download(2), want)
    b.close()
    for j in jobs:
        j.close()
    for s in sessions:
        s.close()


MIXED = dict(cf.H264_C4, width=352, height=288, intra_percent=30, intra_device=1)


@pytest.mark.parametrize("cip", [0, 1])
@pytest.mark.parametrize("levels", [0, 1])
def test_mixed_inter_and_device_intra(device, cip, levels):
    cfg = dict(MIXED, constrained_intra_pred=cip, levels=levels, scaling=levels)
    _run([(cfg, rh.subseed(cf.BASE_SEED, 41, cip, levels))], rh.RV_LAUNCH_PLAIN, f"mixed cip={cip} levels={levels}")


def test_all_intra_picture(device):
    # Every macroblock device-intra: level x + 2y, the whole wavefront.
    cfg = dict(MIXED, width=256, height=128, intra_percent=100, b_picture=0, weighting=0, poc=1, ref_poc=[0, 0])
    _run([(cfg, rh.subseed(cf.BASE_SEED, 42, 0, 0))], rh.RV_LAUNCH_PLAIN, "all intra")


@pytest.mark.parametrize("w,h", [(16, 16), (16, 96), (160, 16)])
def test_single_row_and_column_pictures(device, w, h):
    cfg = dict(MIXED, width=w, height=h, intra_percent=100, constrained_intra_pred=0)
    _run([(cfg, rh.subseed(cf.BASE_SEED, 43, w, h))], rh.RV_LAUNCH_PLAIN, "edge shapes")


@pytest.mark.parametrize("mode", [rh.RV_LAUNCH_BATCHED, rh.RV_LAUNCH_GRAPH])
def test_batch_of_pictures_with_different_levels(device, mode):
    specs = [(dict(MIXED, intra_percent=60, constrained_intra_pred=1), rh.subseed(cf.BASE_SEED, 44, 0, 0)),
             (dict(MIXED, intra_percent=0), rh.subseed(cf.BASE_SEED, 44, 1, 0)),
             (dict(MIXED, width=176, height=144, intra_percent=100, levels=1), rh.subseed(cf.BASE_SEED, 44, 2, 0)),
             (dict(MIXED, intra_percent=10, intra_device=0), rh.subseed(cf.BASE_SEED, 44, 3, 0))]
    _run(specs, mode, "batch")


def test_config4_1080p_with_device_intra_batched(device):
    _, cfg, seed = cf.single(4)
    _run([(dict(cfg, intra_device=1), seed), (dict(cfg, intra_device=1, constrained_intra_pred=1), seed + 1)],
         rh.RV_LAUNCH_BATCHED, "config 4 with device intra")


# ------------------------------------------------------------ refusals

def _prepare_rc(g, cfg, words):
    s = rh.Session(rh.RV_CODEC_H264, cfg["width"], cfg["height"], 8, 3)
    try:
        with pytest.raises(rh.RivetError) as e:
            s.prepare(g, 2)
        assert words in str(e.value), str(e.value)
        return e.value.code
    finally:
        s.close()


def test_modes_without_their_neighbours_are_refused(device):
    cfg = dict(MIXED, width=48, height=48, intra_percent=100, constrained_intra_pred=0)
    g = rh.GenPicture(rh.RV_CODEC_H264, cfg, 7)
    m = g.pic.contents.mbs[0]  # top-left macroblock: no neighbours
    m.intra_kind, m.transform_8x8, m.chroma_mode = _abi.RV_H264_I16X16, 0, 0
    m.modes[0] = 0  # vertical without the row above
    assert _prepare_rc(g, cfg, "Intra_16x16") == RV_E_INVALID
    m.modes[0], m.chroma_mode = 2, 3  # chroma plane without neighbours
    assert _prepare_rc(g, cfg, "chroma") == RV_E_INVALID
    m.chroma_mode = 0
    m.intra_kind, m.transform_8x8 = _abi.RV_H264_I4X4, 1  # 8x8 transform on Intra_4x4
    assert _prepare_rc(g, cfg, "transform_8x8") == RV_E_RANGE
