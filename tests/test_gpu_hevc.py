"""HEVC picture parity: the HIP path against the oracle, byte for byte.

Each picture is checked at three cut points — after prediction + residual
(deblocking and SAO switched off), after deblocking (SAO off) and whole —
so a mismatch names the stage. Run through the C ABI (rivet_hip).
"""
import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import configs as cf

pytestmark = pytest.mark.gpu


def _run_gpu(rh, cfg, seed, mask, refs):
    g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, seed)
    s = rh.Session(rh.RV_CODEC_HEVC, cfg["width"], cfg["height"], cfg["bit_depth"], 3)
    for k in range(2):
        s.upload(k, refs[k])
    rh.set_stage_mask(mask)
    try:
        j = s.prepare(g, 2)
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


def _check(rh, cfg, seed, label):
    bd = cfg["bit_depth"]
    refs = [rh.gen_planes(cfg["width"], cfg["height"], bd, cf.ref_seed(seed, k)) for k in range(2)]
    for mask, stop in ((3, 2), (2, 3), (0, 0)):
        g, got = _run_gpu(rh, cfg, seed, mask, refs)
        want = oracle.reconstruct(0, g.pic, refs + [None], stop)
        for c in range(3):
            if not np.array_equal(got[c], want[c]):
                bad = np.argwhere(got[c] != want[c])
                y, x = bad[0]
                raise AssertionError(
                    f"{label} stop={stop} plane {c}: {len(bad)} samples differ, first at (x={x}, y={y}): "
                    f"got {got[c][y, x]} want {want[c][y, x]}")


@pytest.mark.parametrize("seed_k", [0, 1, 2])
def test_config1_p_416x240(device, seed_k):
    _check(device, cf.HEVC_C1, rh.subseed(cf.BASE_SEED, 1, seed_k, 1), f"C1 seed {seed_k}")


def test_config2_main10_b_1080p(device):
    _, cfg, seed = cf.single(2)
    _check(device, cfg, seed, "C2")


@pytest.mark.parametrize("stream", [0, 1])
def test_config3_stream_4k(device, stream):
    cfg, seed = cf.c3_stream(stream)
    _check(device, cfg, seed, f"C3 stream {stream}")


def test_small_b_with_intra_8bit(device):
    cfg = dict(cf.HEVC_C2, width=416, height=240, bit_depth=8)
    _check(device, cfg, 12345, "8-bit B WP intra")
