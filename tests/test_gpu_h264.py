"""H.264 picture parity: the HIP path against the oracle, byte for byte,
after prediction + residual (deblocking off) and whole (deblocked by the
macroblock wavefront). The wavefront's own cases: one macroblock row or
column, rows past the 64 of one pass (slots reused), pictures wide enough
that rows start 3 or 5 steps apart, and non-zero alpha / beta offsets
with strong intra edges."""
import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import configs as cf

pytestmark = pytest.mark.gpu


def _check(rh, cfg, seed, label, offsets=None):
    w, h = cfg["width"], cfg["height"]
    refs = [rh.gen_planes(w, h, 8, cf.ref_seed(seed, k)) for k in range(2)]
    for mask, stop in ((1, 2), (0, 0)):
        g = rh.GenPicture(rh.RV_CODEC_H264, cfg, seed)
        if offsets:
            g.pic.contents.offset_a, g.pic.contents.offset_b = offsets
        s = rh.Session(rh.RV_CODEC_H264, w, h, 8, 3)
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
        got = s.download(2)
        want = oracle.reconstruct(1, g.pic, refs + [None], stop)
        for c in range(3):
            if not np.array_equal(got[c], want[c]):
                bad = np.argwhere(got[c] != want[c])
                y, x = bad[0]
                raise AssertionError(f"{label} stop={stop} plane {c}: {len(bad)} samples differ, first at "
                                     f"(x={x}, y={y}) got {got[c][y, x]} want {want[c][y, x]}")
        b.close()
        j.close()
        s.close()


def test_config4_b_implicit_1080p(device):
    _, cfg, seed = cf.single(4)
    _check(device, cfg, seed, "C4")


@pytest.mark.parametrize("kind", ["p", "b_default", "b_mixed"])
def test_small_pictures(device, kind):
    cfg = dict(cf.H264_C4, width=352, height=288)
    if kind == "p":
        cfg.update(b_picture=0, weighting=0, poc=1, ref_poc=[0, 0])
    elif kind == "b_default":
        cfg.update(weighting=0, bi_percent=50)
    else:
        cfg.update(bi_percent=50, intra_percent=30)
    _check(device, cfg, rh.subseed(77, {"p": 1, "b_default": 2, "b_mixed": 3}[kind]), kind)


@pytest.mark.parametrize("w,h", [(16, 16), (32, 16), (16, 48), (48, 32), (2032, 1072), (4096, 48)])
def test_deblock_wavefront_shapes(device, w, h):
    # 2032 x 1072: 127 x 67 macroblocks, rows 3 steps apart, 3 rows past the
    # first 64; 4096 x 48: 256 macroblocks wide, rows 5 steps apart.
    cfg = dict(cf.H264_C4, width=w, height=h, intra_percent=30, mv_range=min(512, 4 * min(w, h)))
    _check(device, cfg, rh.subseed(91, w, h), f"{w}x{h}")


def test_deblock_offsets_and_strong_edges(device):
    cfg = dict(cf.H264_C4, width=352, height=288, intra_percent=60, qp_min=24, qp_max=51, bi_percent=50)
    _check(device, cfg, rh.subseed(92, 1), "offsets", offsets=(6, -4))
    _check(device, cfg, rh.subseed(92, 2), "offsets-neg", offsets=(-6, 6))
