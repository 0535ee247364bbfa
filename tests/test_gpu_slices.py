"""Slices, tiles and loop-filter exemptions on the device against the
oracle, bit-exact: per-slice deblocking switches and offsets, edges and SAO
neighbours across slice and tile boundaries as their flags allow
(src/hevc/deblock.rs:105-201, src/hevc/sao.rs:193-238), and samples of
PCM / bypass / palette CUs left alone (src/hevc/sao.rs:140-151,
src/hevc/deblock.rs:274-275); sliced and one-slice pictures share one
batch, and the argument checks of the per-CTB tables hold."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import configs as cf

pytestmark = pytest.mark.gpu

BASE = dict(cf.HEVC_C1, intra_percent=20, nz_permille=200, qp_min=30, qp_max=45)

CASES = [
    ("tiles+slices, no crossing", dict(n_slices=4, tile_cols=3, tile_rows=2, lf_across_slices=0, lf_across_tiles=0)),
    ("slices only, seeded flags", dict(n_slices=6, lf_across_slices=-1, slice_vary=1)),
    ("tiles crossing, slices not", dict(n_slices=3, tile_cols=2, tile_rows=2, lf_across_slices=0, lf_across_tiles=1)),
    ("everything seeded + exempt", dict(n_slices=5, tile_cols=3, tile_rows=2, lf_across_slices=-1,
                                        lf_across_tiles=-1, slice_vary=1, exempt_percent=25)),
    ("one slice, exempt CUs only", dict(exempt_percent=40)),
]


def _check(cfg, seed):
    bd = cfg.get("bit_depth", 8)
    refs = [rh.gen_planes(cfg["width"], cfg["height"], bd, cf.ref_seed(seed, k)) for k in range(2)]
    g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, seed)
    s = rh.Session(rh.RV_CODEC_HEVC, cfg["width"], cfg["height"], bd, 3)
    for k in range(2):
        s.upload(k, refs[k])
    j = s.prepare(g, 2)
    b = rh.Batch([j])
    b.launch(rh.RV_LAUNCH_BATCHED)
    rh.sync()
    got = s.download(2)
    want = oracle.reconstruct(rh.RV_CODEC_HEVC, g.pic, refs + [None], 0)
    for c in range(3):
        assert (got[c] == want[c]).all(), f"plane {c}: {int((got[c] != want[c]).sum())} samples differ"
    b.close()
    j.close()
    s.close()


@pytest.mark.parametrize("bd", [8, 10])
@pytest.mark.parametrize("name,extra", CASES, ids=[c[0] for c in CASES])
def test_sliced_picture_matches_oracle(device, name, extra, bd):
    seed = rh.subseed(0x51CE, bd, CASES.index((name, extra)))
    print(f"sliced picture: {name}, {bd}-bit, seed {seed}", flush=True)
    _check(dict(BASE, bit_depth=bd, **extra), seed)


def test_sliced_1080p_b_picture(device):
    cfg = dict(cf.HEVC_C2, n_slices=8, tile_cols=4, tile_rows=3, lf_across_slices=-1, lf_across_tiles=-1,
               slice_vary=1, exempt_percent=5)
    print("sliced 1920x1080 Main10 B-picture", flush=True)
    _check(cfg, 0x51CF)


def test_sliced_and_plain_pictures_in_one_batch(device):
    cfgs = [dict(BASE, n_slices=4, tile_cols=2, tile_rows=2, lf_across_slices=0, lf_across_tiles=0, slice_vary=1,
                 exempt_percent=20), dict(BASE), dict(BASE, bit_depth=10, exempt_percent=30)]
    print("sliced and plain pictures in one batch", flush=True)
    sessions, jobs, want, keep = [], [], [], []
    for k, cfg in enumerate(cfgs):
        bd = cfg.get("bit_depth", 8)
        seed = rh.subseed(0x51D0, k)
        refs = [rh.gen_planes(cfg["width"], cfg["height"], bd, cf.ref_seed(seed, r)) for r in range(2)]
        g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, seed)
        s = rh.Session(rh.RV_CODEC_HEVC, cfg["width"], cfg["height"], bd, 3)
        for r in range(2):
            s.upload(r, refs[r])
        jobs.append(s.prepare(g, 2))
        sessions.append(s)
        keep.append(g)
        want.append(oracle.reconstruct(rh.RV_CODEC_HEVC, g.pic, refs + [None], 0))
    b = rh.Batch(jobs)
    b.launch(rh.RV_LAUNCH_BATCHED)
    rh.sync()
    for s, w in zip(sessions, want):
        got = s.download(2)
        assert all((got[c] == w[c]).all() for c in range(3))
    b.close()
    for j in jobs:
        j.close()
    for s in sessions:
        s.close()


def test_ctb_tables_are_checked(device):
    g = rh.GenPicture(rh.RV_CODEC_HEVC, dict(BASE, n_slices=3), 7)
    s = rh.Session(rh.RV_CODEC_HEVC, BASE["width"], BASE["height"], 8, 3)
    p = g.pic.contents
    job = C.c_void_p()
    keep = p.ctb_slice[0]
    p.ctb_slice[0] = p.n_slices  # names no slice
    assert rh.lib.rv_hevc_prepare(s._h, g.pic, 2, C.byref(job)) == -5
    p.ctb_slice[0] = keep
    ts = C.cast(p.ctb_ts, C.c_void_p).value  # an address: the field's own object would see the reset
    p.ctb_ts = C.POINTER(C.c_uint32)()
    assert rh.lib.rv_hevc_prepare(s._h, g.pic, 2, C.byref(job)) == -1
    p.ctb_ts = C.cast(ts, C.POINTER(C.c_uint32))
    assert rh.lib.rv_hevc_prepare(s._h, g.pic, 2, C.byref(job)) == 0, rh.lib.rv_last_error()
    rh.lib.rv_job_free(job)
    s.close()
