"""Kernel variants and the autotuner, through the C ABI: every variant a
batch allows reconstructs the same bytes as the oracle (plain batched launch
and graph replay), the tuner only picks variants the batch's sample depths
allow, and variants they rule out are refused."""
import numpy as np
import pytest

import oracle
import rivet_hip as rh
from rivet_hip import configs as cf

pytestmark = pytest.mark.gpu

RV_E_RANGE = -5


def _setup(specs):
    """specs: (cfg, seed) per picture; one session per picture, references in
    slots 0 and 1, output in slot 2."""
    items = []
    for cfg, seed in specs:
        bd = cfg["bit_depth"]
        refs = [rh.gen_planes(cfg["width"], cfg["height"], bd, cf.ref_seed(seed, k)) for k in range(2)]
        g = rh.GenPicture(rh.RV_CODEC_HEVC, cfg, seed)
        s = rh.Session(rh.RV_CODEC_HEVC, cfg["width"], cfg["height"], bd, 3)
        for k in range(2):
            s.upload(k, refs[k])
        j = s.prepare(g, 2)
        want = oracle.reconstruct(0, g.pic, refs + [None], 0)
        items.append((g, s, j, want))
    return items


def _check(batch, items, mode, label):
    batch.launch(mode)
    rh.sync()
    for k, (_, s, _, want) in enumerate(items):
        got = s.download(2)
        for c in range(3):
            assert np.array_equal(got[c], want[c]), f"{label}: picture {k} plane {c} differs"


def _close(batch, items):
    batch.close()
    for _, s, j, _ in items:
        j.close()
        s.close()


def _depths():
    """depth class of each variant per stage: 0 any, 1 8-bit only, 2 deeper only."""
    names = rh.variant_names()
    return [[0 if n.startswith("any") else 1 if n.startswith("8-bit") else 2 for n in stage] for stage in names]


def _sweep(batch, items, cls, label):
    for st, deps in enumerate(_depths()):
        for v, d in enumerate(deps):
            if d not in (0, cls):
                with pytest.raises(rh.RivetError) as e:
                    batch.set_variant(st, v)
                assert e.value.code == RV_E_RANGE
                continue
            batch.set_variant(st, v)
            _check(batch, items, rh.RV_LAUNCH_BATCHED, f"{label} stage {st} variant {v}")
        batch.set_variant(st, 0)


def test_every_variant_matches_the_oracle_8bit(device):
    specs = [(cf.HEVC_C1, rh.subseed(cf.BASE_SEED, 9, k, 1)) for k in range(2)]
    specs.append((dict(cf.HEVC_C2, width=416, height=240, bit_depth=8), 4242))
    items = _setup(specs)
    b = rh.Batch([j for _, _, j, _ in items])
    _sweep(b, items, 1, "8-bit batch")
    chosen = b.tune(2)
    deps = _depths()
    assert all(deps[s][v] in (0, 1) for s, v in enumerate(chosen)), chosen
    for rep in range(3):  # graph capture and replays with the tuned variants
        _check(b, items, rh.RV_LAUNCH_GRAPH, f"tuned graph replay {rep}")
    _close(b, items)


def test_every_variant_matches_the_oracle_main10(device):
    cfg = dict(cf.HEVC_C2, width=416, height=240)
    items = _setup([(cfg, 777), (cfg, 778)])
    b = rh.Batch([j for _, _, j, _ in items])
    _sweep(b, items, 2, "Main10 batch")
    chosen = b.tune(2)
    deps = _depths()
    assert all(deps[s][v] in (0, 2) for s, v in enumerate(chosen)), chosen
    _check(b, items, rh.RV_LAUNCH_GRAPH, "tuned Main10 graph")
    _close(b, items)


def test_mixed_depth_batch_keeps_the_generic_variants(device):
    items = _setup([(cf.HEVC_C1, 31337), (dict(cf.HEVC_C2, width=416, height=240), 31338)])
    b = rh.Batch([j for _, _, j, _ in items])
    _sweep(b, items, 0, "mixed batch")
    # The tuner may pick any variant that serves every depth (a timing
    # choice); none specialised to one depth.
    names = rh.variant_names()
    chosen = b.tune(1)
    assert all(names[s][v].startswith("any") for s, v in enumerate(chosen)), chosen
    _check(b, items, rh.RV_LAUNCH_GRAPH, "mixed graph")
    _close(b, items)
