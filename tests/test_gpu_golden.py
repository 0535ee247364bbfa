"""The device reconstruction of every golden case equals the committed
output hashes (tests/golden/reconstruction.json), plane by plane, through
the batched picture API. This is independent of running the oracle here."""
import pytest

from tests.golden import cases

pytestmark = pytest.mark.gpu

GOLD = cases.load()


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_device_matches_golden(device, name):
    rh = device
    codec, cfg, seed = cases.CASES[name]
    g = rh.GenPicture(codec, cfg, seed)
    assert cases.worklist_digest(codec, g.pic) == GOLD[name]["worklist_sha256"]
    refs = cases.references(codec, cfg, seed)
    s = rh.Session(codec, cfg["width"], cfg["height"], cfg.get("bit_depth", 8), 3)
    for k in range(2):
        s.upload(k, refs[k])
    j = s.prepare(g, 2)
    b = rh.Batch([j])
    b.launch(rh.RV_LAUNCH_BATCHED)
    rh.sync()
    out = s.download(2)
    b.close()
    j.close()
    s.close()
    assert cases.planes_digest(out) == GOLD[name]["output_sha256"]
