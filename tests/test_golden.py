"""The committed golden fixtures (tests/golden/reconstruction.json) are
reproduced on the CPU: the generator gives the same work lists and
reference pictures from the same seeds, and the oracle reconstructs the
same pictures. A deliberate change to either regenerates the file with
tests/golden/make_golden.py."""
import pytest

from tests.golden import cases

GOLD = cases.load()


def test_every_case_has_a_fixture():
    assert set(GOLD) == set(cases.CASES)


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_fixture_reproduces(name):
    got = cases.compute(name)
    want = GOLD[name]
    for k in ("codec", "seed", "width", "height", "bit_depth"):
        assert got[k] == want[k], k
    assert got["worklist_sha256"] == want["worklist_sha256"], "generator work lists changed"
    assert got["references_sha256"] == want["references_sha256"], "reference pictures changed"
    assert got["output_sha256"] == want["output_sha256"], "oracle reconstruction changed"
