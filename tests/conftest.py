"""Shared test set-up: import paths, the `gpu` marker, device fixture."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "rivet-h26x-codecs_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def device():
    import rivet_hip as rh

    rh.init(0)
    return rh
