#!/usr/bin/env python3
"""Write tests/golden/reconstruction.json: for every case in cases.py, the
hashes of its work lists, reference pictures and oracle reconstruction.
TEST INFRASTRUCTURE: run from the repository root after
`python -c "import __graft_entry__ as g; g.build()"`."""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
for p in (ROOT, ROOT / "rivet-h26x-codecs_amd"):
    sys.path.insert(0, str(p))

from tests.golden import cases  # noqa: E402


def main():
    data = {name: cases.compute(name) for name in cases.CASES}
    cases.GOLDEN.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")
    print(f"wrote {len(data)} cases to {cases.GOLDEN}")


if __name__ == "__main__":
    main()
