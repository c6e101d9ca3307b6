"""Writes tests/golden/setup_digests.json: digests of the kinetic setup's tables for the cases in
tests/setup_digest.py, as this library builds them. tests/test_setup_fixture.py requires every later
build to reproduce them bit for bit (the cell numbering, sweep orders and face tables fix every
multi-block trajectory). Run from the repository root: python tests/golden/make_setup_fixture.py"""

from __future__ import annotations

import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import conftest  # noqa: E402
import setup_digest  # noqa: E402


def main() -> None:
    g = conftest.load_package()
    import gsis_sim_amd.meshgen as mg

    with tempfile.TemporaryDirectory() as d:
        digests = setup_digest.build_digests(g, mg, d)
    with open(setup_digest.FIXTURE, "w") as f:
        json.dump(digests, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {len(digests)} cases to {setup_digest.FIXTURE}")


if __name__ == "__main__":
    main()
