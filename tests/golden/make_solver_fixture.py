"""Writes tests/golden/solver_fixture.npz: the outputs of the solver runs in tests/solver_cases.py
(CIS iterations of the library's orchestration on the CPU oracle backend) as this library computes
them; tests/test_solver_fixture.py requires later builds to reproduce them bit for bit. Run from the
repository root: python tests/golden/make_solver_fixture.py"""

from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import conftest  # noqa: E402
import solver_cases  # noqa: E402


def main() -> None:
    g = conftest.load_package()
    import gsis_sim_amd.meshgen as mg
    import oracle

    arrays = {}
    with tempfile.TemporaryDirectory() as d:
        for name in solver_cases.CASES:
            for k, v in solver_cases.run_case(g, oracle, mg, d, name).items():
                arrays[f"{name}/{k}"] = v
    path = os.path.join(HERE, "solver_fixture.npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {len(solver_cases.CASES)} cases to {path}")


if __name__ == "__main__":
    main()
