"""The solver orchestration reproduces tests/golden/solver_fixture.npz bit for bit (the runs of
tests/solver_cases.py on the CPU oracle backend): moments, boundary fluxes, face states, patch mass
flows and the per-iteration epsilon / mass / residual records."""

from __future__ import annotations

import os

import numpy as np
import pytest

import solver_cases

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "solver_fixture.npz")


@pytest.mark.parametrize("name", sorted(solver_cases.CASES))
def test_solver_outputs_match_fixture(gsk, meshgen, oracle_mod, tmp_path, name):
    expected = np.load(FIXTURE)
    got = solver_cases.run_case(gsk, oracle_mod, meshgen, str(tmp_path), name)
    for k, v in got.items():
        e = expected[f"{name}/{k}"]
        assert v.shape == e.shape, k
        diff = float(np.max(np.abs(v - e))) if v.size else 0.0
        assert np.array_equal(v, e), f"{name}/{k}: max difference {diff:.3e}"
