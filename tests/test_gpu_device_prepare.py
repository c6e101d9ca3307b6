"""The prepare of each CIS iteration on the device (the HIP backend's control_prepare: moments,
positivity check, mass and closed-domain rescaling, eq. (59) epsilon, volume means, the
pressure-face reconstruction with the characteristic condition and its discrete-equilibrium
correction, the boundary rounds into face fluxes, face states, wall densities and patch mass flows)
on every case of tests/solver_cases.py, which between them take every boundary branch of the
orchestration (pressure and mass-flow inlets, pressure outlets, vacuum, diffuse and curved Maxwell
walls, symmetry, a closed domain, wall relaxation, several boundary rounds, cell reconstruction, no
limiter, the continuous equilibrium):

- against the same HIP backend with the prepare on the host (the library's host orchestration,
  src/kinetic/kinetic_solver.cpp:172-531 restated): the only differences are the order of the sums
  over cells and faces, so everything agrees to 1e-12 of its scale (patch flows on the scale of the
  one-sided mass fluxes, since a closed box's net flows are zero up to round-off);
- against the CPU oracle's runs of the same cases (tests/golden/solver_fixture.npz) at the FP64
  parity bar, 1e-10 of each quantity's scale."""

from __future__ import annotations

import os

import numpy as np
import pytest

import solver_cases

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "solver_fixture.npz")
HISTORY = ("epsilon", "total mass", "mass scale", "residual")


def _compare(got, ref, bound, what):
    for k, e in ref.items():
        v = got[k]
        assert v.shape == e.shape, (what, k)
        if k == "history":  # each column on its own scale
            for j, col in enumerate(HISTORY):
                scale = max(float(np.max(np.abs(e[:, j]))), 1e-300)
                err = float(np.max(np.abs(v[:, j] - e[:, j]))) / scale
                assert err <= bound, (what, col, err)
            continue
        scale = max(float(np.max(np.abs(e))), 1e-300) if e.size else 1.0
        if k == "flows":  # net flows that cancel to round-off (a closed box) compare on the one-sided fluxes
            scale = max(scale, float(np.sum(np.abs(ref["f_plus"][:, 0]))))
        err = float(np.max(np.abs(v - e))) / scale if e.size else 0.0
        assert err <= bound, (what, k, err)


@pytest.mark.parametrize("name", sorted(solver_cases.CASES))
def test_device_prepare_matches_host_prepare_and_oracle(gsk, meshgen, oracle_mod, tmp_path, name):
    print(f"[gpu] device prepare vs host prepare vs oracle fixture: {name}", flush=True)
    dev = solver_cases.run_case(gsk, oracle_mod, meshgen, str(tmp_path / "dev"), name, device_prepare=True)
    host = solver_cases.run_case(gsk, oracle_mod, meshgen, str(tmp_path / "host"), name, device_prepare=False)
    _compare(dev, host, 1e-12, f"{name}: device vs host prepare")
    fixture = np.load(FIXTURE)
    oracle = {k: fixture[f"{name}/{k}"] for k in dev}
    _compare(dev, oracle, 1e-10, f"{name}: device prepare vs oracle")
