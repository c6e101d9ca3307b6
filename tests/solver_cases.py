"""Solver runs whose outputs tests/golden/solver_fixture.npz records (tests/golden/make_solver_fixture.py)
and tests/test_solver_fixture.py requires bit for bit: a few CIS iterations of the library's FP64
orchestration (prepare / sweep: moments, closed-domain scaling, eq. (59) epsilon, wall densities,
pressure-inlet reconstruction and its discrete-equilibrium correction, boundary sums into face
fluxes, face states and patch mass flows, mass-flow inlet pressure updates) on the CPU oracle
backend. Every boundary branch of the orchestration is exercised by at least one case."""

from __future__ import annotations

import os

import numpy as np

import conftest

T0 = conftest.T0


def _mass_flow_box(mg, d):
    """Box with a mass-flow inlet (zero tangential velocity) and a pressure outlet, diffuse walls."""
    mg.write_pmesh(os.path.join(d, "mfbox.pmesh"), mg.box_mesh(5, 3, 2, distort=0.2, seed=5))
    bnd = [dict(patch="xmin", type="mass_flow_inlet", mass_flow=2.0e-9, temperature=T0, tangential_velocity="zero"),
           dict(patch="xmax", type="pressure_outlet", pressure=0.3),
           dict(patch="ymin", type="diffuse_wall", temperature=T0), dict(patch="ymax", type="diffuse_wall", temperature=T0),
           dict(patch="zmin", type="symmetry"), dict(patch="zmax", type="symmetry")]
    path = os.path.join(d, "mfbox.toml")
    mg.write_case(path, "mfbox.pmesh", "mfbox", bnd, [8] * 3, [(-5.0, 5.0)] * 3,
                  dict(L_ref=1.0e-3, T0=T0, delta_rp=1.0), {"precision": "fp64"})
    return path


# name -> (case builder (meshgen, directory) -> toml path, iterations, update mass-flow pressure)
CASES = {
    "box_mixed": (lambda mg, d: conftest.box_case(mg, d, kinetic={"precision": "fp64"}), 5, False),
    "box_mixed_cell_reconstruction_relaxed": (
        lambda mg, d: conftest.box_case(mg, d, kinetic={"precision": "fp64", "pressure_boundary_reconstruction": "cell",
                                                        "boundary_iterations": 3, "wall_relaxation": 0.7,
                                                        "inner_sweeps": 2}), 4, False),
    "box_mixed_no_limiter_continuous": (
        lambda mg, d: conftest.box_case(mg, d, kinetic={"precision": "fp64", "limiter": "none",
                                                        "equilibrium": "continuous"}), 4, False),
    "box_closed": (lambda mg, d: conftest.box_case(mg, d, bcs="closed", kinetic={"precision": "fp64"}), 5, False),
    "box_vacuum_3blocks": (lambda mg, d: conftest.box_case(mg, d, bcs="vacuum", kinetic={"precision": "fp64",
                                                                                          "sweep_blocks": 3}), 5, False),
    "tube_curved_maxwell": (lambda mg, d: mg.tube_case(d, n_core=4, n_ring=2, n_axial=3, n_vel=8, vrange=5.0, delta=5.0,
                                                       accommodation=0.9, kinetic={"precision": "fp64"}), 5, False),
    "mass_flow_box": (_mass_flow_box, 5, True),
}


def run_case(g, oracle_mod, mg, directory, name, device_prepare=None):
    """The case on the CPU oracle backend (device_prepare None), or on the HIP backend with each
    iteration's prepare on the device (True) or on the host (False)."""
    make, iterations, update_pressure = CASES[name]
    d = os.path.join(directory, name)
    os.makedirs(d, exist_ok=True)
    setup = g.Setup(make(mg, d))
    cb = oracle_mod.CpuBackend(setup, "oracle", ops_type=g.BackendOps) if device_prepare is None else None
    sv = g.Solver(setup, cb.ops) if cb is not None else g.Solver(setup)
    if device_prepare is not None:
        sv.set_device_prepare(device_prepare)
    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    hist, where = [], []
    for _ in range(iterations):
        p = sv.prepare()
        where.append(sv.prepared_on_device())
        s = sv.sweep(update_pressure)
        hist.append([p["epsilon"] if p["has_epsilon"] else 0.0, p["total_mass"], p["mass_scale"], s["residual"]])
    fp, fm, fs = sv.boundary_fluxes()
    out = {"moments": sv.moments(), "f_plus": fp, "f_minus": fm, "face_state": fs, "flows": sv.patch_mass_flow(),
           "history": np.array(hist)}
    sv.close()
    if cb is not None:
        cb.close()
    if device_prepare is not None:
        assert where == [device_prepare] * iterations, where
    return out
