"""The GSIS correction f <- f + f_eq(W_new) - f_eq(W_half) and its negative-value replacement against
their closed forms, restating the reference's tests/kinetic/test_kinetic_cpu.cpp:201-273 (continuous
equilibrium, "created" rule, 2 x 2 x 1 walled box, 12^3 velocities on [-6, 6]): with
f = f_eq(W_half) + df for a smooth df, the corrected f equals f_eq(W_new) + df to 1e-6 of its scale
and no cell is replaced; moving one cell's velocity far enough makes f negative there, and that
cell, only, becomes f_eq(W_new) (1e-7). On the oracle and on the HIP backend, FP64 and FP32."""

from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import kinetic_case

BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]


def maxwellian(rho, u, T, xi):
    c2 = np.sum((xi - np.asarray(u)) ** 2, axis=1)
    return rho / (2 * math.pi * T) ** 1.5 * np.exp(-c2 / (2 * T))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("backend", BACKENDS)
def test_gsis_correction_closed_form(gsk, meshgen, oracle_mod, tmp_path, backend, precision):
    print(f"[{backend}] GSIS correction closed form {precision}", flush=True)
    mesh = meshgen.box_mesh(2, 2, 1, size=(1.0, 1.0, 0.5), names=["wall"] * 6)
    path = kinetic_case(meshgen, str(tmp_path), mesh, [dict(patch="wall", type="diffuse_wall")], n_vel=12, vmax=6.0,
                        kinetic=dict(precision=precision, equilibrium="continuous", gsis_negative_check="created"),
                        name="gsis")
    setup = gsk.Setup(path)
    keep = None
    if backend == "oracle":
        keep = oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps)
        sv = gsk.Solver(setup, keep.ops)
    else:
        sv = gsk.Solver(setup)
    v = setup.view()
    nv, nc = setup.nv, setup.n_cells
    xs = np.column_stack([np.ctypeslib.as_array(p, (nv,)) for p in (v.xi_x, v.xi_y, v.xi_z)])
    xi = np.empty_like(xs)
    xi[setup.storage_to_grid()] = xs  # grid order
    wh = np.array([[1.0 + 0.1 * c, 0.1, -0.05 * c, 0.0, 1.0 + 0.05 * c] for c in range(nc)])
    wn = np.column_stack([wh[:, 0] * 1.02, np.full(nc, 0.12), wh[:, 2] + 0.01, np.full(nc, 0.02), wh[:, 4] * 0.99])
    feq = np.array([maxwellian(w[0], w[1:4], w[4], xi) for w in wh])
    delta = 0.05 * feq * xi[:, 0] * xi[:, 1] / (1.0 + np.sum(xi * xi, axis=1))
    f = feq + delta
    sv.set_distribution(f)
    assert sv.apply_gsis_correction(wh, wn) == 0
    g = sv.distribution()
    expect = np.array([maxwellian(w[0], w[1:4], w[4], xi) for w in wn]) + delta
    assert np.max(np.abs(g - expect)) < 1e-6 * np.max(np.abs(expect))

    sv.set_distribution(f)
    wbig = wn.copy()
    wbig[1, 1] = 2.0  # only cell 1: f - f_eq(W_half) + f_eq(W_new) turns negative there
    assert sv.apply_gsis_correction(wh, wbig) == 1
    h = sv.distribution()
    assert np.max(np.abs(h[1] - maxwellian(wbig[1, 0], wbig[1, 1:4], wbig[1, 4], xi))) < 1e-7
    others = [c for c in range(nc) if c != 1]
    assert np.max(np.abs(h[others] - expect[others])) < 1e-6 * np.max(np.abs(expect))
    sv.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_boundary_iterations_repeat_the_boundary_update(gsk, meshgen, oracle_mod, tmp_path, backend):
    """boundary_iterations = 8 in one prepare equals eight single-iteration prepares separated by a
    zero GSIS correction (which leaves f bit for bit and only invalidates the prepared state), and
    the repeated updates do move the inflow fluxes (the lag they remove is real); restates
    tests/kinetic/test_kinetic_cpu.cpp:511-560 on a distorted box of tetrahedra."""
    print(f"[{backend}] boundary iterations", flush=True)
    mesh = meshgen.tet_box_mesh(3, 2, 2, size=(3.0, 2.0, 2.0), distort=0.2, seed=3,
                                names=["inlet", "outlet", "wall", "wall", "wall", "wall"])
    bcs = [dict(patch="inlet", type="pressure_inlet"), dict(patch="outlet", type="vacuum"),
           dict(patch="wall", type="diffuse_wall")]
    solvers, keep = [], []
    for its in (1, 8):
        path = kinetic_case(meshgen, str(tmp_path), mesh, bcs, n_vel=10, name=f"bi{its}",
                            kinetic=dict(precision="fp64", sweep_blocks=1, positivity_limiter=True,
                                         boundary_iterations=its))
        setup = gsk.Setup(path)
        if backend == "oracle":
            keep.append(oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps))
            solvers.append(gsk.Solver(setup, keep[-1].ops))
        else:
            solvers.append(gsk.Solver(setup))
    a, b = solvers
    centers = np.empty((a.setup.n_cells, 3))
    centers[a.setup.cell_new_to_old()] = a.setup.cell_centers()  # m, L_ref = 1 m
    x, y, z = centers.T
    w = np.column_stack([0.8 + 0.15 * x, 0.05 + 0.03 * y, 0.02 * z, np.zeros_like(x), 0.9 + 0.04 * z])
    a.initialize_equilibrium(w)
    b.initialize_equilibrium(w)
    f = a.distribution()
    first = a.boundary_fluxes()
    m = np.ascontiguousarray(a.moments()[:, :5])
    repeated = first
    for _ in range(1, 8):
        a.apply_gsis_correction(m, m)
        repeated = a.boundary_fluxes()
    batched = b.boundary_fluxes()
    assert np.array_equal(a.distribution(), f) and np.array_equal(b.distribution(), f)
    for k in (0, 1):  # F+ and F-
        assert np.max(np.abs(batched[k] - repeated[k])) <= 1e-12
    assert np.max(np.abs(first[1] - repeated[1])) > 1e-8
    for sv in solvers:
        sv.close()
