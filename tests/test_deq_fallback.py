"""A cell whose moments the velocity grid cannot hold: the reference's known-answer case for
[kinetic].deq_fallback (tests/kinetic/test_kinetic_phase5.cpp:264-327). On a 2 x 1 x 1 box with a
12^3 grid on [-6, 6] (points +-0.5, +-1.5, ...), cell 0 holds a distribution with negative values and
T = 0.05, below the lowest temperature a non-negative distribution on that grid can have (about
0.083 at its velocity), so the discrete-equilibrium Newton cannot converge there; cell 1 holds a
Maxwellian. After prepare exactly one cell has failed. The GSIS correction (density +2% in both
cells) then corrects cell 1 to its new density (1e-6), and for cell 0 either leaves f untouched and
counts it as replaced ("collisionless"; a following sweep stays finite) or rebuilds it from the
continuous Maxwellian ("continuous"). Run through the oracle here and through the HIP backend in
the GPU suite."""

from __future__ import annotations

import math
import os

import numpy as np
import pytest

from conftest import T0


def build_case(gsk, mg, directory, fallback):
    """make_kinetic_box({2,1,1}, (0,0,0)-(1,0.5,0.5)) with patches inlet / outlet / wall, and
    make_kinetic_case with Kn = 0.5, L_ref = 1 m, a pressure inlet at p_ref, T0 (kinetic_test_cases.cpp)."""
    xs, ys, zs = np.array([0.0, 0.5, 1.0]), np.array([0.0, 0.5]), np.array([0.0, 0.5])
    pts = np.array([[x, y, z] for x in xs for y in ys for z in zs])
    nid = lambda i, j, k: (i * 2 + j) * 2 + k  # noqa: E731
    cells = np.array([[nid(i + (c & 1), (c >> 1) & 1, (c >> 2) & 1) for c in range(8)] for i in range(2)])

    def patch_of(fc, n):
        return np.where(fc[:, 0] < 1e-12, 0, np.where(fc[:, 0] > 1.0 - 1e-12, 1, 2))

    mesh = mg.hex_to_poly(pts, cells, patch_of, ["inlet", "outlet", "wall"], ["patch"] * 3)
    mg.write_pmesh(os.path.join(directory, "fb.pmesh"), mesh, {"generator": "tests/test_deq_fallback.py"})
    path = os.path.join(directory, f"fb_{fallback}.toml")

    def write(p_inlet):
        mg.write_case(path, "fb.pmesh", "deq_fallback",
                      [dict(patch="inlet", type="pressure_inlet", temperature=T0, pressure=p_inlet),
                       dict(patch="outlet", type="vacuum"), dict(patch="wall", type="diffuse_wall", temperature=T0)],
                      [12, 12, 12], [(-6.0, 6.0)] * 3, dict(L_ref=1.0, T0=T0, Kn=0.5),
                      dict(precision="fp64", deq_fallback=fallback))

    write(1.0)  # probe for p_ref, as make_kinetic_case does
    write(gsk.Setup(path).describe()["reference"]["pressure_Pa"])
    return path


def grid_velocities(setup):
    v = setup.view()
    nv = setup.nv
    xi = np.column_stack([np.ctypeslib.as_array(v.xi_x, (nv,)), np.ctypeslib.as_array(v.xi_y, (nv,)),
                          np.ctypeslib.as_array(v.xi_z, (nv,))])
    out = np.empty_like(xi)
    out[setup.storage_to_grid()] = xi
    return out, v.weight


def check_fallback(gsk, setup, ops, fallback):
    xi, w = grid_velocities(setup)
    f = np.zeros((2, setup.nv))
    f[1] = (2 * math.pi) ** -1.5 * np.exp(-0.5 * np.sum(xi * xi, axis=1))  # Maxwellian (1, 0, 1)
    x, y, z = xi[:, 0], xi[:, 1], xi[:, 2]
    ax = np.where(np.abs(x - 1.5) < 1e-9, 1.1, np.where((np.abs(x - 0.5) < 1e-9) | (np.abs(x - 2.5) < 1e-9), -0.05, 0.0))
    by = np.where((np.abs(y - 0.5) < 1e-9) | (np.abs(y - 1.5) < 1e-9), 0.5, 0.0)
    cz = np.where(np.abs(z - 0.5) < 1e-9, 1.0, 0.0)
    f[0] = ax * by * cz / w
    sv = gsk.Solver(setup, ops) if ops is not None else gsk.Solver(setup)
    sv.set_distribution(f)
    info = sv.prepare()
    m = sv.moments()
    assert m[0, 4] == pytest.approx(0.05, rel=1e-12) and np.allclose(m[0, 1:4], [1.5, 1.0, 0.5], atol=1e-12)
    assert info["deq_failed"] == 1, info
    wh = np.ascontiguousarray(m[:, :5])
    wn = wh.copy()
    wn[:, 0] *= 1.02
    replaced = sv.apply_gsis_correction(wh, wn)
    m2 = sv.moments()
    assert abs(m2[1, 0] / wn[1, 0] - 1.0) < 1e-6
    d0 = np.max(np.abs(sv.distribution()[0] - f[0]))
    if fallback == "collisionless":
        assert replaced == 1
        assert d0 == 0.0  # f of the unrepresentable cell is left as it is
        sv.sweep()
        assert np.all(np.isfinite(sv.distribution()))
    else:
        assert d0 > 0.0  # rebuilt from the continuous Maxwellian
    sv.close()


@pytest.mark.parametrize("fallback", ["collisionless", "continuous"])
def test_unrepresentable_cell_falls_back_oracle(gsk, meshgen, oracle_mod, tmp_path, fallback):
    setup = gsk.Setup(build_case(gsk, meshgen, str(tmp_path), fallback))
    cb = oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps)
    check_fallback(gsk, setup, cb.ops, fallback)


@pytest.mark.gpu
@pytest.mark.parametrize("fallback", ["collisionless", "continuous"])
def test_unrepresentable_cell_falls_back_hip(gsk, meshgen, tmp_path, fallback):
    print(f"[gpu] discrete-equilibrium fallback {fallback}", flush=True)
    setup = gsk.Setup(build_case(gsk, meshgen, str(tmp_path), fallback))
    check_fallback(gsk, setup, None, fallback)
