"""The tanh-stretched velocity grid ([velocity_grid].stretch > 0), an extension with no reference
counterpart: per-axis nodes xi_i = c + (h / b) atanh(s_i tanh b) at the midpoints s_i of (-1, 1), each
with the weight (dxi/ds) ds. The reference has no such grid (it is uniform, midpoint-rule, one scalar
weight: src/velocity/velocity_grid.cpp:17-24), so nothing here is pinned against it. The grid is
pinned by what the mathematics fixes (exact quadrature of Gaussian moments, odd symmetry, the b -> 0
limit), and the oracle's weighted sums by closed forms (Maxwellian moments, exact conservation of the
discrete equilibrium, equilibrium preservation)."""

from __future__ import annotations

import math

import numpy as np

from conftest import box_case
from test_oracle import OpsCaller


def _grid(setup):
    v = setup.view()
    nv = setup.nv
    xi = np.column_stack([np.ctypeslib.as_array(v.xi_x, shape=(nv,)), np.ctypeslib.as_array(v.xi_y, shape=(nv,)),
                          np.ctypeslib.as_array(v.xi_z, shape=(nv,))])
    w = v.weight * (np.ctypeslib.as_array(v.weight_rel, shape=(nv,)) if v.weight_rel else np.ones(nv))
    n = list(v.grid_n)
    nodes = np.ctypeslib.as_array(v.grid_nodes, shape=(sum(n),)).copy() if v.grid_nodes else None
    return xi, w, n, nodes


def test_stretched_nodes_odd_increasing_and_clustered(gsk, meshgen, tmp_path):
    setup = gsk.Setup(box_case(meshgen, str(tmp_path), n_vel=24, vrange=6.0, stretch=1.5, kinetic={"precision": "fp64"}))
    xi, w, n, nodes = _grid(setup)
    assert nodes is not None and n == [24, 24, 24]
    for a in range(3):
        nd = nodes[sum(n[:a]):sum(n[:a + 1])]
        assert np.array_equal(nd, -nd[::-1])  # odd about the centre, bit for bit (mirror orbits)
        assert np.all(np.diff(nd) > 0) and -6.0 < nd[0] and nd[-1] < 6.0
        d = np.diff(nd)
        assert d[len(d) // 2] < 0.5 < d[0]  # finer than the uniform 0.5 at the centre, coarser at the ends
        assert np.all(np.diff(d[: len(d) // 2]) < 0)  # spacing shrinks monotonically towards the centre
    # every storage point sits on the node lattice
    for a in range(3):
        assert np.all(np.isin(xi[:, a], nodes[sum(n[:a]):sum(n[:a + 1])]))


def test_stretched_weights_integrate_gaussian_moments(gsk, meshgen, tmp_path):
    """sum_p w_p g(xi_p) against closed forms over the box [-6, 6]^3 for g = exp(-|xi|^2 / 2) times
    1, xi_x^2, |xi|^4. The 24-point quadrature errors are 8e-9, 1e-7 and 7e-7; a wrong node or weight
    formula is off by far more than the 2e-6 allowed (the stretched nodes with the uniform weight,
    i.e. a dropped Jacobian, miss the first integral by a factor 3.8)."""
    setup = gsk.Setup(box_case(meshgen, str(tmp_path), n_vel=24, vrange=6.0, stretch=1.5, kinetic={"precision": "fp64"}))
    xi, w, _, _ = _grid(setup)
    r2 = np.sum(xi * xi, axis=1)
    g = np.exp(-0.5 * r2)
    e = math.erf(6.0 / math.sqrt(2.0))
    tail = 12.0 * math.exp(-18.0)  # 2 * 6 * exp(-36/2): boundary term of the truncated moments
    I0 = math.sqrt(2 * math.pi) * e  # int_{-6}^{6} exp(-x^2/2) dx
    I2 = I0 - tail  # int x^2 exp(-x^2/2)
    I4 = 3 * I2 - 6.0 ** 3 * 2 * math.exp(-18.0)  # int x^4 exp(-x^2/2)
    assert abs(np.sum(w * g) / I0 ** 3 - 1) < 2e-6
    assert abs(np.sum(w * g * xi[:, 0] ** 2) / (I2 * I0 ** 2) - 1) < 2e-6
    exact4 = 3 * I4 * I0 ** 2 + 6 * I2 ** 2 * I0  # |xi|^4 = sum x_a^4 + 2 sum_{a<b} x_a^2 x_b^2
    assert abs(np.sum(w * g * r2 * r2) / exact4 - 1) < 2e-6


def test_small_stretch_tends_to_the_uniform_grid(gsk, meshgen, tmp_path):
    uni = _grid(gsk.Setup(box_case(meshgen, str(tmp_path), n_vel=12, vrange=6.0, name="u", kinetic={"precision": "fp64"})))
    st = _grid(gsk.Setup(box_case(meshgen, str(tmp_path), n_vel=12, vrange=6.0, name="s", stretch=1e-4,
                                  kinetic={"precision": "fp64"})))
    assert np.max(np.abs(st[0] - uni[0])) < 1e-7
    assert np.max(np.abs(st[1] / uni[1] - 1)) < 1e-7


def test_stretched_maxwellian_moments_closed_form(gsk, meshgen, oracle_mod, tmp_path):
    """Continuous Maxwellians (the oracle's weighted moment sums) give (rho, u, T), zero stress and heat
    flux to 1e-10 on a 40^3 stretched grid over [-12, 12]."""
    path = box_case(meshgen, str(tmp_path), n=(3, 2, 2), distort=0.0, n_vel=40, vrange=12.0, stretch=1.5,
                    kinetic={"precision": "fp64", "equilibrium": "continuous"})
    setup = gsk.Setup(path)
    x = OpsCaller(oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps), setup)
    rng = np.random.default_rng(11)
    st = np.column_stack([rng.uniform(0.5, 2, setup.n_cells), rng.uniform(-1.5, 1.5, (setup.n_cells, 3)),
                          rng.uniform(0.5, 2, setup.n_cells)])
    x.set_equilibrium(st)
    m = x.moments()
    assert np.allclose(m[:, 0], st[:, 0], rtol=1e-10)
    assert np.allclose(m[:, 1:4], st[:, 1:4], atol=1e-10)
    assert np.allclose(m[:, 4], st[:, 4], rtol=1e-10)
    assert np.max(np.abs(m[:, 6:15])) < 1e-10


def test_stretched_discrete_equilibrium_conserves_exactly(gsk, meshgen, oracle_mod, tmp_path):
    """The weighted Newton sums: the discrete Maxwellian's grid moments equal (rho, u, T) to 1e-10 on a
    coarse stretched [-6, 6] grid where the continuous Maxwellian is visibly off."""
    states = np.array([[1.0, 0.0, 0.0, 0.0, 0.3], [0.5, 1.2, -0.4, 0.3, 0.6], [2.0, -2.0, 0.0, 0.5, 1.5],
                       [1.3, 0.2, 0.1, -0.1, 0.15]])
    res = {}
    for eq in ("discrete", "continuous"):
        path = box_case(meshgen, str(tmp_path), n=(2, 2, 1), distort=0.0, n_vel=12, vrange=6.0, name=f"s_{eq}",
                        stretch=1.5, kinetic={"precision": "fp64", "equilibrium": eq})
        setup = gsk.Setup(path)
        x = OpsCaller(oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps), setup)
        x.set_equilibrium(states)
        res[eq] = x.moments()
    m = res["discrete"]
    assert np.allclose(m[:, 0], states[:, 0], rtol=1e-10, atol=0)
    assert np.allclose(m[:, 1:4], states[:, 1:4], rtol=0, atol=1e-10)
    assert np.allclose(m[:, 4], states[:, 4], rtol=1e-10, atol=0)
    c = res["continuous"]
    assert np.max(np.abs(c[:, 4] - states[:, 4]) / states[:, 4]) > 1e-6  # the correction is doing work


def test_stretched_equilibrium_is_preserved(gsk, meshgen, oracle_mod, tmp_path):
    """Uniform rest gas between walls at its own temperature stays put on a stretched grid."""
    path = box_case(meshgen, str(tmp_path), bcs="closed", n_vel=12, vrange=6.0, wall_T=(1.0, 1.0), stretch=1.5,
                    kinetic={"precision": "fp64"})
    setup = gsk.Setup(path)
    cb = oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps)
    sv = gsk.Solver(setup, cb.ops)
    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    f0 = sv.distribution()
    sv.prepare()
    info = sv.sweep()
    assert info["residual"] < 1e-6
    assert np.max(np.abs(sv.distribution() - f0)) < 1e-9 * np.max(f0)
    sv.close()
