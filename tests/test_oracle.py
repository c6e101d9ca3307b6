"""CPU tests that pin the oracle (oracle/liboracle.so):
  * against the reference's own kernels compiled from the reference tree (oracle/_ref) on identical
    inputs, method by method and over whole CIS iterations -- bitwise;
  * against closed forms and invariants the reference tests hold (Maxwellian moments, discrete
    equilibrium conservation, equilibrium preservation, closed-domain mass / zero wall flux);
  * against the committed golden vectors of the slab benchmark (tests/golden/slab_c1.npz).
A planted mistake in the oracle (a dropped term, a wrong sign) breaks the bitwise comparison."""

from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import box_case, perturbed_distribution, random_inflow, rel_err

HERE = os.path.dirname(os.path.abspath(__file__))


def _pair(gsk, oracle_mod, setup, precision=-1):
    if not oracle_mod.ref_available():
        pytest.skip("reference kernels (oracle/_ref) not built: reference tree absent")
    a = oracle_mod.CpuBackend(setup, "oracle", precision, ops_type=gsk.BackendOps)
    b = oracle_mod.CpuBackend(setup, "ref", precision, ops_type=gsk.BackendOps)
    return a, b


class OpsCaller:
    """Call a gsk_backend_ops table from Python (used on the oracle backends)."""

    def __init__(self, cb, setup):
        import ctypes as C
        self.C = C
        self.cb = cb
        self.s = setup
        o = cb.ops
        self.ctx = C.c_void_p(o.ctx)
        P = C.POINTER
        d = P(C.c_double)
        self.f_set_f = C.CFUNCTYPE(C.c_int, C.c_void_p, d, C.c_size_t)(o.set_distribution)
        self.f_get_f = C.CFUNCTYPE(C.c_int, C.c_void_p, d, C.c_size_t)(o.get_distribution)
        self.f_mom = C.CFUNCTYPE(C.c_int, C.c_void_p, d, C.c_size_t)(o.compute_moments)
        self.f_params = C.CFUNCTYPE(C.c_int, C.c_void_p)(o.update_cell_params)
        self.f_inflow = C.CFUNCTYPE(C.c_int, C.c_void_p, d, C.c_size_t)(o.set_inflow)
        self.f_bsums = C.CFUNCTYPE(C.c_int, C.c_void_p, d, C.c_size_t)(o.boundary_sums)
        self.f_sweep = C.CFUNCTYPE(C.c_int, C.c_void_p, d, C.c_size_t)(o.sweep)
        self.f_eq = C.CFUNCTYPE(C.c_int, C.c_void_p, d, C.c_size_t)(o.set_equilibrium)
        self.f_gsis = C.CFUNCTYPE(C.c_int, C.c_void_p, d, d, C.c_size_t, P(C.c_int32), C.c_size_t,
                                  P(C.c_longlong))(o.gsis_correction)
        self.f_flux = C.CFUNCTYPE(C.c_int, C.c_void_p, P(C.c_int32), P(C.c_int32), C.c_size_t, d)(
            o.internal_face_mass_flux)

    def _d(self, a):
        return a.ctypes.data_as(self.C.POINTER(self.C.c_double))

    def call(self, fn, *args):
        assert fn(self.ctx, *args) == 0

    def set_distribution(self, f):
        f = np.ascontiguousarray(f, np.float64).ravel()
        self.call(self.f_set_f, self._d(f), f.size)

    def get_distribution(self):
        out = np.empty(self.s.n_cells * self.s.nv)
        self.call(self.f_get_f, self._d(out), out.size)
        return out

    def moments(self):
        out = np.empty(self.s.n_cells * 15)
        self.call(self.f_mom, self._d(out), out.size)
        return out.reshape(-1, 15)

    def update_cell_params(self):
        self.call(self.f_params)

    def set_inflow(self, inflow):
        a = np.ascontiguousarray(inflow, np.float64)
        self.call(self.f_inflow, self._d(a), len(a))

    def boundary_sums(self):
        out = np.empty(self.s.n_bfaces * 20)
        self.call(self.f_bsums, self._d(out), out.size)
        return out.reshape(-1, 20)

    def sweep(self):
        out = np.empty(self.s.n_cells)
        self.call(self.f_sweep, self._d(out), out.size)
        return out

    def set_equilibrium(self, states):
        st = np.ascontiguousarray(states, np.float64).ravel()
        self.call(self.f_eq, self._d(st), st.size)

    def gsis_correction(self, wh, wn):
        C = self.C
        wh = np.ascontiguousarray(wh, np.float64).ravel()
        wn = np.ascontiguousarray(wn, np.float64).ravel()
        flags = np.zeros(self.s.n_cells, np.int32)
        rep = C.c_longlong()
        self.call(self.f_gsis, self._d(wh), self._d(wn), wh.size, flags.ctypes.data_as(C.POINTER(C.c_int32)),
                  flags.size, C.byref(rep))
        return rep.value, flags

    def face_flux(self, cells, entries):
        C = self.C
        c = np.ascontiguousarray(cells, np.int32)
        e = np.ascontiguousarray(entries, np.int32)
        out = np.empty(c.size)
        self.call(self.f_flux, c.ctypes.data_as(C.POINTER(C.c_int32)), e.ctypes.data_as(C.POINTER(C.c_int32)), c.size,
                  self._d(out))
        return out


def exercise(x, setup, f, inflow, seed=3):
    """One pass through every backend method on fixed inputs; returns the outputs."""
    out = {}
    x.set_distribution(f)
    out["moments"] = x.moments()
    x.update_cell_params()
    x.set_inflow(inflow)
    out["bsums"] = x.boundary_sums()
    out["residual"] = x.sweep()
    out["f_after_sweep"] = x.get_distribution()
    m = x.moments()
    out["moments2"] = m
    rng = np.random.default_rng(seed)
    wh = m[:, [0, 1, 2, 3, 4]].copy()
    wn = wh * (1.0 + 0.02 * rng.uniform(-1, 1, wh.shape))
    wn[:, 1:4] = wh[:, 1:4] + 0.05 * rng.uniform(-1, 1, (len(wh), 3))
    out["gsis"] = x.gsis_correction(wh, wn)
    out["f_after_gsis"] = x.get_distribution()
    v = setup.view()
    off = np.ctypeslib.as_array(v.csr_offsets, (setup.n_cells + 1,))
    ent = np.ctypeslib.as_array(v.csr, (int(off[-1]),))
    cells, entries = [], []
    for c in range(setup.n_cells):
        for k in range(off[c], off[c + 1]):
            if ent[k]["other"] > c:
                cells.append(c)
                entries.append(k)
    out["face_flux"] = x.face_flux(np.array(cells), np.array(entries))
    x.set_equilibrium(m[:, :5].copy())
    out["f_equilibrium"] = x.get_distribution()
    return out


CASES = {
    "box_mixed_fp64": dict(kinetic={"precision": "fp64"}),
    "box_mixed_fp64_block_rcm_3blocks": dict(kinetic={"precision": "fp64", "sweep_blocks": 3, "cell_ordering": "rcm",
                                                      "gsis_negative_check": "all", "limiter_eps": "absolute"}),
    "box_mixed_fp64_continuous_limit": dict(kinetic={"precision": "fp64", "equilibrium": "continuous",
                                                     "gsis_negative_check": "limit", "sweep_order": "dir13",
                                                     "positivity_limiter": True, "block_passes": 3,
                                                     "velocity_chunk": 64}),
    "box_mixed_fp32": dict(kinetic={"precision": "fp32", "sweep_blocks": 2}),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_kernels_per_method(gsk, meshgen, oracle_mod, tmp_path, name):
    setup = gsk.Setup(box_case(meshgen, str(tmp_path), **CASES[name]))
    a, b = _pair(gsk, oracle_mod, setup)
    f = perturbed_distribution(setup)
    inflow = random_inflow(setup)
    oa = exercise(OpsCaller(a, setup), setup, f, inflow)
    ob = exercise(OpsCaller(b, setup), setup, f, inflow)
    for k in oa:
        if k == "gsis":
            assert oa[k][0] == ob[k][0] and np.array_equal(oa[k][1], ob[k][1])
        else:
            assert np.array_equal(oa[k], ob[k]), (k, rel_err(oa[k], ob[k]))


def test_oracle_matches_reference_kernels_curved_maxwell_tube(gsk, meshgen, oracle_mod, tmp_path):
    """Curved Maxwell wall (alpha = 0.9): the trilinear specular-reflection path."""
    path = meshgen.tube_case(str(tmp_path), n_core=4, n_ring=2, n_axial=3, n_vel=8, vrange=5.0, delta=5.0,
                             accommodation=0.9, kinetic={"precision": "fp64"})
    setup = gsk.Setup(path)
    a, b = _pair(gsk, oracle_mod, setup)
    f = perturbed_distribution(setup, seed=5)
    inflow = random_inflow(setup, seed=6)
    oa = exercise(OpsCaller(a, setup), setup, f, inflow)
    ob = exercise(OpsCaller(b, setup), setup, f, inflow)
    for k in oa:
        if k != "gsis":
            assert np.array_equal(oa[k], ob[k]), k


@pytest.mark.parametrize("bcs", ["mixed", "closed", "vacuum"])
def test_solver_iterations_oracle_vs_reference_kernels(gsk, meshgen, oracle_mod, tmp_path, bcs):
    setup = gsk.Setup(box_case(meshgen, str(tmp_path), bcs=bcs, kinetic={"precision": "fp64"}))
    a, b = _pair(gsk, oracle_mod, setup)
    runs = []
    for cb in (a, b):
        sv = gsk.Solver(setup, cb.ops)
        sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
        hist = []
        for _ in range(4):
            sv.prepare()
            hist.append(sv.sweep()["residual"])
        runs.append((sv.moments(), np.array(hist), sv.boundary_fluxes()[1], sv.patch_mass_flow()))
        sv.close()
    for x, y in zip(runs[0], runs[1]):
        assert np.array_equal(x, y)


def test_maxwellian_moments_closed_form(gsk, meshgen, oracle_mod, tmp_path):
    """Continuous Maxwellians on a resolved grid give (rho, u, T), zero stress and heat flux to 1e-8
    (reference tests/test_physics.cpp:41-77)."""
    path = box_case(meshgen, str(tmp_path), n=(3, 2, 2), distort=0.0, n_vel=40, vrange=12.0,
                    kinetic={"precision": "fp64", "equilibrium": "continuous"})
    setup = gsk.Setup(path)
    x = OpsCaller(oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps), setup)
    rng = np.random.default_rng(11)
    st = np.column_stack([rng.uniform(0.5, 2, setup.n_cells), rng.uniform(-1.5, 1.5, (setup.n_cells, 3)),
                          rng.uniform(0.5, 2, setup.n_cells)])
    x.set_equilibrium(st)
    m = x.moments()
    assert np.allclose(m[:, 0], st[:, 0], rtol=1e-8)
    assert np.allclose(m[:, 1:4], st[:, 1:4], atol=1e-8)
    assert np.allclose(m[:, 4], st[:, 4], rtol=1e-8)
    assert np.allclose(m[:, 5], st[:, 0] * st[:, 4], rtol=1e-8)
    assert np.max(np.abs(m[:, 6:15])) < 1e-8


def test_discrete_equilibrium_conserves_exactly(gsk, meshgen, oracle_mod, tmp_path):
    """The discrete Maxwellian's grid moments equal (rho, u, T) to 1e-10 on a coarse [-6, 6] grid where
    the continuous one is off by ~1e-4 (reference tests/kinetic/test_kinetic_phase5.cpp:77-133)."""
    states = np.array([[1.0, 0.0, 0.0, 0.0, 0.3], [0.5, 1.2, -0.4, 0.3, 0.6], [2.0, -2.0, 0.0, 0.5, 1.5],
                       [1.3, 0.2, 0.1, -0.1, 0.15]])
    res = {}
    for eq in ("discrete", "continuous"):
        path = box_case(meshgen, str(tmp_path), n=(2, 2, 1), distort=0.0, n_vel=16, vrange=6.0, name=f"b_{eq}",
                        kinetic={"precision": "fp64", "equilibrium": eq})
        setup = gsk.Setup(path)
        x = OpsCaller(oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps), setup)
        x.set_equilibrium(states)
        res[eq] = x.moments()
    m = res["discrete"]
    assert np.allclose(m[:, 0], states[:, 0], rtol=1e-10, atol=0)
    assert np.allclose(m[:, 1:4], states[:, 1:4], rtol=0, atol=1e-10)
    assert np.allclose(m[:, 4], states[:, 4], rtol=1e-10, atol=0)
    c = res["continuous"]
    assert np.max(np.abs(c[:, 4] - states[:, 4]) / states[:, 4]) > 1e-5  # the correction is doing work


def test_equilibrium_is_preserved(gsk, meshgen, oracle_mod, tmp_path):
    """Uniform rest gas between walls at its own temperature stays put (residual ~ 0;
    reference tests/kinetic/test_kinetic_cpu.cpp:71-97)."""
    path = box_case(meshgen, str(tmp_path), bcs="closed", n_vel=12, vrange=6.0, wall_T=(1.0, 1.0),
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


def test_closed_domain_mass_and_zero_wall_flux(gsk, meshgen, oracle_mod, tmp_path):
    """Walls at different temperatures: total mass is held and each diffuse wall's net mass flux is 0
    (reference tests/kinetic/test_kinetic_cpu.cpp:99-140)."""
    setup = gsk.Setup(box_case(meshgen, str(tmp_path), bcs="closed", n_vel=10, vrange=6.0,
                               kinetic={"precision": "fp64"}))
    cb = oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps)
    sv = gsk.Solver(setup, cb.ops)
    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    masses = []
    for _ in range(5):
        info = sv.prepare()
        masses.append(info["total_mass"] * info["mass_scale"])
        fp, fm, _ = sv.boundary_fluxes()
        sv.sweep()
    assert np.allclose(masses, masses[0], rtol=1e-13)
    flows = sv.patch_mass_flow()
    assert np.all(np.abs(flows[:4]) < 1e-13)


def test_oracle_matches_golden_slab(gsk, meshgen, oracle_mod, tmp_path):
    """First 10 iterations of the slab benchmark against the committed golden vectors (bitwise: the
    golden run is the oracle itself, cross-checked against oracle/_ref when it was written)."""
    gold = os.path.join(HERE, "golden", "slab_c1.npz")
    if not os.path.exists(gold):
        pytest.skip("golden vectors not generated")
    g = np.load(gold)
    setup = gsk.Setup(meshgen.slab_case(str(tmp_path), kinetic={"precision": "fp64"}))
    cb = oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps)
    sv = gsk.Solver(setup, cb.ops)
    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    hist = []
    for it in range(1, 11):
        sv.prepare()
        info = sv.sweep()
        hist.append((info["residual"], info["epsilon"]))
        if it == 1:
            assert np.array_equal(sv.moments(), g["m1"])
    assert np.array_equal(sv.moments(), g["m10"])
    assert np.array_equal(np.array(hist), g["history"][:10])
