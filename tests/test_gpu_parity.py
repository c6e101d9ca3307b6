"""GPU parity tests: the HIP backend (through the C ABI) against the CPU oracle on identical inputs.

FP64 (the drop-in form): every backend method within 1e-12 relative (the only differences are the
order of FP64 reductions over velocities and last-ulp differences of exp), and whole CIS iterations
within 1e-10 relative (the bar the reference's own GPU-vs-CPU test uses with default settings,
tests/kinetic/test_kinetic_gpu.cpp:133-158). FP32 f: every method within 2e-6 and five CIS
iterations within 1e-6 of the FP32 oracle (measured 8.5e-7 and 9e-8; both compute in the reference's
operation order, so only reduction order and last-ulp exp / division differ; the reference's own
FP32 tolerance is 2e-5). Also: determinism and velocity-chunk independence, bitwise."""

from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import box_case, perturbed_distribution, random_inflow, rel_err
from test_oracle import OpsCaller, exercise

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


class HipCaller:
    """The HIP Backend with the method names exercise() uses."""

    def __init__(self, b):
        self.b = b

    def set_distribution(self, f):
        self.b.set_distribution(f)

    def get_distribution(self):
        return self.b.get_distribution().ravel()

    def moments(self):
        return self.b.compute_moments()

    def update_cell_params(self):
        self.b.update_cell_params()

    def set_inflow(self, inflow):
        self.b.set_inflow(inflow)

    def boundary_sums(self):
        return self.b.boundary_sums()

    def sweep(self):
        return self.b.sweep()

    def set_equilibrium(self, st):
        self.b.set_equilibrium(st)

    def gsis_correction(self, wh, wn):
        return self.b.gsis_correction(wh, wn)

    def face_flux(self, cells, entries):
        return self.b.internal_face_mass_flux(cells, entries)


METHOD_CASES = {
    "box_mixed_fp64": (dict(kinetic={"precision": "fp64"}), 1e-12),
    "box_mixed_fp64_rcm_3blocks_all": (dict(kinetic={"precision": "fp64", "sweep_blocks": 3, "cell_ordering": "rcm",
                                                     "gsis_negative_check": "all", "limiter_eps": "absolute"}), 1e-12),
    "box_mixed_fp64_dir13_limit_chunked": (dict(kinetic={"precision": "fp64", "equilibrium": "continuous",
                                                         "gsis_negative_check": "limit", "sweep_order": "dir13",
                                                         "positivity_limiter": True, "block_passes": 3,
                                                         "velocity_chunk": 64}), 1e-12),
    "box_mixed_fp32": (dict(kinetic={"precision": "fp32", "sweep_blocks": 2}), 2e-6),
    # tanh-stretched velocity grid (an extension; the oracle carries the same per-velocity weights)
    "box_mixed_fp64_stretched": (dict(kinetic={"precision": "fp64"}, stretch=1.5), 1e-12),
    "box_mixed_fp32_stretched": (dict(kinetic={"precision": "fp32", "sweep_blocks": 2}, stretch=1.5), 2e-6),
    # the same box split into tetrahedra (four faces per cell: padded slots in the gathered kernels)
    "tet_box_mixed_fp64": (dict(kinetic={"precision": "fp64", "sweep_blocks": 3}, tet=True), 1e-12),
    "tet_box_mixed_fp64_positivity": (dict(kinetic={"precision": "fp64", "positivity_limiter": True}, tet=True), 1e-12),
    "tet_box_mixed_fp32": (dict(kinetic={"precision": "fp32"}, tet=True), 2e-6),
}


@pytest.mark.parametrize("name", sorted(METHOD_CASES))
def test_backend_methods_match_oracle(gsk, meshgen, oracle_mod, tmp_path, name):
    kw, tol = METHOD_CASES[name]
    print(f"[gpu] backend methods {name}", flush=True)
    setup = gsk.Setup(box_case(meshgen, str(tmp_path), **kw))
    f = perturbed_distribution(setup)
    inflow = random_inflow(setup)
    cpu = exercise(OpsCaller(oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps), setup), setup, f, inflow)
    gpu = exercise(HipCaller(gsk.Backend(setup)), setup, f, inflow)
    for k in cpu:
        if k == "gsis":
            assert gpu[k][0] == cpu[k][0] and np.array_equal(gpu[k][1], cpu[k][1])
        else:
            e = rel_err(gpu[k], cpu[k])
            assert e <= tol, (k, e)


def _run_solver(gsk, setup, ops, iterations, snapshot_every=None):
    sv = gsk.Solver(setup, ops) if ops is not None else gsk.Solver(setup)
    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    hist = []
    for it in range(iterations):
        sv.prepare()
        info = sv.sweep()
        hist.append((info["residual"], info["epsilon"]))
    out = dict(moments=sv.moments(), hist=np.array(hist), flows=sv.patch_mass_flow(),
               f=sv.distribution(), fluxes=sv.boundary_fluxes())
    sv.close()
    return out


SOLVER_CASES = {
    "box_mixed": lambda mg, d: box_case(mg, d, kinetic={"precision": "fp64"}),
    "box_closed": lambda mg, d: box_case(mg, d, bcs="closed", kinetic={"precision": "fp64"}),
    # relaxed diffuse-wall density between inner sweeps (the HIP solver refreshes walls on the
    # device, the oracle on the host)
    "box_mixed_wall_relaxation": lambda mg, d: box_case(mg, d, kinetic={"precision": "fp64", "wall_relaxation": 0.6,
                                                                        "inner_sweeps": 3}),
    "box_vacuum_3blocks": lambda mg, d: box_case(mg, d, bcs="vacuum", kinetic={"precision": "fp64", "sweep_blocks": 3}),
    "tube_curved_maxwell": lambda mg, d: mg.tube_case(d, n_core=4, n_ring=2, n_axial=3, n_vel=8, vrange=5.0, delta=5.0,
                                                     accommodation=0.9, kinetic={"precision": "fp64"}),
    "slab_200_24cubed": lambda mg, d: mg.slab_case(d, kinetic={"precision": "fp64"}),
    # tanh-stretched velocity grid; the curved Maxwell wall interpolates between stretched nodes
    "box_mixed_stretched": lambda mg, d: box_case(mg, d, kinetic={"precision": "fp64"}, stretch=1.5),
    "tet_box_mixed": lambda mg, d: box_case(mg, d, kinetic={"precision": "fp64", "sweep_blocks": 3}, tet=True),
    "tet_box_mixed_positivity": lambda mg, d: box_case(mg, d, kinetic={"precision": "fp64", "positivity_limiter": True},
                                                       tet=True),
    "tube_curved_maxwell_stretched": lambda mg, d: mg.tube_case(d, n_core=4, n_ring=2, n_axial=3, n_vel=8, vrange=5.0,
                                                               delta=5.0, accommodation=0.9, stretch=1.2,
                                                               kinetic={"precision": "fp64"}),
}


@pytest.mark.parametrize("name", sorted(SOLVER_CASES))
def test_solver_iterations_match_oracle(gsk, meshgen, oracle_mod, tmp_path, name):
    print(f"[gpu] solver iterations {name}", flush=True)
    setup = gsk.Setup(SOLVER_CASES[name](meshgen, str(tmp_path)))
    its = 3 if name.startswith("slab") else 6
    cb = oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps)
    cpu = _run_solver(gsk, setup, cb.ops, its)
    gpu = _run_solver(gsk, setup, None, its)
    assert rel_err(gpu["moments"], cpu["moments"]) <= 1e-10
    assert rel_err(gpu["f"], cpu["f"]) <= 1e-10
    assert rel_err(gpu["hist"][:, 0], cpu["hist"][:, 0]) <= 1e-10
    for a, b in zip(gpu["fluxes"], cpu["fluxes"]):
        assert rel_err(a, b) <= 1e-10
    assert np.allclose(gpu["flows"], cpu["flows"], rtol=1e-10, atol=1e-12 * max(1.0, np.max(np.abs(cpu["flows"]))))


@pytest.mark.parametrize("name", ["box_mixed", "tube_curved_maxwell"])
def test_csr_path_cells_match_oracle(gsk, meshgen, oracle_mod, tmp_path, monkeypatch, name):
    """Cells with more faces than the gathered (ELL) kernels hold are recomputed from the CSR after
    those kernels; GSK_ELL_MAX_FACES=0 sends every cell down that path (the box mixes hex cells
    with boundary faces of every kind)."""
    print(f"[gpu] CSR-path cells {name}", flush=True)
    monkeypatch.setenv("GSK_ELL_MAX_FACES", "0")
    setup = gsk.Setup(SOLVER_CASES[name](meshgen, str(tmp_path)))
    cb = oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps)
    cpu = _run_solver(gsk, setup, cb.ops, 4)
    gpu = _run_solver(gsk, setup, None, 4)
    assert rel_err(gpu["f"], cpu["f"]) <= 1e-10
    assert rel_err(gpu["hist"][:, 0], cpu["hist"][:, 0]) <= 1e-10


def test_split_launches_match_oracle(gsk, meshgen, oracle_mod, tmp_path, monkeypatch):
    """Large meshes exceed what one launch may hold: the gradient / residual then take several
    launches and the CSR-path kernels loop; GSK_GRID_CAP=16 forces both on a small mesh."""
    print("[gpu] split launches box_mixed", flush=True)
    monkeypatch.setenv("GSK_GRID_CAP", "16")
    monkeypatch.setenv("GSK_ELL_MAX_FACES", "5")
    setup = gsk.Setup(SOLVER_CASES["box_mixed"](meshgen, str(tmp_path)))
    cb = oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps)
    cpu = _run_solver(gsk, setup, cb.ops, 3)
    gpu = _run_solver(gsk, setup, None, 3)
    assert rel_err(gpu["f"], cpu["f"]) <= 1e-10
    assert rel_err(gpu["hist"][:, 0], cpu["hist"][:, 0]) <= 1e-10


def test_slab_benchmark_300_iterations_match_golden(gsk, meshgen, tmp_path):
    """The 1D slab benchmark (200 cells, 24^3 on [-6, 6], delta = 1, 300 CIS iterations, FP64) against
    the committed oracle vectors (tests/golden/slab_c1.npz): moments within 1e-10 relative."""
    g = np.load(os.path.join(HERE, "golden", "slab_c1.npz"))
    n = int(g["iterations"])
    print(f"[gpu] slab benchmark {n} iterations vs golden", flush=True)
    setup = gsk.Setup(meshgen.slab_case(str(tmp_path), kinetic={"precision": "fp64"}))
    sv = gsk.Solver(setup)
    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    hist = []
    for it in range(1, n + 1):
        sv.prepare()
        info = sv.sweep()
        hist.append((info["residual"], info["epsilon"]))
        if it == 1:
            assert rel_err(sv.moments(), g["m1"]) <= 1e-12
        if it == 10:
            assert rel_err(sv.moments(), g["m10"]) <= 1e-11
    m = sv.moments()
    assert rel_err(m, g["mN"]) <= 1e-10
    for k in (0, 1, 4):  # rho, u_x, T individually
        assert rel_err(m[:, k], g["mN"][:, k]) <= 1e-10, k
    assert rel_err(np.array(hist)[:, 0], g["history"][:, 0]) <= 1e-9
    fp, fm, _ = sv.boundary_fluxes()
    assert rel_err(fp, g["f_plus"]) <= 1e-10 and rel_err(fm, g["f_minus"]) <= 1e-10
    sv.close()


def test_fp32_form_tracks_fp32_oracle(gsk, meshgen, oracle_mod, tmp_path):
    print("[gpu] fp32 form vs fp32 oracle", flush=True)
    setup = gsk.Setup(box_case(meshgen, str(tmp_path), kinetic={"precision": "fp32"}))
    cb = oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps)
    cpu = _run_solver(gsk, setup, cb.ops, 5)
    gpu = _run_solver(gsk, setup, None, 5)
    assert rel_err(gpu["moments"], cpu["moments"]) <= 1e-6
    assert rel_err(gpu["f"], cpu["f"]) <= 2e-6
    assert rel_err(gpu["hist"][:, 0], cpu["hist"][:, 0]) <= 1e-6


def test_deterministic_and_chunk_independent(gsk, meshgen, tmp_path):
    print("[gpu] determinism and chunk independence", flush=True)
    path = box_case(meshgen, str(tmp_path), kinetic={"precision": "fp64"})
    runs = []
    for chunk in (0, 0, 64):
        setup = gsk.Setup(path, {"velocity_chunk": chunk})
        runs.append(_run_solver(gsk, setup, None, 4))
    for r in runs[1:]:
        assert np.array_equal(r["moments"], runs[0]["moments"])
        assert np.array_equal(r["f"], runs[0]["f"])


def _moment_errors(a, b):
    """Per-column max error of moments a vs b: rho, u, T, p relative to the column's largest magnitude;
    the stress deviator (6 columns) and heat flux (3) relative to the largest pressure, the size of the
    tensors they belong to (they are themselves small differences of it)."""
    p_scale = np.max(np.abs(b[:, 5]))
    out = [rel_err(a[:, k], b[:, k]) for k in range(6)]
    out += [float(np.max(np.abs(a[:, k] - b[:, k])) / p_scale) for k in range(6, a.shape[1])]
    return out


@pytest.mark.parametrize("stretch", [0.0, 1.5])
def test_mfma_moment_sums_track_fp64_sums(gsk, meshgen, tmp_path, stretch):
    """FP32 form: the matrix-core moment contraction (FP32 products and accumulation, FP64 block
    sums) against the FP64 sums of the same FP32 f, on 24^3 velocities (uniform and stretched).
    Stated errors: rho, u, T, p within 2e-6 of their scale, stress deviator and heat flux within 2e-6
    of the pressure."""
    print(f"[gpu] mfma moment sums, stretch {stretch}", flush=True)
    path = box_case(meshgen, str(tmp_path), n=(4, 3, 2), n_vel=24, vrange=6.0, kinetic={"precision": "fp32"},
                    stretch=stretch)
    ref = gsk.Backend(gsk.Setup(path))
    mma = gsk.Backend(gsk.Setup(path, {"moment_sums": "mfma"}))
    f = perturbed_distribution(ref.setup, seed=5).astype(np.float32).astype(np.float64)
    errs = []
    for _ in range(2):  # first about zero shifts, then about each cell's velocity
        ref.set_distribution(f)
        mma.set_distribution(f)
        e = _moment_errors(mma.compute_moments(), ref.compute_moments())
        errs.append(e)
        print("  per-column errors", ["%.1e" % x for x in e], flush=True)
    for e in errs:
        assert max(e) <= 2e-6, e


def test_mfma_moment_form_tracks_fp32_oracle(gsk, meshgen, oracle_mod, tmp_path):
    print("[gpu] mfma moment form vs FP32 oracle", flush=True)
    path = box_case(meshgen, str(tmp_path), kinetic={"precision": "fp32", "moment_sums": "mfma"})
    setup = gsk.Setup(path)
    cpu_setup = gsk.Setup(path, {"moment_sums": "fp64"})
    cb = oracle_mod.CpuBackend(cpu_setup, "oracle", ops_type=gsk.BackendOps)
    cpu = _run_solver(gsk, cpu_setup, cb.ops, 4)
    gpu = _run_solver(gsk, setup, None, 4)
    assert rel_err(gpu["moments"], cpu["moments"]) <= 1e-4
    assert rel_err(gpu["f"], cpu["f"]) <= 1e-4



@pytest.mark.parametrize("form", ["fp32", "mixed"])
def test_fp32_forms_against_fp64_oracle(gsk, meshgen, oracle_mod, tmp_path, form):
    """The FP32 forms (f stored and swept in FP32; FP64 moment sums, discrete-equilibrium Newton,
    residual norm and mass flows) against the all-FP64 oracle on the same case: six CIS iterations of
    the mixed-boundary box on 16^3 velocities. Stated errors: moments 1e-6 relative to each column's
    scale (velocity components to the largest speed, stress and heat flux to the pressure), residual
    norm 1e-5 relative, patch mass flows 1e-6 of the largest flow. Measured: FP32 8.3e-8, 7.5e-7,
    8.9e-8; mixed (matrix-core moment sums) 3.0e-7, 2.4e-6, 1.0e-7."""
    print(f"[gpu] {form} form vs FP64 oracle", flush=True)
    kin = {"precision": "fp32"} if form == "fp32" else {"precision": "fp32", "moment_sums": "mfma"}
    path = box_case(meshgen, str(tmp_path), n=(4, 3, 2), n_vel=16, vrange=6.0, kinetic=kin)
    ref_setup = gsk.Setup(path, {"precision": "fp64", "moment_sums": "fp64"})
    cb = oracle_mod.CpuBackend(ref_setup, "oracle", ops_type=gsk.BackendOps)
    cpu = _run_solver(gsk, ref_setup, cb.ops, 6)
    gpu = _run_solver(gsk, gsk.Setup(path), None, 6)
    e_m = _moment_errors(gpu["moments"], cpu["moments"])
    # velocity components against the largest speed (the symmetry planes keep u_z near zero, so its
    # own column scale would inflate a tiny absolute error)
    u_scale = np.max(np.abs(cpu["moments"][:, 1:4]))
    e_m[1:4] = [float(np.max(np.abs(gpu["moments"][:, k] - cpu["moments"][:, k])) / u_scale) for k in range(1, 4)]
    e_r = np.max(np.abs(gpu["hist"][:, 0] / cpu["hist"][:, 0] - 1))
    e_q = np.max(np.abs(gpu["flows"] - cpu["flows"])) / np.max(np.abs(cpu["flows"]))
    print(f"  moments {['%.1e' % x for x in e_m]} residual {e_r:.1e} flows {e_q:.1e}", flush=True)
    assert max(e_m) <= 1e-6, e_m
    assert e_r <= 1e-5
    assert e_q <= 1e-6
