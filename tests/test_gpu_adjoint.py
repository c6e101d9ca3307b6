"""Adjoint operators (gsk_solver_adjoint_prepare / gsk_solver_adjoint_apply_device), FP64.

* Transpose test: for random z and lam, lam . (J z) = (J^T lam) . z and (dMdot/df) z =
  (dMdot/df)^T . z, where J z and (dMdot/df) z are central differences of the residual and of Mdot on
  the recorded limiter branches (exact to O(h^2) there, linear to ~1e-10; see
  test_gpu_sensitivity.py) and J^T, (dMdot/df)^T are the hand-written transposes. The boundary
  rounds are run to their fixed point (60 per prepare) so that R(f) is the function the adjoint
  differentiates. A mistake in any transposed path (a dropped face, a wrong slot, sign or
  velocity) shows as a relative mismatch far above the 1e-7 bound; the measured mismatches are
  ~1e-10 to 1e-9. Cases cover every boundary kind and both reflection forms:
    - a tube with a curved Maxwell wall (velocity-interpolated reflection), pressure inlet, vacuum;
    - a box with a pressure inlet, vacuum outlet, a moving diffuse wall, a flat Maxwell wall (mirror
      reflection) and two symmetry planes, with the relative and the absolute limiter eps and the
      positivity limiter;
    - the tube on a tanh-stretched velocity grid (per-velocity weights, node interpolation).
* Adjoint sensitivities: dMdot/d delta_rp and dMdot/d alpha from one adjoint solve against the
  tangent solve (within 1e-6 relative) and against central differences of re-converged solves
  (within 1e-4), on the tube at alpha = 0.8 and 1.0.
"""

from __future__ import annotations

import os

import pytest
import torch

from conftest import T0

pytestmark = pytest.mark.gpu


def _tube(gsk, meshgen, d, alpha=0.8, stretch=0.0, kinetic=None):
    k = {"precision": "fp64"}
    k.update(kinetic or {})
    path = meshgen.tube_case(str(d), n_core=6, n_ring=3, n_axial=6, n_vel=10, vrange=5.0, delta=5.0,
                             accommodation=min(alpha, 0.99), kinetic=k, stretch=stretch)
    return gsk.Setup(path), 0, 2


def _box(gsk, meshgen, d, kinetic=None):
    meshgen.write_pmesh(os.path.join(str(d), "box.pmesh"), meshgen.box_mesh(4, 3, 3, distort=0.15))
    bnd = [dict(patch="xmin", type="pressure_inlet", temperature=T0), dict(patch="xmax", type="vacuum"),
           dict(patch="ymin", type="diffuse_wall", temperature=300.0, velocity=[40.0, 0.0, 0.0]),
           dict(patch="ymax", type="maxwell_wall", accommodation=0.7, temperature=T0),
           dict(patch="zmin", type="symmetry"), dict(patch="zmax", type="symmetry")]
    k = {"precision": "fp64"}
    k.update(kinetic or {})
    path = os.path.join(str(d), "box.toml")
    meshgen.write_case(path, "box.pmesh", "box", bnd, [8, 8, 8], [(-5.0, 5.0)] * 3,
                       dict(L_ref=1.0e-3, T0=T0, delta_rp=2.0), k)
    return gsk.Setup(path), 0, 3


CASES = {
    "tube_maxwell": lambda g, m, d: _tube(g, m, d),
    "tube_stretched": lambda g, m, d: _tube(g, m, d, stretch=1.5),
    "box_relative_eps": lambda g, m, d: _box(g, m, d),
    "box_absolute_eps_positivity": lambda g, m, d: _box(g, m, d, {"limiter_eps": "absolute",
                                                                 "positivity_limiter": True}),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_adjoint_is_the_transpose_of_the_tangent(gsk, meshgen, tmp_path, case):
    import gsis_sim_amd.sensitivity as sens

    print(f"[gpu] adjoint transpose test: {case}", flush=True)
    setup, inlet, wall = CASES[case](gsk, meshgen, tmp_path)
    sv = gsk.Solver(setup)
    prob = sens.SteadyProblem(sv, inlet=inlet, mdot_factor=setup.mass_flow_factor(), wall_patch=wall)
    sv.set_newton_tolerances(0.0, 0.0)
    sv.set_boundary_iterations(60)
    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    for _ in range(40):
        sv.prepare()
        sv.sweep()
    sv.freeze_limiters(True)
    prob._capture()  # records the branches at this state
    g = torch.Generator(device="cuda").manual_seed(11)
    pairs = []
    for _ in range(2):
        z = torch.randn(prob.n, dtype=torch.float64, device="cuda", generator=g) * prob.f0
        lam = torch.randn(prob.n, dtype=torch.float64, device="cuda", generator=g)
        jz, mz = prob._jv(z)
        pairs.append((z, lam, jz, mz))
    prob._residual(prob.f0)  # back at the recorded state
    sv.adjoint_prepare()
    gm = prob._mdot_gradient()
    worst = 0.0
    for z, lam, jz, mz in pairs:
        jtl = prob._jtv(lam)
        lhs, rhs = float(torch.dot(lam, jz)), float(torch.dot(jtl, z))
        m_adj = float(torch.dot(gm, z))
        e_r = abs(lhs - rhs) / abs(lhs)
        e_m = abs(m_adj - mz) / abs(mz)
        print(f"  lam.Jz {lhs:.12e} JTlam.z {rhs:.12e} ({e_r:.1e}); dMdot z {mz:.12e} adjoint {m_adj:.12e} ({e_m:.1e})",
              flush=True)
        worst = max(worst, e_r, e_m)
    sv.freeze_limiters(False)
    sv.close()
    assert worst <= 1e-7


@pytest.mark.parametrize("alpha", [0.8, 1.0])
def test_adjoint_sensitivities_match_tangent_and_central_differences(gsk, meshgen, tmp_path, alpha):
    import gsis_sim_amd.sensitivity as sens

    print(f"[gpu] adjoint sensitivities, Maxwell tube, alpha {alpha}", flush=True)
    delta = 5.0
    setup, inlet, wall = _tube(gsk, meshgen, tmp_path, alpha)
    sv = gsk.Solver(setup)
    prob = sens.SteadyProblem(sv, inlet=inlet, mdot_factor=setup.mass_flow_factor(), wall_patch=wall, delta_rp=delta,
                              alpha=alpha)
    prob.set_parameters(delta, alpha)
    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    mdot = prob.solve(residual_target=1e-12)
    adj = prob.adjoint_sensitivities(polish=60)
    a_stats = dict(prob.stats)
    tan = prob.sensitivities(polish=60)
    print(f"  Mdot {mdot:.10f}; adjoint {adj} ({a_stats.get('adjoint_krylov_iterations')} iterations, true residual "
          f"{a_stats.get('adjoint_true_relative_residual'):.1e}); tangent {tan}", flush=True)

    def mdot_at(dl, al):
        prob.set_parameters(dl, al)
        return prob.solve(residual_target=1e-12)

    h = 1e-4
    fd = {"delta_rp": (mdot_at(delta * (1 + h), alpha) - mdot_at(delta * (1 - h), alpha)) / (2 * delta * h),
          "alpha": (mdot_at(delta, alpha + h) - mdot_at(delta, alpha - h)) / (2 * h)}
    sv.close()
    for p in ("delta_rp", "alpha"):
        e_t = abs(adj[p] - tan[p]) / abs(tan[p])
        e_fd = abs(adj[p] - fd[p]) / abs(fd[p])
        print(f"  {p}: adjoint {adj[p]:.8e} tangent {tan[p]:.8e} ({e_t:.1e}) central differences {fd[p]:.8e} ({e_fd:.1e})",
              flush=True)
        assert e_t <= 1e-6
        assert e_fd <= 1e-4
    assert a_stats["adjoint_true_relative_residual"] <= 1e-6
