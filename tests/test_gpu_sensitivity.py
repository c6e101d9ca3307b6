"""Sensitivities of the converged inlet mass flow (gsis-sim_amd/sensitivity.py), FP64.

* Branch freezing (gsk_solver_freeze_limiters): the residual evaluated on the recorded limiter
  branches equals the plain residual at the recorded state (an unconverged state, where R is far
  from 0; to within the plain residual's own repeat-to-repeat difference, which comes from the
  pressure-inlet rounds each evaluation runs again, or 1e-14), and at the converged state the central-difference
  Jacobian products on those branches are linear in the direction (J(a z1 + b z2) = a J z1 + b J z2)
  and independent of the difference step to 1e-7 relative; the same products of the plain residual
  are not (they cross limiter branches).
* dMdot/d delta_rp and dMdot/d alpha from the tangent solves (one per parameter) against central
  differences (step 1e-4, relative in delta_rp, absolute in alpha) of re-converged solves, within
  1e-4 relative, on a tube with a curved Maxwell
  wall at alpha = 0.8, 0.9 and 1.0 (1.0 through the wall's smooth extension: the case is a Maxwell
  wall set to alpha = 1).
"""

from __future__ import annotations

import pytest
import torch

pytestmark = pytest.mark.gpu


def _tube(gsk, meshgen, d, alpha, delta=5.0, n=(6, 3, 6), n_vel=10):
    path = meshgen.tube_case(str(d), n_core=n[0], n_ring=n[1], n_axial=n[2], n_vel=n_vel, vrange=5.0, delta=delta,
                             accommodation=min(alpha, 0.99), kinetic={"precision": "fp64"})
    return gsk.Setup(path)


def test_frozen_branches_reproduce_residual_and_give_linear_products(gsk, meshgen, tmp_path):
    import gsis_sim_amd.sensitivity as sens

    print("[gpu] frozen limiter branches: reproduction and linearity", flush=True)
    setup = _tube(gsk, meshgen, tmp_path, 0.8)
    sv = gsk.Solver(setup)
    prob = sens.SteadyProblem(sv, inlet=0, mdot_factor=setup.mass_flow_factor(), wall_patch=2, delta_rp=5.0, alpha=0.8)
    sv.set_newton_tolerances(0.0, 0.0)
    sv.set_boundary_iterations(10)

    def rel(a, b):
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

    def state():
        f = prob._empty()
        torch.cuda.synchronize()
        sv.distribution_device(f.data_ptr(), prob.n)
        return f

    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    for _ in range(10):  # an unconverged state: R far from 0
        sv.prepare()
        sv.sweep()
    f10 = state()
    plain1, _ = prob._residual(f10)
    plain2, _ = prob._residual(f10)
    sv.freeze_limiters(True)
    recorded, _ = prob._residual(f10)
    followed, _ = prob._residual(f10)
    sv.freeze_limiters(False)
    noise = max(rel(plain2, plain1), 1e-14)
    print(f"  |R| {float(torch.linalg.vector_norm(plain2)):.3e}: plain repeat {rel(plain2, plain1):.2e}, recorded "
          f"{rel(recorded, plain2):.2e}, followed {rel(followed, recorded):.2e}", flush=True)
    assert rel(recorded, plain2) <= 10 * noise and rel(followed, recorded) <= 10 * noise

    prob.solve(residual_target=1e-9)
    f0 = state()
    sv.freeze_limiters(True)
    prob._capture()  # records at the converged state
    g = torch.Generator(device="cuda").manual_seed(3)
    z1 = torch.randn(prob.n, dtype=torch.float64, device="cuda", generator=g) * f0
    z2 = torch.randn(prob.n, dtype=torch.float64, device="cuda", generator=g) * f0

    def products(eps):
        prob.eps = eps
        j1, _ = prob._jv(z1)
        j2, _ = prob._jv(z2)
        j12, _ = prob._jv(2.0 * z1 - 3.0 * z2)
        return j1, j2, j12

    j1, j2, j12 = products(1e-5)
    j1s, _, _ = products(1e-6)
    frozen = {"linearity": rel(j12, 2.0 * j1 - 3.0 * j2), "step": rel(j1s, j1)}
    sv.freeze_limiters(False)
    j1p, j2p, j12p = products(1e-5)
    j1ps, _, _ = products(1e-6)
    plain_errs = {"linearity": rel(j12p, 2.0 * j1p - 3.0 * j2p), "step": rel(j1ps, j1p)}
    print(f"  frozen branches {frozen}; plain residual {plain_errs}", flush=True)
    assert frozen["linearity"] <= 1e-7 and frozen["step"] <= 1e-7
    sv.close()


@pytest.mark.parametrize("alpha", [0.8, 0.9, 1.0])
def test_sensitivities_match_central_differences(gsk, meshgen, tmp_path, alpha):
    import gsis_sim_amd.sensitivity as sens

    print(f"[gpu] sensitivities vs central differences, Maxwell tube, alpha {alpha}", flush=True)
    delta = 5.0
    setup = _tube(gsk, meshgen, tmp_path, alpha, delta)
    sv = gsk.Solver(setup)
    prob = sens.SteadyProblem(sv, inlet=0, mdot_factor=setup.mass_flow_factor(), wall_patch=2, delta_rp=delta,
                              alpha=alpha)
    d_t = torch.tensor(delta, dtype=torch.float64, requires_grad=True)
    a_t = torch.tensor(alpha, dtype=torch.float64, requires_grad=True)
    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    mdot = sens.converged_mdot(d_t, a_t, prob, residual_target=1e-12)
    prob.polish = 60
    prob.method = "tangent"  # one tangent solve per parameter (the adjoint has its own tests)
    mdot.backward()
    g_delta, g_alpha = float(d_t.grad), float(a_t.grad)
    print(f"  Mdot {float(mdot.detach()):.10f} dMdot/ddelta {g_delta:.8e} dMdot/dalpha {g_alpha:.8e} {prob.stats}",
          flush=True)

    def mdot_at(dl, al):
        prob.set_parameters(dl, al)
        return prob.solve(residual_target=1e-12)

    h = 1e-4
    fd_delta = (mdot_at(delta * (1 + h), alpha) - mdot_at(delta * (1 - h), alpha)) / (2 * delta * h)
    fd_alpha = (mdot_at(delta, alpha + h) - mdot_at(delta, alpha - h)) / (2 * h)
    e_delta, e_alpha = abs(g_delta - fd_delta) / abs(fd_delta), abs(g_alpha - fd_alpha) / abs(fd_alpha)
    print(f"  central differences: ddelta {fd_delta:.8e} ({e_delta:.1e}) dalpha {fd_alpha:.8e} ({e_alpha:.1e})",
          flush=True)
    assert e_delta <= 1e-4
    assert e_alpha <= 1e-4
    assert prob.stats["alpha_true_relative_residual"] <= 1e-6
    sv.close()
