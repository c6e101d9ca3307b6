"""Sensitivities of the converged inlet mass flow (gsis-sim_amd/sensitivity.py) against central
finite differences of re-converged solves, on a small tube with a curved Maxwell wall (FP64)."""

from __future__ import annotations

import pytest
import torch


@pytest.mark.gpu
def test_sensitivities_match_central_differences(gsk, meshgen, tmp_path):
    import gsis_sim_amd.sensitivity as sens

    print("[gpu] sensitivities vs central differences, small Maxwell tube", flush=True)
    delta, alpha = 5.0, 0.9
    path = meshgen.tube_case(str(tmp_path), n_core=4, n_ring=2, n_axial=3, n_vel=8, vrange=5.0, delta=delta,
                             accommodation=alpha, kinetic={"precision": "fp64"})
    setup = gsk.Setup(path)
    sv = gsk.Solver(setup)
    prob = sens.SteadyProblem(sv, inlet=0, mdot_factor=setup.mass_flow_factor(), wall_patch=2, delta_rp=delta,
                              alpha=alpha)
    d_t = torch.tensor(delta, dtype=torch.float64, requires_grad=True)
    a_t = torch.tensor(alpha, dtype=torch.float64, requires_grad=True)
    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    mdot = sens.converged_mdot(d_t, a_t, prob, residual_target=1e-12)
    mdot.backward()
    g_delta, g_alpha = float(d_t.grad), float(a_t.grad)
    print(f"  Mdot {float(mdot.detach()):.10f} dMdot/ddelta {g_delta:.8e} dMdot/dalpha {g_alpha:.8e} {prob.stats}", flush=True)

    def mdot_at(dl, al):
        prob.set_parameters(dl, al)
        return prob.solve(residual_target=1e-12)

    h = 1e-4
    fd_delta = (mdot_at(delta * (1 + h), alpha) - mdot_at(delta * (1 - h), alpha)) / (2 * delta * h)
    fd_alpha = (mdot_at(delta, alpha + h) - mdot_at(delta, alpha - h)) / (2 * h)
    print(f"  central differences: ddelta {fd_delta:.8e} dalpha {fd_alpha:.8e}", flush=True)
    assert abs(g_delta - fd_delta) <= 1e-4 * abs(fd_delta)
    assert abs(g_alpha - fd_alpha) <= 1e-4 * abs(fd_alpha)
    sv.close()
