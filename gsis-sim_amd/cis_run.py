"""CIS-only run loop over the kinetic solver (the `gsis-solver cis-run` loop,
src/app/cmd_cis_run.cpp:280-332 of the reference): prepare -> convergence check (eq. 59 epsilon
below the tolerance, after iteration 0) -> sweep, recording per iteration the residual, epsilon,
the inlet / outlet mass flows (nondimensional Mdot, cmd_cis_run.cpp:451-455) and the wall time.

Also stops on a residual target (the benchmark's "converge to residual 1e-6").
"""

from __future__ import annotations

import time

import numpy as np


def patch_index(setup_toml_patches, name):
    return setup_toml_patches.index(name)


def run(solver, mdot_factor: float, inlet: int | None, outlet: int | None, tol: float = 0.0,
        residual_target: float = 0.0, max_iterations: int = 1000, init=(1.0, (0.0, 0.0, 0.0), 1.0),
        log=None):
    """Run CIS iterations; returns a dict of arrays (iteration, epsilon, residual, mdot_in, mdot_out,
    t_wall) and the stop reason. mdot_factor = gsk_setup_mass_flow_scale (nondimensional patch flow ->
    Mdot); inlet flows are negative (outward positive) so Mdot_in = -flow * factor."""
    rho, u, T = init
    solver.initialize_uniform(rho, u, T)
    hist = {k: [] for k in ("iteration", "epsilon", "residual", "mdot_in", "mdot_out", "t_wall")}
    t0 = time.perf_counter()
    reason = "max_iterations"
    for it in range(max_iterations + 1):
        info = solver.prepare()
        flows = solver.patch_mass_flow()
        if it > 0 and tol > 0.0 and info["has_epsilon"] and info["epsilon"] < tol:
            reason = "epsilon"
            break
        if it == max_iterations:
            break
        info = solver.sweep()
        hist["iteration"].append(it)
        hist["epsilon"].append(info["epsilon"] if info["has_epsilon"] else np.nan)
        hist["residual"].append(info["residual"])
        hist["mdot_in"].append(-flows[inlet] * mdot_factor if inlet is not None else np.nan)
        hist["mdot_out"].append(flows[outlet] * mdot_factor if outlet is not None else np.nan)
        hist["t_wall"].append(time.perf_counter() - t0)
        if log is not None and (it % 10 == 0):
            log(f"iteration {it}: residual {info['residual']:.3e} Mdot_in {hist['mdot_in'][-1]:.6f}")
        if residual_target > 0.0 and info["residual"] <= residual_target:
            reason = "residual"
            break
    out = {k: np.array(v) for k, v in hist.items()}
    out["reason"] = reason
    return out


def time_to_within(hist: dict, target: float, rel: float = 1e-4, key: str = "mdot_in"):
    """Wall time after which |Mdot - target| <= rel |target| holds for every later iteration
    (None if never)."""
    x = hist[key]
    ok = np.abs(x - target) <= rel * abs(target)
    if not ok[-1]:
        return None
    bad = np.nonzero(~ok)[0]
    first = 0 if bad.size == 0 else bad[-1] + 1
    return float(hist["t_wall"][first]), int(hist["iteration"][first])
