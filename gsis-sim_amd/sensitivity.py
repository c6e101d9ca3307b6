"""Sensitivities of the converged inlet mass flow Mdot to the rarefaction parameter delta_rp and to
a Maxwell wall's accommodation alpha, behind a torch.autograd.Function (no counterpart in the
reference, which has no derivative capability).

At the converged state f* the steady residual vanishes, R(f*; p) = 0, so with J = dR/df at f*
    adjoint (default):  J^T nu = (dMdot/df)^T,  dMdot/dp = dMdot/dp|_f - nu . dR/dp   (one solve)
    tangent:            J s = -dR/dp,           dMdot/dp = dMdot/dp|_f + (dMdot/df) s  (per parameter)
R is the solver's full steady residual (transport with the limited reconstruction, Shakhov
collision with the discrete equilibrium, every boundary condition, wall densities and inflow states
recomputed from f), evaluated on the device through gsk_solver_residual_device. The adjoint applies
the hand-written transposed Jacobian (gsk_solver_adjoint_apply_device, which also gives
(dMdot/df)^T) and is preconditioned by LU-SGS of the reversed velocities (the transposed first-order
transport of xi is the transport of -xi); the tangent solve takes difference Jacobian-vector
products and LU-SGS (gsk_solver_precondition_device, the level-scheduled kernels of the forward
solve). Both solve by BiCGStab, then restarted GMRES on the true residual.

Jacobian-vector products. R is only piecewise smooth in f: every limited gradient takes the largest
and the most negative face projection, the extreme neighbour values and the smaller of two limiter
values, clamped to [0, 1], and the pressure-inlet reconstruction has its own limiter. A difference
product that crosses one of those branches is not linear in the direction (with differences of the
plain residual the true linear residual of the Krylov solve stayed at 1-4% and the sensitivities
depended on the step). So at f* every branch is recorded (gsk_solver_freeze_limiters), and every
evaluation of the solve follows the record: R is then smooth around f*, equal to R on f*'s branch,
and central differences J z = (R(f* + h z) - R(f* - h z)) / 2h give that branch's Jacobian products to
O(h^2) (h = 1e-5 |f*| / |z| by default; the rounding part is ~1e-11 of the product). The solve
further sets the discrete-equilibrium Newton tolerances to 0 (iterate until the step stops improving,
so R does not jump with the iteration count) and runs 10 inflow / boundary-sum rounds per prepare
(the pressure-inlet state depends on the previous round's face states), and restores the case's
settings afterwards.

The adjoint transposes R on the same recorded branches, so the two forms differentiate the same
function (they agree to ~1e-9). dR/dp is a central difference in the parameter in both. delta_rp
enters only through the nondimensional viscosity scale mu0 = 1 / (sqrt(2) delta_rp); alpha only
through the specular fraction 1 - alpha of the Maxwell-wall faces.
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch


class SteadyProblem:
    """A converged forward solve and the device-resident operators around it.

    solver: gsis_sim_amd.Solver (HIP backend); inlet: patch index of the inlet; mdot_factor:
    Setup.mass_flow_factor(); wall_patch: the Maxwell-wall patch whose accommodation is a
    parameter (None if alpha is not needed)."""

    def __init__(self, solver, inlet: int, mdot_factor: float, wall_patch: int | None = None,
                 delta_rp: float | None = None, alpha: float | None = None, eps: float = 1e-5):
        self.sv = solver
        self.inlet = inlet
        self.factor = mdot_factor
        self.wall_patch = wall_patch
        self.n = solver.setup.n_cells * solver.setup.nv
        self.dtype = torch.float64 if solver.element_size() == 8 else torch.float32
        self.delta_rp = delta_rp if delta_rp is not None else 1.0 / (math.sqrt(2.0) * solver.mu0())
        self.alpha = alpha
        self.eps = eps
        # the case's own settings, restored after a sensitivity solve changes them
        params = solver.setup.describe()["params"]
        self.boundary_iterations = int(params["boundary_iterations"])
        self.newton_tolerances = (solver.setup.view().deq_tolerance, 1e-10)
        self.log = None  # optional progress callback (message) for long solves
        self.forward_solver = None  # set by use_forward_solver: the solver the forward solve runs on
        self.f0 = None
        self.R0 = None
        self.mdot0 = None
        self.stats = {}

    # -- forward --------------------------------------------------------------------------------
    def set_parameters(self, delta_rp: float | None = None, alpha: float | None = None) -> None:
        if delta_rp is not None:
            self.delta_rp = float(delta_rp)
            self.sv.set_mu0(1.0 / (math.sqrt(2.0) * self.delta_rp))
        if alpha is not None and self.wall_patch is not None:
            self.alpha = float(alpha)
            self.sv.set_accommodation(self.wall_patch, self.alpha)

    def mdot(self) -> float:
        return -self.sv.patch_mass_flow()[self.inlet] * self.factor

    def use_forward_solver(self, solver) -> None:
        """Run the forward solve on another solver of the same case, typically the FP32 / mixed form
        (FP32 f): forward() converges it and hands its f to this problem's FP64 solver, which then
        carries the sensitivity solve at that state. The FP32 state solves the FP64 equations to the
        forward residual target, so the sensitivities differ from those at the FP64 solution by
        about that much relative to their size."""
        self.forward_solver = solver

    def forward(self, residual_target: float, max_iterations: int = 5000) -> float:
        """The converged Mdot: on this problem's solver, or on the forward solver (then copied over)."""
        if self.forward_solver is None:
            return self.solve(residual_target=residual_target, max_iterations=max_iterations)
        other = self.forward_solver
        if self.delta_rp is not None:
            other.set_mu0(1.0 / (math.sqrt(2.0) * self.delta_rp))
        if self.alpha is not None and self.wall_patch is not None:
            other.set_accommodation(self.wall_patch, self.alpha)
        t0 = time.perf_counter()
        it = 0
        for it in range(max_iterations):
            other.prepare()
            info = other.sweep()
            if self.log is not None and it % 25 == 0:
                self.log(f"forward (element size {other.element_size()}) iteration {it}: residual {info['residual']:.3e}")
            if info["residual"] <= residual_target:
                break
        other.prepare()
        mdot = -other.patch_mass_flow()[self.inlet] * self.factor
        self.sv.set_distribution(other.distribution())  # original numbering, grid order, as float64
        self.sv.prepare()
        self.stats["forward_seconds"] = time.perf_counter() - t0
        self.stats["forward_iterations"] = it + 1
        self.stats["forward_mdot"] = mdot
        self.f0 = None
        return mdot

    def solve(self, residual_target: float = 1e-10, max_iterations: int = 5000, init=None, log=None) -> float:
        """CIS iterations until the residual norm is below residual_target; returns Mdot."""
        log = log or self.log
        if init is not None:
            self.sv.initialize_uniform(*init)
        t0 = time.perf_counter()
        it = 0
        for it in range(max_iterations):
            self.sv.prepare()
            info = self.sv.sweep()
            if log is not None and it % 25 == 0:
                log(f"iteration {it}: residual {info['residual']:.3e}")
            if info["residual"] <= residual_target:
                break
        self.sv.prepare()
        self.stats["forward_seconds"] = time.perf_counter() - t0
        self.stats["forward_iterations"] = it + 1
        self.f0 = None
        return self.mdot()

    # -- device operators -----------------------------------------------------------------------
    def _empty(self) -> torch.Tensor:
        return torch.empty(self.n, dtype=self.dtype, device="cuda")

    def _set_f(self, f: torch.Tensor) -> None:
        torch.cuda.synchronize()
        self.sv.set_distribution_device(f.data_ptr(), self.n)

    def _residual(self, f: torch.Tensor) -> tuple[torch.Tensor, float]:
        self._set_f(f)
        out = self._empty()
        self.sv.residual_device(out.data_ptr(), self.n)
        return out, self.mdot()

    def _capture(self) -> None:
        self.f0 = self._empty()
        torch.cuda.synchronize()
        self.sv.distribution_device(self.f0.data_ptr(), self.n)
        self.R0, self.mdot0 = self._residual(self.f0)
        self.f0_norm = float(torch.linalg.vector_norm(self.f0))

    def _jv(self, z: torch.Tensor) -> tuple[torch.Tensor, float]:
        """J z and (dMdot/df) z by central differences of R and Mdot at f*, on f*'s limiter branches."""
        zn = float(torch.linalg.vector_norm(z))
        if zn == 0.0:
            return torch.zeros_like(z), 0.0
        eps = self.eps * self.f0_norm / zn
        Rp, mp = self._residual(self.f0 + eps * z)
        Rm, mm = self._residual(self.f0 - eps * z)
        return (Rp - Rm) / (2.0 * eps), (mp - mm) / (2.0 * eps)

    def _precondition(self, r: torch.Tensor) -> torch.Tensor:
        """~ J^-1 r: the LU-SGS passes return -(D + L)^-1 D (D + U)^-1 r."""
        out = self._empty()
        torch.cuda.synchronize()
        self.sv.precondition_device(r.data_ptr(), out.data_ptr(), self.n)
        return -out

    def _param_derivative(self, which: str) -> tuple[torch.Tensor, float]:
        """dR/dp and dMdot/dp at fixed f* by a central difference in the parameter."""
        if which == "delta_rp":
            mu0 = self.sv.mu0()
            h = self.eps * mu0
            self.sv.set_mu0(mu0 + h)
            Rp, mp = self._residual(self.f0)
            self.sv.set_mu0(mu0 - h)
            Rm, mm = self._residual(self.f0)
            self.sv.set_mu0(mu0)
            dmu_ddelta = -mu0 / self.delta_rp
            return (Rp - Rm) / (2.0 * h) * dmu_ddelta, (mp - mm) / (2.0 * h) * dmu_ddelta
        if which == "alpha":
            h = self.eps
            self.sv.set_accommodation(self.wall_patch, self.alpha + h)
            Rp, mp = self._residual(self.f0)
            self.sv.set_accommodation(self.wall_patch, self.alpha - h)
            Rm, mm = self._residual(self.f0)
            self.sv.set_accommodation(self.wall_patch, self.alpha)
            return (Rp - Rm) / (2.0 * h), (mp - mm) / (2.0 * h)
        raise ValueError(which)

    # -- adjoint operators ----------------------------------------------------------------------
    def _reversed_velocities(self) -> torch.Tensor:
        """Storage position of -xi for every velocity position. The first-order transport of velocity
        xi, transposed, is the transport of -xi (each face coefficient xi . S changes sign, and a
        closed cell's outflow sum equals its inflow sum), so LU-SGS applied to the velocity-reversed
        vector, reversed back, is to J^T what LU-SGS is to J: the same level-scheduled sweeps, run
        for the opposite direction, i.e. in the reversed level order."""
        if getattr(self, "_reverse", None) is None:
            v = self.sv.setup.view()
            nv = self.sv.setup.nv
            xi = np.stack([np.ctypeslib.as_array(a, shape=(nv,)) for a in (v.xi_x, v.xi_y, v.xi_z)], axis=1)
            scale = 1e9 / max(float(np.abs(xi).max()), 1e-300)
            keys = {tuple(k): p for p, k in enumerate(np.rint(xi * scale).astype(np.int64))}
            rev = [keys.get(tuple(k)) for k in np.rint(-xi * scale).astype(np.int64)]
            if any(r is None for r in rev):
                raise ValueError("the velocity grid is not symmetric under xi -> -xi: the adjoint solve needs one")
            self._reverse = torch.tensor(rev, dtype=torch.int64, device="cuda")
        return self._reverse

    def _reverse_velocities(self, x: torch.Tensor) -> torch.Tensor:
        nc = self.sv.setup.n_cells
        return x.view(nc, -1)[:, self._reversed_velocities()].reshape(-1)

    def _precondition_t(self, r: torch.Tensor) -> torch.Tensor:
        """~ J^-T r: LU-SGS on the velocity-reversed r, reversed back."""
        return self._reverse_velocities(self._precondition(self._reverse_velocities(r).contiguous()))

    def _jtv(self, lam: torch.Tensor) -> torch.Tensor:
        """J^T lam at the prepared state (gsk_solver_adjoint_apply_device)."""
        out = self._empty()
        torch.cuda.synchronize()
        self.sv.adjoint_apply_device(lam.data_ptr(), [0.0] * self.sv.setup.n_patches, out.data_ptr(), self.n)
        return out

    def _mdot_gradient(self) -> torch.Tensor:
        """(dMdot/df)^T at the prepared state: the transposed boundary model with the inlet's weight."""
        w = [0.0] * self.sv.setup.n_patches
        w[self.inlet] = -self.factor
        out = self._empty()
        torch.cuda.synchronize()
        self.sv.adjoint_apply_device(0, w, out.data_ptr(), self.n)
        return out

    def _gmres(self, b: torch.Tensor, tol: float, restart: int, max_iterations: int, op=None, prec=None):
        """Right-preconditioned restarted GMRES for A s = b (A = J and LU-SGS unless op / prec are
        given); returns s, iterations, relative residual."""
        op = op or (lambda z: self._jv(z)[0])
        prec = prec or self._precondition
        bnorm = float(torch.linalg.vector_norm(b))
        s = torch.zeros_like(b)
        if bnorm == 0.0:
            return s, 0, 0.0
        total = 0
        rel = 1.0
        r = b.clone()
        while total < max_iterations:
            beta = float(torch.linalg.vector_norm(r))
            rel = beta / bnorm
            if rel <= tol:
                break
            m = min(restart, max_iterations - total)
            # the basis as separate vectors: each reuses a cached block of the solve's own vector size,
            # where one (m + 1) x n tensor would need a contiguous m + 1 vectors' worth (129 GiB at
            # the 40k-cell tube with m = 30) that a fragmented cache may not have
            torch.cuda.empty_cache()
            V = [r / beta]
            H = np.zeros((m + 1, m))
            cs, sn = np.zeros(m), np.zeros(m)
            g = np.zeros(m + 1)
            g[0] = beta
            k_used = 0
            for k in range(m):
                w = op(prec(V[k]))
                for _ in range(2):  # classical Gram-Schmidt, twice
                    h = torch.stack([torch.dot(V[i], w) for i in range(k + 1)]).double().cpu().numpy()
                    for i in range(k + 1):
                        w.add_(V[i], alpha=-float(h[i]))
                    H[: k + 1, k] += h
                H[k + 1, k] = float(torch.linalg.vector_norm(w))
                V.append(w / H[k + 1, k] if H[k + 1, k] > 0.0 else torch.zeros_like(w))
                for i in range(k):  # apply the previous Givens rotations
                    t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
                    H[i + 1, k] = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
                    H[i, k] = t
                d = math.hypot(H[k, k], H[k + 1, k])
                cs[k], sn[k] = (H[k, k] / d, H[k + 1, k] / d) if d > 0.0 else (1.0, 0.0)
                H[k, k] = d
                H[k + 1, k] = 0.0
                g[k + 1] = -sn[k] * g[k]
                g[k] = cs[k] * g[k]
                k_used = k + 1
                total += 1
                rel = abs(g[k + 1]) / bnorm
                if self.log is not None and total % 10 == 0:
                    self.log(f"GMRES iteration {total}: relative residual {rel:.2e}")
                if rel <= tol:
                    break
            y = np.linalg.solve(np.triu(H[:k_used, :k_used]), g[:k_used])
            z = torch.zeros_like(s)
            for i in range(k_used):
                z.add_(V[i], alpha=float(y[i]))
            s = s + prec(z)
            del V, z
            r = b - op(s)
        return s, total, rel

    def _bicgstab(self, b: torch.Tensor, tol: float, max_iterations: int, op=None, prec=None):
        """Right-preconditioned BiCGStab for A s = b (A = J and LU-SGS unless op / prec are given;
        short recurrences: a few vectors of state, two products with A and two preconditioner passes
        per iteration)."""
        op = op or (lambda z: self._jv(z)[0])
        prec = prec or self._precondition
        bnorm = float(torch.linalg.vector_norm(b))
        x = torch.zeros_like(b)
        if bnorm == 0.0:
            return x, 0, 0.0
        r = b.clone()
        rhat = b.clone()
        rho = alpha = omega = 1.0
        v = torch.zeros_like(b)
        p = torch.zeros_like(b)
        rel = 1.0
        it = 0
        for it in range(1, max_iterations + 1):
            rho_new = float(torch.dot(rhat, r))
            if rho_new == 0.0:
                break
            beta = (rho_new / rho) * (alpha / omega)
            p = r + beta * (p - omega * v)
            y = prec(p)
            v = op(y)
            alpha = rho_new / float(torch.dot(rhat, v))
            x = x + alpha * y
            r = r - alpha * v
            rel = float(torch.linalg.vector_norm(r)) / bnorm
            if rel <= tol:
                break
            z = prec(r)
            t = op(z)
            tt = float(torch.dot(t, t))
            omega = float(torch.dot(t, r)) / tt if tt > 0.0 else 0.0
            x = x + omega * z
            r = r - omega * t
            rel = float(torch.linalg.vector_norm(r)) / bnorm
            rho = rho_new
            if self.log is not None and it % 10 == 0:
                self.log(f"BiCGStab iteration {it}: relative residual {rel:.2e}")
            if rel <= tol or omega == 0.0:
                break
        return x, it, rel

    def sensitivities(self, params=("delta_rp", "alpha"), tol: float = 1e-8, restart: int = 30,
                      max_iterations: int = 400, method: str = "bicgstab", polish: int = 0) -> dict:
        """dMdot/dp at the converged state for each named parameter."""
        t0 = time.perf_counter()
        if "alpha" in params and self.wall_patch is None:
            raise ValueError("dMdot/dalpha needs the Maxwell-wall patch (wall_patch)")
        self.sv.set_newton_tolerances(0.0, 0.0)
        self.sv.set_boundary_iterations(10)
        out = {}
        try:
            self.sv.freeze_limiters(True)  # recorded by the residual evaluation at f* in _capture
            self._capture()
            for p in params:
                dR, dM = self._param_derivative(p)
                b = -dR
                bn = float(torch.linalg.vector_norm(b))
                if method == "gmres":
                    s, its, rel = self._gmres(b, tol, restart, max_iterations)
                else:
                    s, its, rel = self._bicgstab(b, tol, max_iterations)
                # BiCGStab's recurrence residual drifts from the true one b - J s on an ill-conditioned
                # system: polish with restarted GMRES on the true residual (its residual cannot grow)
                r_true = b - self._jv(s)[0]
                true_rel = float(torch.linalg.vector_norm(r_true)) / bn
                if method != "gmres" and polish > 0 and true_rel > tol:
                    d, k, _ = self._gmres(r_true, tol * bn / float(torch.linalg.vector_norm(r_true)), restart, polish)
                    s = s + d
                    its += k
                    true_rel = float(torch.linalg.vector_norm(b - self._jv(s)[0])) / bn
                if self.log is not None:
                    self.log(f"{p}: {its} Krylov iterations, true relative residual {true_rel:.2e}")
                _, dMs = self._jv(s)
                out[p] = dM + dMs
                self.stats[f"{p}_krylov_iterations"] = its
                self.stats[f"{p}_krylov_relative_residual"] = rel
                self.stats[f"{p}_true_relative_residual"] = true_rel
        finally:
            self.sv.freeze_limiters(False)
            self.sv.set_newton_tolerances(*self.newton_tolerances)
            self.sv.set_boundary_iterations(self.boundary_iterations)
            if self.f0 is not None:
                self._set_f(self.f0)  # leave the solver at f*
                self.sv.prepare()
        self.stats["sensitivity_seconds"] = time.perf_counter() - t0
        return out


    def adjoint_sensitivities(self, params=("delta_rp", "alpha"), tol: float = 1e-8, restart: int = 30,
                              max_iterations: int = 400, polish: int = 0) -> dict:
        """dMdot/dp at the converged state for each named parameter from one adjoint solve:
        J^T nu = (dMdot/df)^T, then dMdot/dp = dMdot/dp|_f - nu . dR/dp. J^T is the hand-written
        transposed Jacobian on the limiter branches of f* (gsk_solver_adjoint_apply_device), the
        solve BiCGStab right-preconditioned by LU-SGS of the reversed velocities, with optional
        restarted-GMRES polishing on the true residual."""
        t0 = time.perf_counter()
        if "alpha" in params and self.wall_patch is None:
            raise ValueError("dMdot/dalpha needs the Maxwell-wall patch (wall_patch)")
        self.sv.set_newton_tolerances(0.0, 0.0)
        self.sv.set_boundary_iterations(10)
        out = {}
        try:
            self.sv.freeze_limiters(True)  # recorded by the residual evaluation at f* in _capture
            self._capture()
            derivs = {p: self._param_derivative(p) for p in params}
            self._residual(self.f0)  # back at f* (cell parameters and boundary state of f*)
            self.sv.adjoint_prepare()
            g = self._mdot_gradient()
            gn = float(torch.linalg.vector_norm(g))
            nu, its, rel = self._bicgstab(g, tol, max_iterations, op=self._jtv, prec=self._precondition_t)
            torch.cuda.empty_cache()  # the BiCGStab vectors, before GMRES builds its basis
            r_true = g - self._jtv(nu)
            true_rel = float(torch.linalg.vector_norm(r_true)) / gn if gn > 0.0 else 0.0
            if polish > 0 and true_rel > tol:
                d, k, _ = self._gmres(r_true, tol * gn / float(torch.linalg.vector_norm(r_true)), restart, polish,
                                      op=self._jtv, prec=self._precondition_t)
                nu = nu + d
                its += k
                true_rel = float(torch.linalg.vector_norm(g - self._jtv(nu))) / gn
            if self.log is not None:
                self.log(f"adjoint: {its} Krylov iterations, true relative residual {true_rel:.2e}")
            for p, (dR, dM) in derivs.items():
                out[p] = dM - float(torch.dot(nu, dR))
            self.stats["adjoint_krylov_iterations"] = its
            self.stats["adjoint_krylov_relative_residual"] = rel
            self.stats["adjoint_true_relative_residual"] = true_rel
        finally:
            self.sv.freeze_limiters(False)
            self.sv.set_newton_tolerances(*self.newton_tolerances)
            self.sv.set_boundary_iterations(self.boundary_iterations)
            if self.f0 is not None:
                self._set_f(self.f0)  # leave the solver at f*
                self.sv.prepare()
        self.stats["sensitivity_seconds"] = time.perf_counter() - t0
        return out


class ConvergedMdot(torch.autograd.Function):
    """Mdot of the converged kinetic solve as a function of (delta_rp, alpha); backward returns the
    sensitivities from one adjoint solve (SteadyProblem.adjoint_sensitivities; problem.method =
    "tangent" uses one tangent solve per parameter instead, SteadyProblem.sensitivities)."""

    @staticmethod
    def forward(ctx, delta_rp: torch.Tensor, alpha: torch.Tensor, problem: SteadyProblem, residual_target: float):
        problem.set_parameters(float(delta_rp), float(alpha) if problem.wall_patch is not None else None)
        mdot = problem.forward(residual_target)
        ctx.problem = problem
        return torch.tensor(mdot, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        prob = ctx.problem
        params = ("delta_rp", "alpha") if prob.wall_patch is not None else ("delta_rp",)
        opts = dict(params=params, polish=getattr(prob, "polish", 0), restart=getattr(prob, "restart", 30))
        if getattr(prob, "method", "adjoint") == "tangent":
            sens = prob.sensitivities(**opts)
        else:
            sens = prob.adjoint_sensitivities(**opts)
        g = float(grad)
        return (torch.tensor(g * sens["delta_rp"], dtype=torch.float64),
                torch.tensor(g * sens.get("alpha", 0.0), dtype=torch.float64), None, None)


def converged_mdot(delta_rp: torch.Tensor, alpha: torch.Tensor, problem: SteadyProblem,
                   residual_target: float = 1e-10) -> torch.Tensor:
    return ConvergedMdot.apply(delta_rp, alpha, problem, residual_target)
