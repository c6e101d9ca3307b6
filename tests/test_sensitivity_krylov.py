"""The Krylov solvers of gsis-sim_amd/sensitivity.py (restarted GMRES and BiCGStab, both right-
preconditioned) on a small dense nonsymmetric system on the CPU, against numpy's direct solve. The
solvers only see the problem through _jv (J v) and _precondition (M^-1 v), so a stand-in supplies
those; the kinetic operators themselves are covered by tests/test_gpu_sensitivity.py."""

from __future__ import annotations

import numpy as np
import pytest
import torch


class _DenseProblem:
    """J = I + 0.4 A (A random, nonsymmetric); preconditioner = the inverse of J's diagonal."""

    def __init__(self, n: int = 60, seed: int = 3):
        rng = np.random.default_rng(seed)
        self.J = np.eye(n) + 0.4 * rng.standard_normal((n, n)) / np.sqrt(n)
        self.J[np.arange(n), np.arange(n)] *= rng.uniform(0.5, 2.0, n)
        self.n = n
        self.dtype = torch.float64
        self.log = None
        self._Jt = torch.as_tensor(self.J)
        self._dinv = torch.as_tensor(1.0 / np.diag(self.J))

    def _jv(self, z):
        return self._Jt @ z, 0.0

    def _precondition(self, r):
        return self._dinv * r


@pytest.mark.parametrize("restart", [5, 60])
def test_gmres_solves_dense_system(gsk, restart):
    import gsis_sim_amd.sensitivity as sens

    p = _DenseProblem()
    b = torch.as_tensor(np.random.default_rng(9).standard_normal(p.n))
    x, its, rel = sens.SteadyProblem._gmres(p, b, 1e-12, restart, 400)
    exact = np.linalg.solve(p.J, b.numpy())
    assert rel <= 1e-12
    assert np.max(np.abs(x.numpy() - exact)) <= 1e-9 * np.max(np.abs(exact))
    assert np.max(np.abs(p.J @ x.numpy() - b.numpy())) <= 1e-10 * np.max(np.abs(b.numpy()))


def test_bicgstab_solves_dense_system(gsk):
    import gsis_sim_amd.sensitivity as sens

    p = _DenseProblem()
    b = torch.as_tensor(np.random.default_rng(9).standard_normal(p.n))
    x, its, rel = sens.SteadyProblem._bicgstab(p, b, 1e-12, 400)
    exact = np.linalg.solve(p.J, b.numpy())
    assert rel <= 1e-12
    assert np.max(np.abs(x.numpy() - exact)) <= 1e-9 * np.max(np.abs(exact))
