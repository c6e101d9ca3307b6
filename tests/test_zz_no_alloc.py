"""The GSIS correction entry point that allocates no device memory
(gsk_backend_gsis_correction_noalloc) against the allocating one (gsk_backend_gsis_correction):
the same kernel, so f, the replaced flags and the replaced count must agree bit for bit. Two calls
in a row, the second on a smaller case (the fixed scratch keeps the first call's contents beyond
the second's size), then a case above the scratch's cell limit, which must be refused."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import box_case

pytestmark = pytest.mark.gpu


def _corrected(backend, f0, w_half, w_new, noalloc):
    backend.set_distribution(f0)
    if noalloc:
        rep, flags = backend.gsis_correction_noalloc(w_half, w_new)
    else:
        rep, flags = backend.gsis_correction(w_half, w_new)
    return rep, flags, backend.get_distribution()


def _states(backend, seed, n_cells):
    """w_half = the cells' own (rho, u, T); w_new moved from it, strongly in a quarter of the cells
    so that the negative-value rule replaces some of them."""
    m = backend.compute_moments()
    w_half = np.stack([m[:, 0], m[:, 1], m[:, 2], m[:, 3], m[:, 4]], axis=1)
    rng = np.random.default_rng(seed)
    w_new = w_half.copy()
    w_new[:, 0] *= 1.0 + 0.02 * rng.standard_normal(n_cells)
    w_new[:, 1:4] += 0.02 * rng.standard_normal((n_cells, 3))
    w_new[:, 4] *= 1.0 + 0.02 * rng.standard_normal(n_cells)
    strong = rng.random(n_cells) < 0.25
    w_new[strong, 4] *= 0.3
    w_new[strong, 1] += 0.8
    return w_half, w_new


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_gsis_correction_matches_allocating_entry_point(gsk, meshgen, tmp_path, precision):
    for call, (n, n_vel) in enumerate([((6, 4, 3), 10), ((3, 2, 2), 8)]):  # second call: smaller input
        print(f"noalloc case: call {call}, box {n}, {n_vel}^3 velocities, {precision}", flush=True)
        d = tmp_path / f"c{call}"
        d.mkdir()
        setup = gsk.Setup(box_case(meshgen, str(d), n=n, n_vel=n_vel, kinetic={"precision": precision}))
        sv = gsk.Solver(setup)
        sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
        for _ in range(2):
            sv.prepare()
            sv.sweep()
        be = sv.backend()
        f0 = be.get_distribution()
        w_half, w_new = _states(be, 11 + call, setup.n_cells)
        rep_a, flags_a, f_a = _corrected(be, f0, w_half, w_new, noalloc=False)
        rep_b, flags_b, f_b = _corrected(be, f0, w_half, w_new, noalloc=True)
        assert rep_a == rep_b
        assert rep_a == int(np.count_nonzero(flags_a))
        assert 0 < rep_a < setup.n_cells  # both branches of the replacement rule ran
        np.testing.assert_array_equal(flags_a, flags_b)
        np.testing.assert_array_equal(f_a, f_b)
        assert not np.array_equal(f_a, f0)
        sv.close()


def test_noalloc_gsis_correction_refuses_cases_above_its_scratch(gsk, meshgen, tmp_path):
    nx = gsk.NOALLOC_MAX_CELLS + 1
    print(f"noalloc limit case: {nx}-cell slab", flush=True)
    path = meshgen.slab_case(str(tmp_path), nx=nx, n_vel=2, vrange=4.0)
    setup = gsk.Setup(path, {"precision": "fp64"})
    assert setup.n_cells == nx
    be = gsk.Backend(setup)
    be.set_equilibrium(np.tile([1.0, 0.0, 0.0, 0.0, 1.0], (nx, 1)))
    w = np.tile([1.0, 0.0, 0.0, 0.0, 1.0], (nx, 1))
    with pytest.raises(gsk.GskError, match="fixed scratch"):
        be.gsis_correction_noalloc(w, w)
    # the allocating entry point still takes it (on this 2^3 grid every cell takes the
    # discrete-equilibrium fallback, flag 2, so every cell counts as replaced)
    rep, flags = be.gsis_correction(w, w)
    assert rep == int(np.count_nonzero(flags))
    be.close()
