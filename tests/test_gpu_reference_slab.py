"""The HIP forms on the reference's own slab verification runs (tests/golden/slab_reference_results.csv;
see tests/test_reference_slab.py): heat transfer and Couette flow, linear and nonlinear, Kn 0.1, 1
and 10, 50 / 100 / 200 cells, every side periodic, run by cis-run to epsilon < 2e-7 and reduced to
the volume-weighted mean P_xy or q_x as the reference's compare_ref1d.py does.

Bounds: FP64 within 2e-6 of the reference's value (measured up to 1.4e-6: the reference ran FP32
and the two runs stop at different iterates); FP32 (the reference's own precision) and the mixed
form (FP32 with matrix-core moment sums, nx = 100) within 5e-6. FP32 epsilon can sit at its
round-off floor near 2e-7, so those runs stop at 3x the reference's iteration count + 20."""

from __future__ import annotations

import pytest

from conftest import slab_reference_rows
from test_reference_slab import run_case

pytestmark = pytest.mark.gpu

CASES = sorted(name for name, r in slab_reference_rows().items() if float(r["Kn"]) >= 0.1 and r["sides"] == "periodic")
BOUND = {"fp64": 2e-6, "fp32": 5e-6, "mixed": 5e-6}


@pytest.mark.parametrize("case", CASES)
def test_hip_forms_reach_the_reference_slab_result(gsk, meshgen, tmp_path, case):
    ref = slab_reference_rows()[case]
    want = float(ref["value"])
    forms = ["fp64", "fp32"] + (["mixed"] if ref["nx"] == "100" else [])
    for form in forms:
        print(f"[gpu] reference slab {case} {form}", flush=True)
        val, s = run_case(gsk, meshgen, tmp_path / form, ref, form)
        assert abs(val - want) <= BOUND[form] * abs(want), (form, val, want, s["iterations"])


DEFAULT_SIDES = sorted(name for name, r in slab_reference_rows().items() if float(r["Kn"]) >= 0.1 and r["sides"] == "default")


@pytest.mark.parametrize("case", DEFAULT_SIDES)
def test_symmetry_sides_converge_to_the_reference_periodic_result(gsk, meshgen, tmp_path, case):
    """The reference's default-side runs (symmetry planes; Couette: sym_y periodic) were stopped at
    epsilon < 1e-6 by a build whose symmetry planes converged more slowly (its published iteration
    counts are about twice those of its current kernels, which this library's solver reproduces),
    so their values still carry up to 6e-4 of truncation error. The converged solution is the same
    as with periodic sides (cases/verification/kinetic/make_cases.py): FP64 to epsilon < 1e-8 with
    symmetry planes lands within 3e-6 of the reference's periodic result (measured up to 2.0e-6)."""
    import sys

    import gsis_sim_amd.cis_run  # noqa: F401
    from conftest import slab_verification_case
    from test_reference_slab import mean_flux_from_cells

    cr = sys.modules["gsis_sim_amd.cis_run"]
    rows = slab_reference_rows()
    ref, per = rows[case], rows[case + "_per"]
    print(f"[gpu] reference slab {case} symmetry sides, FP64 to 1e-8", flush=True)
    path = slab_verification_case(meshgen, str(tmp_path), ref["kind"], float(ref["Kn"]), int(ref["nx"]), "fp64", "default")
    s = cr.cis_run(path, str(tmp_path / "run"), tol=1e-8, max_iter=3000, cells_csv=True)
    assert s["converged"]
    val = mean_flux_from_cells(str(tmp_path / "run" / "cells.csv"), ref["kind"])
    want = float(per["value"])
    assert abs(val - want) <= 3e-6 * abs(want), (val, want, s["iterations"])
