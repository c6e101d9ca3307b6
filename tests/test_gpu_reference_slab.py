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

CASES = sorted(name for name, r in slab_reference_rows().items() if float(r["Kn"]) >= 0.1)
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
