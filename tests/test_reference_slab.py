"""Converged slab solutions against the reference's own results (tests/golden/slab_reference_results.csv,
taken from its cases/verification/kinetic/results/slab_comparison.csv): the reference ran
`gsis-solver cis-run` (FP32 CUDA backend) on heat-transfer and Couette slabs with periodic sides to
epsilon < 2e-7 and recorded the volume-weighted mean P_xy or q_x of cells.csv. The same cases, run by
this driver to the same tolerance and reduced the same way (compare_ref1d.py), must land on the
same value. On the CPU the FP64 oracle takes the two cheapest cases; the GPU suite runs the HIP
forms on more (tests/test_gpu_reference_slab.py). The bound, 5e-6 relative, is what separates two
converged solutions of one discretisation stopped at epsilon < 2e-7 (the runs differ in precision and
stop at different iterates); the 1D reference solution (ref_vmid30) differs from them by 1e-6 to 1e-4."""

from __future__ import annotations

import csv
import os
import sys

import numpy as np
import pytest

from conftest import slab_reference_rows, slab_verification_case

REL = 5e-6


def mean_flux_from_cells(path, kind):
    with open(path) as f:
        rows = list(csv.DictReader(f))
    vol = np.array([float(r["volume"]) for r in rows])
    q = np.array([float(r["sxy" if kind.startswith("couette") else "qx"]) for r in rows])
    return float(np.sum(vol / vol.sum() * q))


def run_case(gsk, meshgen, tmp_path, ref, precision, backend_factory=None):
    """cis-run to the reference's stop: epsilon < 2e-7 with periodic sides, < 1e-6 with the default
    sides. FP64 must get there; FP32 epsilon can sit at its round-off floor near 2e-7 (the
    reference's FP32 run stopped at 1.95e-7 on some cases), so the FP32 forms stop at 3x the
    reference's iterations + 20 at the latest."""
    import gsis_sim_amd.cis_run  # noqa: F401
    cr = sys.modules["gsis_sim_amd.cis_run"]
    kind, kn, nx = ref["kind"], float(ref["Kn"]), int(ref["nx"])
    os.makedirs(str(tmp_path), exist_ok=True)
    mixed = precision == "mixed"  # FP32 f with matrix-core moment sums
    case = slab_verification_case(meshgen, str(tmp_path), kind, kn, nx, "fp32" if mixed else precision, ref["sides"])
    out = str(tmp_path / f"run_{precision}")
    its = int(ref["iterations"])
    cap = max(2000, 2 * its) if precision == "fp64" else 3 * its + 20
    tol = 2e-7 if ref["sides"] == "periodic" else 1e-6
    s = cr.cis_run(case, out, tol=tol, max_iter=cap, cells_csv=True, backend_factory=backend_factory,
                   overrides={"moment_sums": "mfma"} if mixed else None)
    assert s["converged"] or precision != "fp64", ref["case"]
    return mean_flux_from_cells(os.path.join(out, "cells.csv"), kind), s


def test_fixture_covers_the_reference_runs():
    rows = slab_reference_rows()
    assert len(rows) == 64
    assert {r["kind"] for r in rows.values()} == {"heat_lin", "heat_nl", "couette_lin", "couette_nl"}
    assert sum(r["sides"] == "periodic" for r in rows.values()) == 48
    assert all(float(r["epsilon"]) < (2e-7 if r["sides"] == "periodic" else 1e-6) for r in rows.values())


@pytest.mark.parametrize("case", ["slab_couette_lin_Kn10_nx100_per", "slab_heat_lin_Kn1_nx100_per"])
def test_oracle_reaches_the_reference_slab_result(gsk, meshgen, oracle_mod, tmp_path, case):
    ref = slab_reference_rows()[case]
    val, s = run_case(gsk, meshgen, tmp_path, ref, "fp64",
                      lambda st: oracle_mod.CpuBackend(st, "oracle", ops_type=gsk.BackendOps))
    want = float(ref["value"])
    assert abs(val - want) <= REL * abs(want), (val, want)
    assert abs(s["iterations"] - int(ref["iterations"])) <= 3
