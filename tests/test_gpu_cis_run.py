"""The cis-run driver over the HIP backend against the same driver over the FP64 oracle: history.csv
and summary.json of a few CIS iterations on the mixed-boundary box agree to the CIS-iteration bar
(1e-10 relative) in every physics column; timings, memory and the backend name are the only
columns allowed to differ."""

from __future__ import annotations

import csv
import json
import os
import sys

import numpy as np
import pytest

from conftest import box_case

pytestmark = pytest.mark.gpu

SKIP_COLUMNS = ("time_", "t_", "gpu_memory_bytes")
SKIP_KEYS = {"program", "wall_time_s", "setup_time_s", "time_per_iteration_s", "kernel_time_per_iteration_s", "backend",
             "memory_bytes", "case_file", "mesh", "velocity_chunk"}


def _history(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return rows[0], rows[1:]


def _close(a, b, what, scale=None):
    """Equal structure; reals within 1e-10 of `scale` (default: the value's own magnitude)."""
    if isinstance(a, dict):
        assert list(a) == list(b), what
        for k in a:
            if k not in SKIP_KEYS:
                _close(a[k], b[k], f"{what}.{k}", scale)
    elif isinstance(a, list):
        assert len(a) == len(b), what
        for i, (x, y) in enumerate(zip(a, b)):
            _close(x, y, f"{what}[{i}]", scale)
    elif isinstance(a, float) and isinstance(b, float):
        s = scale if scale is not None else abs(b)
        assert abs(a - b) <= 1e-10 * max(s, 1e-300), (what, a, b)
    else:
        assert a == b, (what, a, b)


def _flat(v):
    return [abs(x) for x in (v if isinstance(v, list) else [v]) if isinstance(x, float)]


def _close_patches(a, b):
    """Each patch field against the largest magnitude that field takes over all patches (a flow
    that is zero up to round-off on one patch is compared on the scale of the others)."""
    assert [p["name"] for p in a] == [p["name"] for p in b]
    for pa, pb in zip(a, b):
        assert list(pa) == list(pb), pb["name"]
    keys = list(dict.fromkeys(k for p in b for k in p))
    for k in keys:
        scale = max((x for p in b for x in _flat(p.get(k))), default=0.0)
        for pa, pb in zip(a, b):
            if k in pb:
                _close(pa[k], pb[k], f"patches.{pb['name']}.{k}", scale)


def test_cis_run_files_match_oracle(gsk, meshgen, oracle_mod, tmp_path):
    print("[gpu] cis-run files, HIP vs oracle, box_mixed 6 iterations", flush=True)
    import gsis_sim_amd.cis_run  # noqa: F401
    cr = sys.modules["gsis_sim_amd.cis_run"]
    case = box_case(meshgen, str(tmp_path), kinetic={"precision": "fp64"})
    with open(case, "a") as f:  # section flows in both files (src/mesh/section_cut.hpp)
        f.write('\n[output]\nmass_flow_sections = "0.2, 0.5 0.8"\n')
    cpu = cr.cis_run(case, str(tmp_path / "cpu"), max_iter=6,
                     backend_factory=lambda st: oracle_mod.CpuBackend(st, "oracle", ops_type=gsk.BackendOps))
    gpu = cr.cis_run(case, str(tmp_path / "gpu"), max_iter=6)
    assert gpu["backend"] != cpu["backend"]
    _close_patches(gpu.pop("patches"), cpu.pop("patches"))
    _close(gpu, cpu, "summary")
    hc, rc = _history(str(tmp_path / "cpu" / "history.csv"))
    hg, rg = _history(str(tmp_path / "gpu" / "history.csv"))
    assert hg == hc and len(rg) == len(rc) == 7
    def family(name):  # columns of one quantity share a scale (a component can be zero up to round-off)
        if name.startswith("mdot_"):
            return "mdot_kg_s" if name.endswith("_kg_s") else "mdot_nd"
        if name in ("qx_mean", "qy_mean", "qz_mean"):
            return "q"
        if name.startswith("s") and name.endswith("_mean"):
            return "sigma"
        return name

    cols = {}
    for j, name in enumerate(hc):
        if not name.startswith(SKIP_COLUMNS):
            cols[name] = (np.array([float(r[j]) for r in rg]), np.array([float(r[j]) for r in rc]))
    scales = {}
    for name, (_, b) in cols.items():
        f = family(name)
        scales[f] = max(scales.get(f, 0.0), np.max(np.abs(b[~np.isnan(b)]), initial=0.0))
    for name, (a, b) in cols.items():
        assert np.array_equal(np.isnan(a), np.isnan(b)), name
        ok = ~np.isnan(b)
        assert np.max(np.abs(a[ok] - b[ok]), initial=0.0) <= 1e-10 * max(scales[family(name)], 1e-300), name
    # the script form writes the same files
    out = str(tmp_path / "script")
    assert cr.main([case, "--out", out, "--max-iter", "2", "--print-every", "0"]) == 0
    with open(os.path.join(out, "summary.json")) as f:
        assert json.load(f)["iterations"] == 2
