"""Extract the reference's own converged slab results into a fixture: slab_reference_results.csv.

Source: the reference tree's cases/verification/kinetic/results/slab_comparison.csv, written by its
compare_ref1d.py from `gsis-solver cis-run` runs (cases/verification/kinetic/run_slab.py; the case
TOMLs beside it, [kinetic] defaults, so the FP32 CUDA backend). Every run is kept: the `_per` cases
(every side periodic, converged to epsilon < 2e-7) and the default ones (heat transfer: both sides
symmetry planes; Couette flow: sym_y periodic, sym_z a symmetry plane; converged to epsilon < 1e-6);
the value is the volume-weighted mean of the cells' P_xy (Couette) or q_x (heat transfer). Before
writing, each case's TOML is read and its wall temperatures and velocities are checked against the
case kinds the tests rebuild (T_L/T0, T_R/T0 = 0.95/1.05 and 0.5/1.5; wall u_y = -/+0.1 and -/+1
sqrt(R T0)), its grid against 30 x 24 x 24 on [-10, 10] x [-8, 8]^2 and its side conditions against
the ones above.

Run from the repository root:  python tests/golden/make_slab_reference.py [--reference /root/reference]
"""

from __future__ import annotations

import argparse
import csv
import math
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
T0 = 273.15
C0 = math.sqrt(1.380649e-23 / 6.65e-27 * T0)
KINDS = {"heat_lin": (0.95, 1.05, 0.0), "heat_nl": (0.5, 1.5, 0.0), "couette_lin": (1.0, 1.0, 0.1),
         "couette_nl": (1.0, 1.0, 1.0)}
COLUMNS = ["case", "kind", "Kn", "nx", "sides", "quantity", "value", "ref_vmid30", "iterations", "epsilon"]


def check_case(path: str, kind: str, kn: float, sides: str) -> None:
    text = open(path, encoding="utf-8").read()
    tl, tr, u = KINDS[kind]
    temps = [float(x) for x in re.findall(r"^temperature = ([-0-9.e+]+)", text, re.M)]
    assert len(temps) == 2 and abs(temps[0] - tl * T0) < 1e-9 and abs(temps[1] - tr * T0) < 1e-9, (path, temps)
    vels = [float(x) for x in re.findall(r"^velocity = \[0\.0, ([-0-9.e+]+), 0\.0\]", text, re.M)]
    if u == 0.0:
        assert not vels, path
    else:
        assert len(vels) == 2 and abs(vels[0] + u * C0) < 1e-9 and abs(vels[1] - u * C0) < 1e-9, (path, vels)
    assert re.search(r"^Kn = ([-0-9.e+]+)", text, re.M).group(1) and \
        float(re.search(r"^Kn = ([-0-9.e+]+)", text, re.M).group(1)) == kn, path
    sy = re.search(r'patch = "sym_y"\ntype = "(\w+)"', text).group(1)
    sz = re.search(r'patch = "sym_z"\ntype = "(\w+)"', text).group(1)
    want = ("periodic_slab", "periodic_slab") if sides == "periodic" else \
        ("periodic_slab" if kind.startswith("couette") else "symmetry", "symmetry")
    assert (sy, sz) == want, (path, sy, sz)
    assert "n = [30, 24, 24]" in text and "range = [[-10.0, 10.0], [-8.0, 8.0], [-8.0, 8.0]]" in text, path
    assert "sweep_blocks = 1" in text and "precision" not in text, path


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args()
    kdir = os.path.join(a.reference, "cases", "verification", "kinetic")
    rows = []
    with open(os.path.join(kdir, "results", "slab_comparison.csv"), encoding="utf-8") as f:
        for r in csv.DictReader(f):
            tol = 2e-7 if r["sides"] == "periodic" else 1e-6
            assert r["converged"] == "True" and float(r["epsilon"]) < tol, r["case"]
            check_case(os.path.join(kdir, r["case"] + ".toml"), r["kind"], float(r["Kn"]), r["sides"])
            rows.append({k: r[k] for k in COLUMNS})
    out = os.path.join(HERE, "slab_reference_results.csv")
    with open(out, "w", newline="", encoding="utf-8") as f:
        w = csv.DictWriter(f, COLUMNS)
        w.writeheader()
        w.writerows(rows)
    print(f"{len(rows)} runs -> {out}")


if __name__ == "__main__":
    main()
