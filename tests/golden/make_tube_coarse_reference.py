"""Writes tests/golden/tube_coarse_reference_results.csv: the reference's own converged CIS runs of
the tube-to-vacuum case on its coarse tube mesh (tests/golden/tube_hex_coarse.npz; cases
tube_coarse_delta{1,10}_cis.toml: helium, L_ref = R = 1 mm, T0 = 273.15 K, pressure inlet, vacuum
outlet, diffuse wall at T0, 30 x 24 x 24 velocities on [-10, 10] x [-8, 8]^2), copied from
cases/tube_vacuum/kinetic/results/param_study.csv (limiter_k = 3, the default),
cases/tube_vacuum/kinetic/results/param_study_relK1.csv (limiter_k = 1) and
cases/tube_vacuum/phase5/results/cis_study_coarse.csv (limiter_k = 3). Each run stopped at the
reference's convergence measure epsilon < 1e-5; the columns kept are the run tag, delta_rp, the
kinetic settings that differ from the defaults, limiter_k, iterations and the inlet / outlet mass
flows (Mdot, mass_flow_scale units). Run from the repository root:
    python tests/golden/make_tube_coarse_reference.py [reference tree, default /root/reference]
"""

from __future__ import annotations

import csv
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main() -> None:
    reference = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    base = os.path.join(reference, "cases", "tube_vacuum")
    sources = [("kinetic/results/param_study.csv", 3.0), ("kinetic/results/param_study_relK1.csv", 1.0),
               ("phase5/results/cis_study_coarse.csv", 3.0)]
    rows = []
    for rel, k in sources:
        with open(os.path.join(base, rel)) as f:
            for r in csv.DictReader(f):
                status = r.get("status") or ("converged" if r.get("converged") == "True" else "max_iter")
                if status != "converged":
                    continue
                settings = r.get("settings", r.get("sets", "")) or ""
                k_run = k
                for item in settings.split():
                    if item.startswith("limiter_k="):
                        k_run = float(item.split("=")[1])
                rows.append(dict(source=rel, tag=r["tag"], delta_rp=r["delta"], settings=settings, limiter_k=k_run,
                                 iterations=r["iterations"],
                                 mdot_inlet=r.get("M_dot_inlet") or r.get("Mdot_inlet"),
                                 mdot_outlet=r.get("M_dot_outlet") or r.get("Mdot_outlet")))
    out = os.path.join(HERE, "tube_coarse_reference_results.csv")
    with open(out, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0]))
        w.writeheader()
        w.writerows(rows)
    print(f"wrote {len(rows)} runs to {out}")


if __name__ == "__main__":
    main()
