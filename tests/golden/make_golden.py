"""Write the committed golden vectors for the 1D slab benchmark case (200 cells, 24^3 velocities on
[-6, 6], delta = 1, walls T = 0.95 / 1.05, kinetic defaults, FP64, f initialised to the global
Maxwellian, 300 CIS iterations).

The expected values come from the CPU oracle (oracle/liboracle.so) driven by the library's solver
orchestration; when the reference's own kernels (oracle/_ref, built from the reference tree where it
exists) is present, the same run is repeated with it and the two must agree bit for bit before
anything is written. Output: slab_c1.npz with moments after iterations 1, 10 and 300 (original cell
numbering), the residual / epsilon history, and the wall fluxes at the end.

Run from the repository root:  python tests/golden/make_golden.py [--iterations 300]
"""

from __future__ import annotations

import argparse
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from conftest import load_package  # noqa: E402

import oracle  # noqa: E402


def run(kind: str, iterations: int, case_dir: str):
    g = load_package()
    import gsis_sim_amd.meshgen as mg
    path = mg.slab_case(case_dir, nx=200, n_vel=24, vrange=6.0, delta=1.0, kinetic={"precision": "fp64"})
    s = g.Setup(path)
    cb = oracle.CpuBackend(s, kind, ops_type=g.BackendOps)
    sv = g.Solver(s, cb.ops)
    sv.initialize_uniform(1.0, (0.0, 0.0, 0.0), 1.0)
    snaps, hist = {}, []
    t0 = time.time()
    for it in range(1, iterations + 1):
        sv.prepare()
        info = sv.sweep()
        hist.append((info["residual"], info["epsilon"]))
        if it in (1, 10, iterations):
            snaps[it] = sv.moments().copy()
        if it % 25 == 0:
            print(f"{kind}: iteration {it} residual {info['residual']:.6e} ({time.time() - t0:.0f} s)", flush=True)
    fp, fm, fs = sv.boundary_fluxes()
    return snaps, np.array(hist), fp, fm


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iterations", type=int, default=300)
    args = ap.parse_args()
    oracle.build()
    with tempfile.TemporaryDirectory() as d:
        snaps, hist, fp, fm = run("oracle", args.iterations, d)
    if oracle.ref_available():
        with tempfile.TemporaryDirectory() as d:
            rs, rh, rfp, rfm = run("ref", args.iterations, d)
        for k in snaps:
            assert np.array_equal(snaps[k], rs[k]), f"oracle and reference kernels differ at iteration {k}"
        assert np.array_equal(hist, rh), "residual histories differ"
        print("oracle == reference kernels (bitwise) over", args.iterations, "iterations")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "slab_c1.npz")
    np.savez_compressed(out, iterations=args.iterations, m1=snaps[1], m10=snaps[10], mN=snaps[args.iterations],
                        history=hist, f_plus=fp, f_minus=fm)
    print("wrote", out)


if __name__ == "__main__":
    main()
