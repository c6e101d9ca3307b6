"""The cis-run driver (gsis-sim_amd/cis_run.py) on the CPU, through the oracle backend: the result
files' formats (history.csv columns of src/kinetic/kinetic_output.cpp:323-389, summary.json keys of
src/app/cmd_cis_run.cpp:360-443, the CSV / JSON number forms of src/io/csv_writer.cpp:22-29 and
src/core/json.cpp:461-545) and what the physics fixes in the summaries (no mass or tangential momentum
through a symmetry plane, the tube mass flow against the setup's own Mdot factor, the convergence
stop of eq. (59))."""

from __future__ import annotations

import csv
import json
import math
import os
import random
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import box_case

TO_CHARS = r"""
#include <charconv>
#include <cstdio>
#include <cstdlib>
int main(int argc, char** argv) {
  for (int i = 1; i < argc; ++i) {
    char b[64];
    auto r = std::to_chars(b, b + sizeof(b), std::strtod(argv[i], nullptr));
    *r.ptr = 0;
    std::puts(b);
  }
}
"""


@pytest.fixture(scope="module")
def cr(gsk):
    import gsis_sim_amd.cis_run as m  # noqa: F401  (registered by load_package)
    import sys
    return sys.modules["gsis_sim_amd.cis_run"]


def test_number_form_matches_to_chars(cr, tmp_path):
    """cpp_number against the C++ library's std::to_chars(double) (what std::format("{}") writes)
    on special values and 6000 random doubles of every magnitude."""
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available to build the std::to_chars reference")
    src = tmp_path / "tc.cpp"
    src.write_text(TO_CHARS)
    exe = tmp_path / "tc"
    subprocess.run(["g++", "-std=c++17", "-O1", str(src), "-o", str(exe)], check=True)
    rng = random.Random(5)
    vals = [0.001, 0.0001, 1e16, 100.0, 123456.0, 1e-5, 0.5, -2.5e-300, 5e-324, 1.7976931348623157e308,
            1234567890123456789.0, 0.1, 1e15, 1e22, 12345678.9, -0.0, 0.0, 1.0, -1.0, 2.0 ** 53, 2.0 ** 60, 1e21,
            123456789012345680000.0]
    while len(vals) < 6000:
        k = rng.random()
        if k < 0.4:
            v = struct.unpack("d", struct.pack("Q", rng.getrandbits(64)))[0]
        elif k < 0.7:
            v = rng.uniform(-1, 1) * 10 ** rng.randint(-30, 30)
        else:
            v = float(round(rng.uniform(-1e6, 1e6), rng.randint(0, 8))) * 10 ** rng.randint(-8, 20)
        if math.isfinite(v):
            vals.append(v)
    out = subprocess.run([str(exe)] + [repr(v) for v in vals], capture_output=True, text=True, check=True).stdout.split()
    assert len(out) == len(vals)
    bad = [(v, want, cr.cpp_number(v)) for v, want in zip(vals, out) if cr.cpp_number(v) != want]
    assert not bad, bad[:5]
    assert cr.cpp_number(float("nan")) == "nan" and cr.cpp_number(float("-inf")) == "-inf"


def test_json_layout(cr):
    """json::dump's layout: reals keep a decimal point or exponent, numeric arrays on one line,
    objects and arrays of objects one entry per line, two-space indent."""
    v = {"a": 1.0, "b": 3, "c": [1.0, 2.5e-05, None, True], "d": {"e": "x\"y", "f": []}, "g": [{"h": 0.1}], "i": {}}
    want = ('{\n  "a": 1.0,\n  "b": 3,\n  "c": [1.0, 2.5e-05, null, true],\n  "d": {\n    "e": "x\\"y",\n'
            '    "f": []\n  },\n  "g": [\n    {\n      "h": 0.1\n    }\n  ],\n  "i": {}\n}')
    assert cr.dump_json(v) == want
    assert json.loads(want) == v
    with pytest.raises(ValueError):
        cr.dump_json({"x": float("nan")})


def _run(gsk, cr, oracle_mod, case, out, **kw):
    return cr.cis_run(case, out, backend_factory=lambda st: oracle_mod.CpuBackend(st, "oracle", ops_type=gsk.BackendOps),
                      **kw)


def _history(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return rows[0], rows[1:]


def test_history_and_summary_files(gsk, meshgen, oracle_mod, cr, tmp_path):
    case = box_case(meshgen, str(tmp_path), kinetic={"precision": "fp64"})
    out = str(tmp_path / "run")
    s = _run(gsk, cr, oracle_mod, case, out, max_iter=4, cells_csv=True)
    header, rows = _history(os.path.join(out, "history.csv"))
    names = ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]
    want = (["iteration", "epsilon", "eps_rho", "eps_u", "eps_T", "residual", "residual_rel", "total_mass", "mass_scale"]
            + [f"mdot_{n}_kg_s" for n in names] + [f"mdot_{n}_nd" for n in names]
            + ["qx_mean", "qy_mean", "qz_mean", "sxx_mean", "syy_mean", "szz_mean", "sxy_mean", "syz_mean", "sxz_mean",
               "time_iter_s", "time_total_s", "t_moments_s", "t_boundary_s", "t_gradient_s", "t_residual_s",
               "t_sweep_forward_s", "t_sweep_backward_s", "t_other_s", "mass_imbalance", "gpu_memory_bytes"])
    assert header == want
    # iterations 0..3 from sweeps, then the stop record of iteration 4 (prepare only)
    assert [int(r[0]) for r in rows] == [0, 1, 2, 3, 4]
    assert rows[0][1] == "nan" and rows[0][6] == "1"  # no epsilon before the second iterate; residual_rel = 1
    assert rows[-1][5] == "nan"  # a prepare-only record carries no residual
    col = {h: i for i, h in enumerate(header)}
    desc = cr.Setup(case).describe()
    mflux, L = desc["reference"]["mass_flux_kg_m2_s"], desc["reference"]["length_m"]
    for r in rows:
        for n in names:  # mass_flow_to_si: nd * mass_flux * L * L (kinetic_output.cpp:25-28)
            nd, si = float(r[col[f"mdot_{n}_nd"]]), float(r[col[f"mdot_{n}_kg_s"]])
            assert si == nd * mflux * L * L
        flows = [float(r[col[f"mdot_{n}_nd"]]) for n in names]
        mb = cr.mass_balance(flows)
        assert float(r[col["mass_imbalance"]]) == mb["imbalance"]

    with open(os.path.join(out, "summary.json")) as f:
        text = f.read()
    js = json.loads(text)
    assert text == cr.dump_json(s) + "\n"
    assert list(js)[:12] == ["program", "case", "case_file", "mesh", "n_cells", "n_boundary_faces", "n_velocity", "Kn",
                             "delta", "iterations", "converged", "stopped"]
    assert js["iterations"] == 4 and js["converged"] is False and js["n_velocity"] == 8 ** 3
    assert isinstance(js["params"]["pseudo_cfl"], float) and js["params"]["inner_sweeps"] == 4
    assert [p["name"] for p in js["patches"]] == names
    P = {p["name"]: p for p in js["patches"]}
    for n in ("zmin", "zmax"):  # specular planes: no mass, energy or tangential momentum through them
        assert P[n]["mdot_nd"] == 0.0 and P[n]["energy_flux_nd"] == 0.0
        assert P[n]["momentum_flux_nd"][0] == 0.0 and P[n]["momentum_flux_nd"][1] == 0.0
    assert P["ymax"]["accommodation"] == 0.8
    # faces and areas: the box's side planes
    assert sum(p["faces"] for p in js["patches"]) == js["n_boundary_faces"]
    # tube mass flow against the setup's own nondimensional -> Mdot factor (an independent path)
    factor = cr.Setup(case).mass_flow_factor()
    assert js["mass_flow"]["M_dot_inlet"] == pytest.approx(-P["xmin"]["mdot_nd"] * factor, rel=1e-14)
    assert js["mass_flow"]["M_dot_outlet"] == pytest.approx(P["xmax"]["mdot_nd"] * factor, rel=1e-14)
    assert js["mass_balance"]["kinetic"]["inflow_nd"] == pytest.approx(-P["xmin"]["mdot_nd"] - min(0.0, P["ymin"]["mdot_nd"])
                                                                     - min(0.0, P["ymax"]["mdot_nd"]), rel=1e-12)
    with open(os.path.join(out, "cells.csv")) as f:
        cells = list(csv.reader(f))
    assert cells[0][:6] == ["cell", "x", "y", "z", "volume", "rho"] and len(cells) == 1 + js["n_cells"]
    vols = np.array([float(r[4]) for r in cells[1:]])
    assert np.all(vols > 0)


def test_stops_at_tolerance(gsk, meshgen, oracle_mod, cr, tmp_path):
    """eq. (59): the run stops at the first prepare after min_iter whose epsilon is below tol."""
    case = box_case(meshgen, str(tmp_path), kinetic={"precision": "fp64"})
    s = _run(gsk, cr, oracle_mod, case, str(tmp_path / "run"), max_iter=50, min_iter=3, tol=0.3)
    header, rows = _history(str(tmp_path / "run" / "history.csv"))
    eps = [float(r[1]) for r in rows]
    assert s["converged"] is True and s["epsilon"] < 0.3
    first = next(i for i, e in enumerate(eps) if i >= 3 and e < 0.3)
    assert s["iterations"] == first == len(rows) - 1


def test_closed_box_has_no_net_mass_flow(gsk, meshgen, oracle_mod, cr, tmp_path):
    """Walls and symmetry planes only: every patch flow is zero to round-off and no tube mass flow
    is written (no inlet or outlet)."""
    case = box_case(meshgen, str(tmp_path), bcs="closed", kinetic={"precision": "fp64"})
    s = _run(gsk, cr, oracle_mod, case, str(tmp_path / "run"), max_iter=3)
    assert "mass_flow" not in s
    assert max(abs(p["mdot_nd"]) for p in s["patches"]) < 1e-12
    assert s["mass_balance"]["kinetic"]["imbalance"] is None or abs(s["mass_balance"]["kinetic"]["outflow_nd"]) < 1e-12


def test_in_memory_loop_matches_the_files(gsk, meshgen, oracle_mod, cr, tmp_path):
    """run() (the loop the time-to-Mdot benchmark times) records the same inlet Mdot per iteration
    as the history.csv of cis_run on the same case, and time_to_within finds the last entry into
    the band."""
    case = box_case(meshgen, str(tmp_path), kinetic={"precision": "fp64"})
    _run(gsk, cr, oracle_mod, case, str(tmp_path / "run"), max_iter=4)
    header, rows = _history(str(tmp_path / "run" / "history.csv"))
    nd = np.array([float(r[header.index("mdot_xmin_nd")]) for r in rows[:4]])
    setup = cr.Setup(case)
    cb = oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps)
    sv = cr.Solver(setup, cb.ops)
    h = cr.run(sv, setup.mass_flow_factor(), inlet=0, outlet=1, max_iterations=4)
    sv.close()
    assert h["reason"] == "max_iterations" and list(h["iteration"]) == [0, 1, 2, 3]
    assert np.array_equal(h["mdot_in"], -nd * setup.mass_flow_factor())
    fake = {"mdot_in": np.array([0.5, 1.2, 0.99995, 1.00001, 1.0]), "t_wall": np.arange(1.0, 6.0),
            "iteration": np.arange(5)}
    assert cr.time_to_within(fake, 1.0, 1e-4) == (3.0, 2)
    assert cr.time_to_within({**fake, "mdot_in": np.array([1.0, 1.0, 1.0, 1.0, 2.0])}, 1.0, 1e-4) is None


def test_restart_from_cells_csv(gsk, meshgen, oracle_mod, cr, tmp_path):
    """--init-from: a run restarted from cells.csv starts from the equilibrium of those cell values
    (the discrete equilibrium reproduces them to round-off), and a cells.csv that lacks a cell or a
    column is rejected."""
    case = box_case(meshgen, str(tmp_path), kinetic={"precision": "fp64"})
    fac = lambda st: oracle_mod.CpuBackend(st, "oracle", ops_type=gsk.BackendOps)  # noqa: E731
    _run(gsk, cr, oracle_mod, case, str(tmp_path / "a"), max_iter=3, cells_csv=True)
    cells = str(tmp_path / "a" / "cells.csv")
    w = cr.read_cells_csv(cells, 30)
    setup = cr.Setup(case)
    keep = fac(setup)
    sv = cr.Solver(setup, keep.ops)
    sv.initialize_equilibrium(w)
    m = sv.moments()
    assert np.allclose(m[:, [0, 4]], w[:, [0, 4]], rtol=1e-10, atol=0)
    assert np.max(np.abs(m[:, 1:4] - w[:, 1:4])) < 1e-10
    sv.close()
    s = cr.cis_run(case, str(tmp_path / "b"), max_iter=1, backend_factory=fac, init_from=cells)
    assert s["iterations"] == 1
    rows = open(cells).read().splitlines()
    open(str(tmp_path / "short.csv"), "w").write("\n".join(rows[:-1]) + "\n")
    with pytest.raises(ValueError, match="no values for cell"):
        cr.read_cells_csv(str(tmp_path / "short.csv"), 30)
    open(str(tmp_path / "nocol.csv"), "w").write(rows[0].replace(",T,", ",Temp,") + "\n")
    with pytest.raises(ValueError, match="no column"):
        cr.read_cells_csv(str(tmp_path / "nocol.csv"), 30)


def test_stop_file_ends_the_run_with_results(gsk, meshgen, oracle_mod, cr, tmp_path):
    """A STOP file in the output directory ends the run after the current iteration: the summary
    says stopped (by the user), and cells.csv is written although it was not asked for
    (src/app/cmd_cis_run.cpp:324-343)."""
    case = box_case(meshgen, str(tmp_path), kinetic={"precision": "fp64"})
    out = tmp_path / "run"
    out.mkdir()
    (out / "STOP").write_text("")
    s = _run(gsk, cr, oracle_mod, case, str(out), max_iter=50)
    assert s["stopped"] and s["stopped_by_user"] and not s["converged"]
    assert s["iterations"] == 1
    assert (out / "cells.csv").exists()
    _, rows = _history(str(out / "history.csv"))
    assert len(rows) == 2  # the sweep of iteration 0 and the final state


def test_checkpoint_and_restart(gsk, meshgen, oracle_mod, cr, tmp_path):
    """checkpoint_every writes f (original numbering, grid order) that restart_from reads back: the
    restarted solver holds exactly the checkpointed f, and a restarted run continues from it."""
    case = box_case(meshgen, str(tmp_path), kinetic={"precision": "fp64"})
    fac = lambda st: oracle_mod.CpuBackend(st, "oracle", ops_type=gsk.BackendOps)  # noqa: E731
    _run(gsk, cr, oracle_mod, case, str(tmp_path / "a"), max_iter=4, checkpoint_every=2)
    ck = str(tmp_path / "a" / "checkpoint.npy")
    f = np.load(ck)
    setup = cr.Setup(case)
    assert f.shape == (setup.n_cells, setup.nv)
    keep = fac(setup)
    sv = cr.Solver(setup, keep.ops)
    sv.set_distribution(f)
    assert np.array_equal(sv.distribution().reshape(f.shape), f)
    sv.close()
    s = cr.cis_run(case, str(tmp_path / "b"), max_iter=2, backend_factory=fac, restart_from=ck)
    assert s["iterations"] == 2
    with pytest.raises(ValueError, match="holds"):
        np.save(str(tmp_path / "bad.npy"), f[:-1])
        cr.cis_run(case, str(tmp_path / "c"), max_iter=1, backend_factory=fac, restart_from=str(tmp_path / "bad.npy"))
