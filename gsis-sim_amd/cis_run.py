"""CIS-only steady solve with the reference's result files: `gsis-solver cis-run` (src/app/cmd_cis_run.cpp).

The loop follows src/app/cmd_cis_run.cpp:280-330: prepare; stop when the epsilon of eq. (59) is below
the tolerance (after min_iter, with the mass balance and mass-flow targets met when those are
required) or at max_iter; otherwise one CIS iteration (`KineticSolver::sweep`). Each record goes to
history.csv with the columns of KineticHistoryWriter (src/kinetic/kinetic_output.cpp:323-389); at the
end summary.json holds the keys of cmd_cis_run.cpp:360-443, with the patch summaries of
summarize_kinetic_patches (kinetic_output.cpp:198-271), the tube mass flow of tube_mass_flow_json
(:273-284) and the balance record of src/app/mass_balance_report.cpp:16-73. Numbers are written the
way the reference writes them: std::format("{}") in the CSV (shortest round-trip form, fixed or
scientific, whichever is shorter), std::to_chars plus ".0" for integral reals in the JSON
(src/io/csv_writer.cpp:22-29, src/core/json.cpp:461-472), and the JSON layout of json::dump (two-space
indent, arrays of numbers on one line).

The section mass flows of [output].mass_flow_sections (src/mesh/section_cut.hpp) go into both files
as the reference writes them (history columns M_dot_sec<k>, summary mass_flow_sections and the
balance record's section spread). Not written: fields.vtu / boundary.vtu (visualisation) and the
STOP-file interrupt. cells.csv is written when asked (`cells_csv=True`, the reference's --cells-csv).

`init_from` restarts from a previous cells.csv: f = the equilibrium of each cell's (rho, u, T)
(read_kinetic_cells_csv, src/kinetic/kinetic_output.cpp:142-196; cmd_cis_run.cpp:215-219).

Run as a script: python gsis-sim_amd/cis_run.py case.toml [--out DIR] [--max-iter N] [--min-iter N]
[--tol X] [--precision fp32|fp64] [--set key=value]... [--print-every N] [--cells-csv]
[--init-from cells.csv]
"""

from __future__ import annotations

import decimal
import math
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):  # run as a script: load the package from this directory
    import importlib.util

    _spec = importlib.util.spec_from_file_location(
        "gsis_sim_amd", os.path.join(os.path.dirname(os.path.abspath(__file__)), "__init__.py"),
        submodule_search_locations=[os.path.dirname(os.path.abspath(__file__))])
    _pkg = importlib.util.module_from_spec(_spec)
    sys.modules["gsis_sim_amd"] = _pkg
    _spec.loader.exec_module(_pkg)
    from gsis_sim_amd import Setup, Solver, lib  # noqa: E402
else:
    from . import Setup, Solver, lib

# ------------------------------------------------------------------ number and JSON formatting


def cpp_number(x: float) -> str:
    """std::format("{}", x) / std::to_chars(x) for a double: the shortest digits that read back to
    x, in fixed or scientific notation, whichever is shorter (fixed on a tie; an integral value in
    fixed form shows its exact digits); exponents with a sign and at least two digits."""
    x = float(x)
    if math.isnan(x):
        return "nan"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    if x == 0.0:
        return "-0" if math.copysign(1.0, x) < 0 else "0"
    sign, digits, exp = decimal.Decimal(repr(x)).as_tuple()
    ds = "".join(map(str, digits)).rstrip("0")
    exp += len(digits) - len(ds)  # value = int(ds) * 10^exp
    n = len(ds)
    e10 = exp + n - 1
    sci = ds[0] + ("." + ds[1:] if n > 1 else "") + "e" + ("-" if e10 < 0 else "+") + f"{abs(e10):02d}"
    if exp >= 0:
        fixed = str(int(abs(x)))  # an integral value prints its exact digits in fixed form
    elif -exp < n:
        fixed = ds[: n + exp] + "." + ds[n + exp:]
    else:
        fixed = "0." + "0" * (-exp - n) + ds
    body = fixed if len(fixed) <= len(sci) else sci
    return ("-" if sign else "") + body


def json_number(x: float) -> str:
    """json::dump of a real (src/core/json.cpp:461-472): to_chars, ".0" appended to an integral
    form; non-finite values are an error."""
    if not math.isfinite(x):
        raise ValueError("a non-finite real cannot be written to JSON")
    t = cpp_number(x)
    return t if any(c in t for c in ".eE") else t + ".0"


def _json_string(s: str) -> str:
    out = ['"']
    for c in s:
        if c == '"':
            out.append('\\"')
        elif c == "\\":
            out.append("\\\\")
        elif c == "\b":
            out.append("\\b")
        elif c == "\f":
            out.append("\\f")
        elif c == "\n":
            out.append("\\n")
        elif c == "\r":
            out.append("\\r")
        elif c == "\t":
            out.append("\\t")
        elif ord(c) < 0x20:
            out.append(f"\\u{ord(c):04x}")
        else:
            out.append(c)
    out.append('"')
    return "".join(out)


def dump_json(v, indent: int = 2, level: int = 0) -> str:
    """json::dump layout (src/core/json.cpp:480-545): one key per line, arrays whose elements are
    all numbers / booleans / nulls on one line separated by ", "."""
    if v is None:
        return "null"
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return json_number(float(v))
    if isinstance(v, str):
        return _json_string(v)
    if isinstance(v, (list, tuple)):
        if not v:
            return "[]"
        scalars = all(e is None or isinstance(e, (bool, int, float, np.integer, np.floating)) for e in v)
        if scalars:
            return "[" + ", ".join(dump_json(e, indent, level + 1) for e in v) + "]"
        pad, inner = "\n" + " " * (indent * level), "\n" + " " * (indent * (level + 1))
        return "[" + ",".join(inner + dump_json(e, indent, level + 1) for e in v) + pad + "]"
    if isinstance(v, dict):
        if not v:
            return "{}"
        pad, inner = "\n" + " " * (indent * level), "\n" + " " * (indent * (level + 1))
        return "{" + ",".join(inner + _json_string(k) + ": " + dump_json(x, indent, level + 1)
                              for k, x in v.items()) + pad + "}"
    raise TypeError(f"cannot write {type(v).__name__} to JSON")


def _csv_field(v) -> str:
    if isinstance(v, str):
        return '"' + v.replace('"', '""') + '"' if any(c in v for c in ',"\n\r') else v
    if isinstance(v, (int, np.integer)) and not isinstance(v, bool):
        return str(int(v))
    return cpp_number(float(v))


# ------------------------------------------------------------------ mass flows and balance
# (src/kinetic/kinetic_output.cpp:15-68)

def mass_flow_to_si(desc: dict, nondimensional: float) -> float:
    r = desc["reference"]
    return nondimensional * r["mass_flux_kg_m2_s"] * r["length_m"] * r["length_m"]


def mass_flow_display(desc: dict, nondimensional: float) -> float:
    if not desc["has_p_in"]:
        return nondimensional
    return mass_flow_to_si(desc, nondimensional) / desc["mass_flow_scale_kg_s"]


def mass_flow_targets_met(desc: dict, flow) -> bool:
    for p, pc in enumerate(desc["patches"]):
        if pc["type"] != "mass_flow_inlet":
            continue
        t = pc["target_mass_flow"]
        if p >= len(flow) or not math.isfinite(flow[p]) or abs(flow[p] + t) > 1e-3 * t:
            return False
    return True


def mass_balance(flows) -> dict:
    """Inflow / outflow totals of outward patch flows and imbalance = (out - in) / max(in, out)."""
    b = {"inflow": 0.0, "outflow": 0.0, "imbalance": 0.0, "valid": False}
    for F in flows:
        if not math.isfinite(F):
            return b
        if F > 0.0:
            b["outflow"] += F
        else:
            b["inflow"] -= F
    scale = max(b["inflow"], b["outflow"])
    if scale > 0.0:
        b["imbalance"] = (b["outflow"] - b["inflow"]) / scale
        b["valid"] = True
    return b


def flow_spread(flows) -> float:
    if len(flows) < 2:
        return math.nan
    lo = hi = flows[0]
    total = 0.0
    for x in flows:
        if not math.isfinite(x):
            return math.nan
        lo, hi = min(lo, x), max(hi, x)
        total += x
    mean = total / len(flows)
    return (hi - lo) / abs(mean) if mean != 0.0 else math.nan


# ------------------------------------------------------------------ history.csv

TIMING_KEYS = ("moments", "boundary", "gradient", "residual", "sweep_forward", "sweep_backward", "other")


def mass_flow_display_prefix(desc: dict) -> str:
    return "M_dot_" if desc["has_p_in"] else "mdot_nd_"


class HistoryWriter:
    """history.csv with KineticHistoryWriter's columns (src/kinetic/kinetic_output.cpp:323-389)."""

    def __init__(self, path: str, desc: dict, gpu_memory_bytes: int):
        self.desc = desc
        self.gpu_memory_bytes = int(gpu_memory_bytes)
        self.first_residual = 0.0
        names = [pc["name"] for pc in desc["patches"]]
        header = ["iteration", "epsilon", "eps_rho", "eps_u", "eps_T", "residual", "residual_rel", "total_mass",
                  "mass_scale"]
        header += [f"mdot_{n}_kg_s" for n in names] + [f"mdot_{n}_nd" for n in names]
        header += ["qx_mean", "qy_mean", "qz_mean", "sxx_mean", "syy_mean", "szz_mean", "sxy_mean", "syz_mean",
                   "sxz_mean"]
        header += ["time_iter_s", "time_total_s"] + [f"t_{k}_s" for k in TIMING_KEYS]
        self.n_sections = len(desc["sections"]["positions"])
        header += ["mass_imbalance"] + [f"{mass_flow_display_prefix(desc)}sec{k}" for k in range(self.n_sections)]
        header += ["gpu_memory_bytes"]
        header += [f"p_b_{pc['name']}_Pa" for pc in desc["patches"] if pc["type"] == "mass_flow_inlet"]
        self.header = header
        self._f = open(path, "w", newline="")
        self._line(header)

    def _line(self, values) -> None:
        self._f.write(",".join(_csv_field(v) for v in values) + "\n")
        self._f.flush()

    def write(self, info: dict, flows, elapsed_total_s: float, sections=()) -> None:
        nan = math.nan
        eps = info["has_epsilon"]
        if info["has_residual"] and self.first_residual == 0.0:
            self.first_residual = info["residual"]
        row = [int(info["iteration"])]
        row += [info["epsilon"] if eps else nan, info["eps_rho"] if eps else nan, info["eps_u"] if eps else nan,
                info["eps_T"] if eps else nan]
        row.append(info["residual"] if info["has_residual"] else nan)
        row.append(info["residual"] / self.first_residual if info["has_residual"] and self.first_residual > 0.0
                   else nan)
        row += [info["total_mass"], info["mass_scale"]]
        row += [mass_flow_to_si(self.desc, F) for F in flows] + [float(F) for F in flows]
        row += list(info["q_mean"]) + list(info["sigma_mean"])
        row += [info["seconds"], elapsed_total_s] + [info["timings"][k] for k in TIMING_KEYS]
        mb = mass_balance(flows)
        row.append(mb["imbalance"] if mb["valid"] else nan)
        row += [mass_flow_display(self.desc, sections[k]) if k < len(sections) else nan for k in range(self.n_sections)]
        row.append(self.gpu_memory_bytes)
        row += [pc["p"] * self.desc["reference"]["pressure_Pa"] for pc in self.desc["patches"]
                if pc["type"] == "mass_flow_inlet"]
        assert len(row) == len(self.header)
        self._line(row)

    def close(self) -> None:
        self._f.close()


# ------------------------------------------------------------------ patch summaries


def summarize_patches(desc: dict, geometry, fluxes) -> dict:
    """summarize_kinetic_patches (src/kinetic/kinetic_output.cpp:198-271): per patch the sums of
    F+ + F- over its faces (in face order), the area-weighted face state, the heat flow, and the
    inlet / outlet totals."""
    patch_of, area_m2, normal = geometry
    fp, fm, fs = fluxes
    L2 = desc["length_scale"] ** 2
    ref = desc["reference"]
    out = {"patches": [], "mdot": [0.0] * len(desc["patches"]), "mdot_in": 0.0, "mdot_out": 0.0,
           "has_in": False, "has_out": False}
    for pi, pc in enumerate(desc["patches"]):
        F = [0.0] * 5
        face = [0.0] * 5
        area = heat = 0.0
        for b in np.flatnonzero(patch_of == pi):
            A = float(area_m2[b]) / L2
            for k in range(5):
                F[k] += float(fp[b, k]) + float(fm[b, k])
                face[k] += A * float(fs[b, k])
            w = fs[b]
            u = (float(w[1]), float(w[2]), float(w[3]))
            n = normal[b]
            q = [float(fp[b, k]) + float(fm[b, k]) for k in range(5)]
            un = u[0] * float(n[0]) + u[1] * float(n[1]) + u[2] * float(n[2])
            heat += (q[4] - u[0] * q[1] - u[1] * q[2] - u[2] * q[3]
                     + 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * q[0] - 1.5 * float(w[0]) * float(w[4]) * un * A)
            area += A
        pj = {"name": pc["name"], "type": pc["type"]}
        if pc["type"] == "maxwell_wall":
            pj["accommodation"] = pc["accommodation"]
        if pc["type"] == "mass_flow_inlet":
            pj["pressure_Pa"] = pc["p"] * ref["pressure_Pa"]
            pj["target_mass_flow_kg_s"] = mass_flow_to_si(desc, pc["target_mass_flow"])
            pj["mass_flow_relative_error"] = abs(F[0] + pc["target_mass_flow"]) / pc["target_mass_flow"]
        pj["faces"] = int(pc["n_faces"])
        pj["area_nd"] = area
        pj["mdot_nd"] = F[0]
        pj["mdot_kg_s"] = mass_flow_to_si(desc, F[0])
        pj["momentum_flux_nd"] = [F[1], F[2], F[3]]
        pj["energy_flux_nd"] = F[4]
        pj["area_m2"] = area * L2
        pj["mass_flow_kg_s"] = -mass_flow_to_si(desc, F[0])
        pj["heat_flow_W"] = -heat * ref["heat_flux_W_m2"] * L2
        pj["energy_flow_W"] = -F[4] * ref["heat_flux_W_m2"] * L2
        pj["force_N"] = [F[k] * ref["pressure_Pa"] * L2 for k in (1, 2, 3)]
        if area > 0.0:
            pj["face_rho_nd"] = face[0] / area
            pj["face_u_nd"] = [face[1] / area, face[2] / area, face[3] / area]
            pj["face_T_nd"] = face[4] / area
        out["patches"].append(pj)
        out["mdot"][pi] = F[0]
        if pc["type"] in ("pressure_outlet", "vacuum"):
            out["mdot_out"] += F[0]
            out["has_out"] = True
        elif pc["type"] in ("pressure_inlet", "mass_flow_inlet"):
            out["mdot_in"] -= F[0]
            out["has_in"] = True
    return out


def tube_mass_flow(desc: dict, ps: dict):
    """tube_mass_flow_json (kinetic_output.cpp:273-284): Mdot = mdot / (sqrt(pi) L^2 p_in / v_m)."""
    if not desc["has_p_in"] or not (ps["has_in"] or ps["has_out"]):
        return None
    scale = desc["mass_flow_scale_kg_s"]
    mf = {"scale_kg_s": scale}
    if ps["has_in"]:
        mf["M_dot_inlet"] = mass_flow_to_si(desc, ps["mdot_in"]) / scale
    if ps["has_out"]:
        mf["M_dot_outlet"] = mass_flow_to_si(desc, ps["mdot_out"]) / scale
    if ps["has_in"] and ps["has_out"]:
        mf["M_dot"] = 0.5 * (mass_flow_to_si(desc, ps["mdot_in"]) + mass_flow_to_si(desc, ps["mdot_out"])) / scale
    return mf


def balance_record(desc: dict, patch_flows, sections=()) -> dict:
    """The kinetic entry of the mass-balance report (src/app/mass_balance_report.cpp:16-73):
    inflow / outflow totals, imbalance, the spread of {inflow, outflow, |sections|}, and warnings
    above the tolerance."""
    b = mass_balance(patch_flows)
    spread = flow_spread([b["inflow"], b["outflow"]] + [abs(x) for x in sections]) if b["valid"] else math.nan
    tol = desc["mass_balance_tolerance"]
    warnings = []
    if b["valid"] and abs(b["imbalance"]) > tol:
        warnings.append(f"kinetic (last f): inlet / outlet mass-flow imbalance {b['imbalance']:+.3e} "
                        f"(tolerance {tol:.1e})")
    if len(sections) and math.isfinite(spread) and spread > tol:
        warnings.append(f"kinetic (last f): section / inlet / outlet mass-flow spread {spread:.3e} (tolerance {tol:.1e})")
    e = {"inflow_nd": b["inflow"], "outflow_nd": b["outflow"]}
    if desc["has_p_in"]:
        e["M_dot_inflow"] = mass_flow_display(desc, b["inflow"])
        e["M_dot_outflow"] = mass_flow_display(desc, b["outflow"])
    e["imbalance"] = b["imbalance"] if b["valid"] and math.isfinite(b["imbalance"]) else None
    e["section_spread"] = spread if math.isfinite(spread) else None
    e["checked"] = True
    return {"tolerance": tol, "required": desc["mass_balance_required"],
            "definition": "imbalance = (outflow - inflow) / max(inflow, outflow) from the outward boundary mass "
                          "flows; section_spread = (max - min) / mean of {inflow, outflow, |sections|}",
            "kinetic": e, "within_tolerance": not warnings, "warnings": warnings}


def sections_record(desc: dict, cuts: dict, flows) -> dict:
    """mass_flow_sections (src/app/mass_balance_report.cpp:75-100) for the kinetic f."""
    out = []
    for k, x in enumerate(desc["sections"]["positions"]):
        e = {"index": k, "position": float(x), "face_position": float(cuts["face_position"][k]),
             "n_faces": int(cuts["n_faces"][k]), "area_nd": float(cuts["area"][k]),
             "projected_area_nd": float(cuts["projected_area"][k]),
             "mdot_kinetic_nd": float(flows[k]) if math.isfinite(flows[k]) else None}
        if desc["has_p_in"]:
            m = mass_flow_display(desc, float(flows[k]))
            e["M_dot_kinetic"] = m if math.isfinite(m) else None
        out.append(e)
    return {"axis": desc["sections"]["axis"],
            "definition": "sum of the discrete face mass fluxes over the internal faces between cells whose centres lie "
                          "before and after the position (positive along the axis); kinetic = the last f",
            "sections": out}


def read_cells_csv(path: str, n_cells: int) -> np.ndarray:
    """(rho, ux, uy, uz, T) per cell from a cells.csv, located by the header's cell, rho, ux, uy, uz
    and T columns; every cell must appear (read_kinetic_cells_csv, kinetic_output.cpp:142-196)."""
    import csv

    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    if not rows:
        raise ValueError(f"{path} is empty")
    head = rows[0]
    need = ("cell", "rho", "ux", "uy", "uz", "T")
    missing = [c for c in need if c not in head]
    if missing:
        raise ValueError(f"{path} has no column(s) {', '.join(missing)}")
    col = [head.index(c) for c in need]
    w = np.zeros((n_cells, 5))
    seen = np.zeros(n_cells, bool)
    for r in rows[1:]:
        if not r:
            continue
        c = int(float(r[col[0]]))
        if not 0 <= c < n_cells:
            raise ValueError(f"{path}: cell number {c} is out of range")
        w[c] = [float(r[k]) for k in col[1:]]
        seen[c] = True
    if not seen.all():
        raise ValueError(f"{path} has no values for cell {int(np.flatnonzero(~seen)[0])}")
    return w


def write_cells_csv(path: str, setup: Setup, solver: Solver, desc: dict) -> None:
    """cells.csv (write_kinetic_cells_csv, kinetic_output.cpp:125-140): centre / L, V / L^3 and the
    15 nondimensional cell values per cell, original numbering, plus the cell residual."""
    L = desc["length_scale"]
    n2o = setup.cell_new_to_old()
    centers = np.empty((setup.n_cells, 3))
    centers[n2o] = setup.cell_centers()
    vol = np.empty(setup.n_cells)
    v = setup.view()
    vol[n2o] = np.ctypeslib.as_array(v.volume, shape=(setup.n_cells,))
    m = solver.moments()
    res = solver.residual_per_cell()
    cols = ["cell", "x", "y", "z", "volume", "rho", "ux", "uy", "uz", "T", "p", "sxx", "syy", "szz", "sxy", "syz",
            "sxz", "qx", "qy", "qz", "residual"]
    with open(path, "w", newline="") as f:
        f.write(",".join(cols) + "\n")
        for c in range(setup.n_cells):
            row = [float(c), *(centers[c] / L), vol[c], *m[c, :15], res[c]]
            f.write(",".join(cpp_number(x) for x in row) + "\n")


# ------------------------------------------------------------------ the run


def cis_run(case_file: str, out_dir: str | None = None, max_iter: int | None = None, min_iter: int = 1,
            tol: float | None = None, precision: str | None = None, overrides: dict | None = None,
            print_every: int = 0, cells_csv: bool = False, backend_factory=None, log=print,
            init_from: str | None = None, checkpoint_every: int = 0, restart_from: str | None = None) -> dict:
    """CIS iterations to eq. (59) convergence (or max_iter); writes history.csv and summary.json
    into out_dir (default runs/kinetic/<case name>) and returns the summary. backend_factory(setup)
    gives an external backend (an ops table, or an object holding one as .ops, kept alive for the
    run); the HIP backend when None.

    A file named STOP in out_dir ends the run after the current iteration, writing the results of
    the current f and cells.csv (src/app/cmd_cis_run.cpp:324-343). checkpoint_every > 0 writes f
    (original cell numbering, velocity-grid order, float64) to out_dir/checkpoint.npy every that many
    iterations and when the run stops; restart_from starts from such a file (an extension: the
    reference restarts from cells.csv, init_from here)."""
    if min_iter < 1:
        raise ValueError("min_iter must be at least 1")
    if max_iter is not None and max_iter < 1:
        raise ValueError("max_iter must be at least 1")
    ov = dict(overrides or {})
    if precision is not None:
        ov["precision"] = precision
    t_total = time.perf_counter()
    setup = Setup(case_file, ov or None)
    desc = setup.describe()
    t_setup = time.perf_counter()
    external = backend_factory(setup) if backend_factory is not None else None
    solver = Solver(setup, getattr(external, "ops", external))
    setup_s = time.perf_counter() - t_setup
    params = desc["params"]
    max_iter = int(max_iter if max_iter is not None else desc["max_iterations"])
    tol = float(tol if tol is not None else desc["tolerance"])
    out_dir = out_dir or os.path.join("runs", "kinetic", desc["case"])
    os.makedirs(out_dir, exist_ok=True)
    if restart_from:
        f0 = np.load(restart_from, allow_pickle=False)
        if f0.shape != (setup.n_cells, setup.nv):
            raise ValueError(f"{restart_from} holds {f0.shape}, the case has {(setup.n_cells, setup.nv)}")
        solver.set_distribution(f0)
    elif init_from:
        solver.initialize_equilibrium(read_cells_csv(init_from, setup.n_cells))
    else:
        solver.initialize_uniform(params["initial_pressure"] / params["initial_temperature"], (0.0, 0.0, 0.0),
                                  params["initial_temperature"])
    backend = solver.backend()
    memory_bytes = backend.memory_bytes()
    history = HistoryWriter(os.path.join(out_dir, "history.csv"), desc,
                            memory_bytes if params["backend"] == "cuda" else 0)
    sec_pos = list(desc["sections"]["positions"])
    sec_axis = "xyz".index(desc["sections"]["axis"])

    def section_flows():
        return list(solver.section_mass_flow(sec_pos, sec_axis)) if sec_pos else []

    def checkpoint():
        tmp = os.path.join(out_dir, "checkpoint.tmp.npy")
        np.save(tmp, solver.distribution().reshape(setup.n_cells, setup.nv))
        os.replace(tmp, os.path.join(out_dir, "checkpoint.npy"))

    converged = stopped = False
    timing_sum = dict.fromkeys(TIMING_KEYS, 0.0)
    timed, iter_time_sum = 0, 0.0
    last = None
    try:
        while True:
            prep = solver.prepare()
            flows = solver.patch_mass_flow()
            sec = section_flows()
            mb = mass_balance(flows)
            balance_ok = (not desc["mass_balance_required"] or not mb["valid"]
                          or abs(mb["imbalance"]) < desc["mass_balance_tolerance"])
            if (prep["iteration"] >= min_iter and prep["has_epsilon"] and prep["epsilon"] < tol and balance_ok
                    and mass_flow_targets_met(desc, flows)):
                converged, last = True, prep
                history.write(prep, flows, time.perf_counter() - t_total, sec)
                break
            if prep["iteration"] >= max_iter:
                last = prep
                history.write(prep, flows, time.perf_counter() - t_total, sec)
                break
            info = solver.sweep()
            history.write(info, flows, time.perf_counter() - t_total, sec)
            if info["iteration"] >= 1:  # the first iteration, right after initialisation, is not timed
                for k in TIMING_KEYS:
                    timing_sum[k] += info["timings"][k]
                iter_time_sum += info["seconds"]
                timed += 1
            if print_every > 0 and (info["iteration"] % print_every == 0 or info["iteration"] < 3):
                log(f"iteration {info['iteration']:6d}  eps {info['epsilon'] if info['has_epsilon'] else 0.0:.3e}  "
                    f"residual {info['residual']:.3e}  {info['seconds']:.3f} s")
            if checkpoint_every > 0 and (info["iteration"] + 1) % checkpoint_every == 0:
                checkpoint()
            if os.path.exists(os.path.join(out_dir, "STOP")):  # a gentle stop with the current f
                stopped = True
                last = solver.prepare()
                history.write(last, solver.patch_mass_flow(), time.perf_counter() - t_total, section_flows())
                log(f"STOP found: stopping at iteration {last['iteration']}")
                break
    finally:
        history.close()
    wall_s = time.perf_counter() - t_total

    ps = summarize_patches(desc, setup.bface_geometry(), solver.boundary_fluxes())
    sec_final = section_flows()
    if cells_csv or stopped:
        write_cells_csv(os.path.join(out_dir, "cells.csv"), setup, solver, desc)
    if checkpoint_every > 0:
        checkpoint()
    scale = 1.0 / timed if timed > 0 else 0.0
    kt = {k: timing_sum[k] * scale for k in TIMING_KEYS}
    kt["total"] = sum(timing_sum[k] for k in TIMING_KEYS) * scale
    eps = last["has_epsilon"]
    root = {
        "program": lib().gsk_version().decode(),
        "case": desc["case"], "case_file": str(case_file), "mesh": desc["mesh"],
        "n_cells": desc["n_cells"], "n_boundary_faces": desc["n_boundary_faces"], "n_velocity": desc["n_velocity"],
        "Kn": desc["Kn"], "delta": desc["delta"],
        "iterations": int(last["iteration"]), "converged": converged, "stopped": stopped, "stopped_by_user": stopped,
        "tolerance": tol,
        "epsilon": last["epsilon"] if eps else None, "eps_rho": last["eps_rho"] if eps else None,
        "eps_u": last["eps_u"] if eps else None, "eps_T": last["eps_T"] if eps else None,
        "wall_time_s": wall_s, "setup_time_s": setup_s,
        "time_per_iteration_s": iter_time_sum / timed if timed > 0 else 0.0,
        "kernel_time_per_iteration_s": kt,
        "backend": backend.name, "precision": params["precision"], "memory_bytes": int(memory_bytes),
        "velocity_chunk": int(backend.chunk_width()), "sweep_blocks": desc["sweep_blocks"],
        "cell_ordering": params["cell_ordering"], "block_cut_faces": desc["block_cut_faces"],
        "params": {k: params[k] for k in (
            "pseudo_cfl", "pseudo_time", "limiter", "limiter_k", "limiter_eps", "lsq_weight_power", "cell_ordering",
            "initial_pressure", "initial_temperature", "convergence_u_floor", "wall_relaxation", "boundary_iterations",
            "interface_damping", "positivity_limiter", "sweep_order", "inner_sweeps", "inner_wall_update",
            "block_passes", "equilibrium", "deq_fallback", "pressure_boundary_reconstruction",
            "cold_bgk_temperature")},
        "reference": dict(desc["reference"]),
        "patches": ps["patches"],
    }
    mf = tube_mass_flow(desc, ps)
    if mf is not None:
        root["mass_flow"] = mf
    root["mass_balance"] = balance_record(desc, ps["mdot"], sec_final)
    if sec_pos:
        root["mass_flow_sections"] = sections_record(desc, solver.section_cuts(sec_pos, sec_axis), sec_final)
    root["residual"] = last["residual"] if math.isfinite(last["residual"]) else 0.0
    with open(os.path.join(out_dir, "summary.json"), "w") as f:
        f.write(dump_json(root) + "\n")
    solver.close()
    return root


# ------------------------------------------------------------------ in-memory loop (benchmarks)


def run(solver, mdot_factor: float, inlet: int | None, outlet: int | None, tol: float = 0.0,
        residual_target: float = 0.0, max_iterations: int = 1000, init=(1.0, (0.0, 0.0, 0.0), 1.0),
        log=None) -> dict:
    """The same loop without files, for timing: prepare -> stop on eq. (59) epsilon below tol (after
    iteration 0) or on a residual target -> sweep; returns arrays per iteration (iteration,
    epsilon, residual, inlet / outlet Mdot, wall time) and the stop reason. mdot_factor =
    Setup.mass_flow_factor() (nondimensional patch flow -> Mdot, cmd_cis_run.cpp:451-455); inlet flows
    are outward-negative, so Mdot_in = -flow * factor."""
    rho, u, T = init
    solver.initialize_uniform(rho, u, T)
    hist = {k: [] for k in ("iteration", "epsilon", "residual", "mdot_in", "mdot_out", "t_wall")}
    t0 = time.perf_counter()
    reason = "max_iterations"
    for it in range(max_iterations + 1):
        info = solver.prepare()
        flows = solver.patch_mass_flow()
        if it > 0 and tol > 0.0 and info["has_epsilon"] and info["epsilon"] < tol:
            reason = "epsilon"
            break
        if it == max_iterations:
            break
        info = solver.sweep()
        hist["iteration"].append(it)
        hist["epsilon"].append(info["epsilon"] if info["has_epsilon"] else np.nan)
        hist["residual"].append(info["residual"])
        hist["mdot_in"].append(-flows[inlet] * mdot_factor if inlet is not None else np.nan)
        hist["mdot_out"].append(flows[outlet] * mdot_factor if outlet is not None else np.nan)
        hist["t_wall"].append(time.perf_counter() - t0)
        if log is not None and (it % 10 == 0):
            log(f"iteration {it}: residual {info['residual']:.3e} Mdot_in {hist['mdot_in'][-1]:.6f}")
        if residual_target > 0.0 and info["residual"] <= residual_target:
            reason = "residual"
            break
    out = {k: np.array(v) for k, v in hist.items()}
    out["reason"] = reason
    return out


def time_to_within(hist: dict, target: float, rel: float = 1e-4, key: str = "mdot_in"):
    """(wall time, iteration) after which |Mdot - target| <= rel |target| holds for every later
    iteration (None if it does not hold at the end)."""
    x = hist[key]
    ok = np.abs(x - target) <= rel * abs(target)
    if not ok[-1]:
        return None
    bad = np.nonzero(~ok)[0]
    first = 0 if bad.size == 0 else bad[-1] + 1
    return float(hist["t_wall"][first]), int(hist["iteration"][first])


def main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(description="CIS-only steady solve (the reference's gsis-solver cis-run)")
    ap.add_argument("case")
    ap.add_argument("--out")
    ap.add_argument("--max-iter", type=int)
    ap.add_argument("--min-iter", type=int, default=1)
    ap.add_argument("--tol", type=float)
    ap.add_argument("--precision", choices=["fp32", "fp64"])
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE", help="[kinetic] override")
    ap.add_argument("--print-every", type=int, default=100)
    ap.add_argument("--cells-csv", action="store_true")
    ap.add_argument("--init-from", help="restart from the cell values of a previous cells.csv")
    ap.add_argument("--checkpoint-every", type=int, default=0, help="write f to <out>/checkpoint.npy every N iterations")
    ap.add_argument("--restart-from", help="start from f in a checkpoint.npy of a previous run")
    a = ap.parse_args(argv)
    ov = {}
    for kv in a.set:
        k, _, v = kv.partition("=")
        if not k or not _:
            ap.error(f"--set expects key=value, got {kv!r}")
        ov[k.strip()] = _parse_value(v.strip())
    s = cis_run(a.case, a.out, a.max_iter, a.min_iter, a.tol, a.precision, ov, a.print_every, a.cells_csv,
                init_from=a.init_from, checkpoint_every=a.checkpoint_every, restart_from=a.restart_from)
    print(f"iterations {s['iterations']} ({'converged' if s['converged'] else 'limit reached'}), "
          f"{s['wall_time_s']:.2f} s")
    return 0


def _parse_value(s: str):
    """--set values: integer, then real, then true / false, else a string (cmd_cis_run.cpp:146-164)."""
    try:
        return int(s)
    except ValueError:
        pass
    try:
        return float(s)
    except ValueError:
        pass
    if s in ("true", "false"):
        return s == "true"
    return s[1:-1] if len(s) >= 2 and s[0] == s[-1] == '"' else s


if __name__ == "__main__":
    sys.exit(main())
