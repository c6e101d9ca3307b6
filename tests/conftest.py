"""Shared fixtures: the package (loaded from gsis-sim_amd/), the oracle backends (oracle/), and
small programmatic cases in the style of the reference's make_kinetic_case / make_box_mesh helpers
(tests/kinetic/kinetic_test_cases.cpp:9-60, tests/support/box_mesh.cpp of the reference)."""

from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "gsis-sim_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_package():
    if "gsis_sim_amd" in sys.modules:
        return sys.modules["gsis_sim_amd"]
    spec = importlib.util.spec_from_file_location("gsis_sim_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gsis_sim_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def gsk():
    return load_package()


@pytest.fixture(scope="session")
def meshgen(gsk):
    import gsis_sim_amd.meshgen as mg  # noqa: F401  (registered by load_package)
    return sys.modules["gsis_sim_amd.meshgen"]


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    if not os.path.exists(oracle.ORACLE_LIB):
        oracle.build()
    return oracle


T0 = 273.15


def box_case(mg, directory, n=(5, 3, 2), distort=0.25, seed=7, n_vel=8, vrange=5.0, kinetic=None, bcs="mixed",
             delta=1.0, name="box", wall_T=(0.95, 1.05), stretch=0.0, tet=False):
    """Distorted box with every boundary kind the kinetic path handles:
    xmin pressure inlet, xmax pressure outlet (p > 0), ymin moving diffuse wall, ymax Maxwell wall
    (alpha = 0.8, axis aligned -> mirror shortcut), zmin / zmax symmetry. tet: the same box split
    into tetrahedra (six per hexahedron)."""
    mesh = (mg.tet_box_mesh if tet else mg.box_mesh)(*n, distort=distort, seed=seed)
    mg.write_pmesh(os.path.join(directory, f"{name}.pmesh"), mesh)
    if bcs == "mixed":
        bnd = [dict(patch="xmin", type="pressure_inlet", temperature=1.1 * T0),
               dict(patch="xmax", type="pressure_outlet", pressure=1.0),  # Pa, tiny vs p_ref -> nearly vacuum-like
               dict(patch="ymin", type="diffuse_wall", temperature=0.9 * T0, velocity=[20.0, 0.0, 0.0]),
               dict(patch="ymax", type="maxwell_wall", accommodation=0.8, temperature=1.05 * T0),
               dict(patch="zmin", type="symmetry"), dict(patch="zmax", type="symmetry")]
    elif bcs == "closed":
        bnd = [dict(patch="xmin", type="diffuse_wall", temperature=wall_T[0] * T0),
               dict(patch="xmax", type="diffuse_wall", temperature=wall_T[1] * T0),
               dict(patch="ymin", type="diffuse_wall", temperature=T0), dict(patch="ymax", type="diffuse_wall", temperature=T0),
               dict(patch="zmin", type="symmetry"), dict(patch="zmax", type="symmetry")]
    elif bcs == "vacuum":
        bnd = [dict(patch="xmin", type="pressure_inlet", temperature=T0), dict(patch="xmax", type="vacuum"),
               dict(patch="ymin", type="diffuse_wall", temperature=T0), dict(patch="ymax", type="diffuse_wall", temperature=T0),
               dict(patch="zmin", type="symmetry"), dict(patch="zmax", type="symmetry")]
    else:
        raise ValueError(bcs)
    path = os.path.join(directory, f"{name}.toml")
    mg.write_case(path, f"{name}.pmesh", name, bnd, [n_vel] * 3, [(-vrange, vrange)] * 3,
                  dict(L_ref=1.0e-3, T0=T0, delta_rp=float(delta)), kinetic, stretch=stretch)
    return path


def perturbed_distribution(setup, seed=1, amp=0.05):
    """Local Maxwellians with varying (rho, u, T) times (1 + amp * noise): an f with nonzero
    gradients, heat flux and stress, in the solver's internal numbering and storage order."""
    rng = np.random.default_rng(seed)
    v = setup.view()
    nv = v.nv
    xi = np.stack([np.ctypeslib.as_array(v.xi_x, (nv,)), np.ctypeslib.as_array(v.xi_y, (nv,)),
                   np.ctypeslib.as_array(v.xi_z, (nv,))], 1)
    nc = setup.n_cells
    rho = rng.uniform(0.7, 1.3, nc)
    T = rng.uniform(0.7, 1.3, nc)
    u = rng.uniform(-0.3, 0.3, (nc, 3))
    c2 = ((xi[None, :, :] - u[:, None, :]) ** 2).sum(-1)
    f = rho[:, None] / (2 * np.pi * T[:, None]) ** 1.5 * np.exp(-c2 / (2 * T[:, None]))
    f *= 1.0 + amp * rng.uniform(-1.0, 1.0, f.shape)
    return f


def random_inflow(setup, seed=2):
    rng = np.random.default_rng(seed)
    nb = setup.n_bfaces
    inflow = np.zeros((nb, 6))
    inflow[:, 0] = rng.uniform(0.5, 1.5, nb)
    inflow[:, 1] = rng.uniform(0.8, 1.2, nb)
    inflow[:, 2:5] = rng.uniform(-0.2, 0.2, (nb, 3))
    inflow[:, 5] = rng.uniform(0.8, 1.2, nb)
    return inflow


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / max(scale, floor)) if a.size else 0.0


# The reference's slab verification cases (cases/verification/kinetic/make_cases.py): walls at
# T_L/T0, T_R/T0 moving with u_y = -/+ u sqrt(R T0); sides all periodic (the `_per` runs) or, by
# default, symmetry planes (heat transfer) / sym_y periodic and sym_z a symmetry plane (Couette flow);
# 30 x 24 x 24 velocities on [-10, 10] x [-8, 8]^2, Kn from eq. (8), one sweep block.
SLAB_KINDS = {"heat_lin": (0.95, 1.05, 0.0), "heat_nl": (0.5, 1.5, 0.0), "couette_lin": (1.0, 1.0, 0.1),
              "couette_nl": (1.0, 1.0, 1.0)}


def slab_verification_case(mg, directory, kind, kn, nx, precision="fp64", sides="periodic"):
    c0 = (1.380649e-23 / 6.65e-27 * T0) ** 0.5
    tl, tr, u = SLAB_KINDS[kind]
    mesh_file = f"slab_nx{nx}.pmesh"
    mg.write_pmesh(os.path.join(directory, mesh_file), mg.slab_mesh(nx), {"generator": "gsis-sim_amd.meshgen.slab_mesh"})
    walls = []
    for patch, t, sgn in (("wall_left", tl, -1.0), ("wall_right", tr, 1.0)):
        w = dict(patch=patch, type="diffuse_wall", temperature=t * T0)
        if u != 0.0:
            w["velocity"] = [0.0, sgn * u * c0, 0.0]
        walls.append(w)
    if sides == "periodic":
        sy, sz = "periodic_slab", "periodic_slab"
    else:
        sy, sz = ("periodic_slab" if kind.startswith("couette") else "symmetry"), "symmetry"
    bnd = walls + [dict(patch="sym_y", type=sy), dict(patch="sym_z", type=sz)]
    name = f"slab_{kind}_Kn{kn:g}_nx{nx}_{sides}_{precision}".replace(".", "p")
    path = os.path.join(directory, name + ".toml")
    mg.write_case(path, mesh_file, name, bnd, [30, 24, 24], [(-10.0, 10.0), (-8.0, 8.0), (-8.0, 8.0)],
                  dict(L_ref=1.0e-3, T0=T0, Kn=float(kn)), dict(sweep_blocks=1, precision=precision))
    return path


def slab_reference_rows():
    """The reference's converged periodic slab runs (tests/golden/slab_reference_results.csv)."""
    import csv
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "slab_reference_results.csv")) as f:
        return {r["case"]: r for r in csv.DictReader(f)}


def slab_mean_flux(solver, setup, kind):
    """Volume-weighted mean over the cells of P_xy (Couette) or q_x (heat transfer): the value the
    reference's compare_ref1d.py takes from cells.csv."""
    m = solver.moments()
    vol = np.empty(setup.n_cells)
    vol[setup.cell_new_to_old()] = np.ctypeslib.as_array(setup.view().volume, shape=(setup.n_cells,))
    w = vol / vol.sum()
    col = 9 if kind.startswith("couette") else 12  # sigma_xy, q_x in the 15-moment row
    return float(np.sum(w * m[:, col]))


def kinetic_case(mg, directory, mesh, bcs, Kn=0.5, n_vel=10, vmax=5.0, kinetic=None, name="kcase"):
    """The reference's make_kinetic_case (tests/kinetic/kinetic_test_cases.cpp:9-60): L_ref = 1 m,
    T0 = 273.15 K, helium, the given Kn, n_vel^3 velocities on [-vmax, vmax]^3; boundary values in
    reference units (T / T0, p / p_ref, u / sqrt(R T0)), converted with the case's own scales."""
    mg.write_pmesh(os.path.join(directory, f"{name}.pmesh"), mesh, {"generator": "tests/conftest.py"})
    path = os.path.join(directory, f"{name}.toml")

    def write(p_ref, v_ref):
        out = []
        for b in bcs:
            d = dict(patch=b["patch"], type=b["type"])
            if b["type"] == "diffuse_wall":
                d["temperature"] = b.get("T", 1.0) * T0
                u = b.get("u", (0.0, 0.0, 0.0))
                if any(x != 0.0 for x in u):
                    d["velocity"] = [x * v_ref for x in u]
            elif b["type"] == "pressure_inlet":
                d["temperature"] = b.get("T", 1.0) * T0
                d["pressure"] = b.get("p", 1.0) * p_ref
            elif b["type"] == "pressure_outlet":
                d["pressure"] = b.get("p", 1.0) * p_ref
            out.append(d)
        mg.write_case(path, f"{name}.pmesh", name, out, [n_vel] * 3, [(-vmax, vmax)] * 3,
                      dict(L_ref=1.0, T0=T0, Kn=float(Kn)), kinetic)

    write(1.0, 1.0)  # probe for the reference scales
    import gsis_sim_amd as g  # registered by load_package
    ref = g.Setup(path).describe()["reference"]
    write(ref["pressure_Pa"], ref["velocity_m_s"])
    return path


def tube_coarse_case(mg, directory, delta, kinetic=None, name=None):
    """The reference's coarse tube-to-vacuum case (cases/tube_vacuum/kinetic/tube_coarse_delta{1,10}_cis.toml)
    on its mesh (tests/golden/tube_hex_coarse.npz, written back as a .pmesh): helium, L_ref = R = 1 mm,
    T0 = 273.15 K, pressure inlet at (p_in, T0), vacuum outlet, diffuse wall at T0, 30 x 24 x 24
    velocities on [-10, 10] x [-8, 8]^2, sweep_blocks = 0 (automatic), the given [kinetic] settings."""
    import json

    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tube_hex_coarse.npz"))
    mesh = dict(points=g["points"], face_nodes=g["face_nodes"], owner=g["owner"], neighbour=g["neighbour"],
                n_cells=int(g["n_cells"]), patches=json.loads(str(g["patches"])))
    mg.write_pmesh(os.path.join(directory, "tube_hex_coarse.pmesh"), mesh, {"generator": "tests/golden/tube_hex_coarse.npz"})
    name = name or f"tube_coarse_delta{delta:g}_cis"
    path = os.path.join(directory, name + ".toml")
    bnd = [dict(patch="inlet", type="pressure_inlet", temperature=T0), dict(patch="outlet", type="vacuum"),
           dict(patch="wall", type="diffuse_wall", temperature=T0)]
    mg.write_case(path, "tube_hex_coarse.pmesh", name, bnd, [30, 24, 24], [(-10.0, 10.0), (-8.0, 8.0), (-8.0, 8.0)],
                  dict(L_ref=1.0e-3, T0=T0, delta_rp=float(delta)), dict({"sweep_blocks": 0}, **(kinetic or {})))
    return path
