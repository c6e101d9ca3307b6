"""Digests of everything the kinetic setup hands the backend and the solver: the velocity layout,
the RCB / RCM cell numbering, LU-SGS blocks and per-direction sweep orders, the CSR face entries,
boundary-face table, per-cell scalars and the case description. Used by
tests/golden/make_setup_fixture.py (which records them) and tests/test_setup_fixture.py (which
requires a rebuilt setup to reproduce them bit for bit).

The cases cover every branch of the setup: RCB (default) / RCM / no reordering, octant / dir13 /
block sweep orders, one to 64 blocks, symmetry planes (mirror orbits in the velocity layout),
periodic slabs, axis-aligned and curved Maxwell walls, pressure inlets and outlets, vacuum,
moving walls, the pseudo-time step, tetrahedra, stretched velocity grids, and the full-size tubes."""

from __future__ import annotations

import ctypes as C
import hashlib
import json
import os

import conftest


def _raw(ptr, count: int, ctype) -> bytes:
    if count == 0 or not ptr:
        return b""
    return C.string_at(ptr, count * C.sizeof(ctype))


def setup_digests(g, setup) -> dict:
    v = setup.view()
    nc, nv, nb = v.n_cells, v.nv, v.n_bfaces
    n_entries = v.csr_offsets[nc]
    tables = {
        "xi": _raw(v.xi_x, nv, C.c_double) + _raw(v.xi_y, nv, C.c_double) + _raw(v.xi_z, nv, C.c_double),
        "csr_offsets": _raw(v.csr_offsets, nc + 1, C.c_int32),
        "csr_entries": _raw(v.csr, n_entries, g.CsrEntry),
        "eps2": _raw(v.eps2, nc, C.c_double),
        "volume": _raw(v.volume, nc, C.c_double),
        "vdt": _raw(v.vdt, nc, C.c_double),
        "bfaces": _raw(v.bfaces, nb, g.BFace),
        "block_offsets": _raw(v.block_offsets, v.n_blocks + 1, C.c_int32),
        "sweep_order": _raw(v.sweep_order, v.n_sweep_dirs * nc, C.c_int32),
        "sweep_pos": _raw(v.sweep_pos, v.n_sweep_dirs * nc, C.c_int32),
        "sweep_dir_of": _raw(v.sweep_dir_of, nv if v.n_sweep_dirs > 0 else 0, C.c_uint8),
        "grid_to_storage": _raw(v.grid_to_storage, nv, C.c_int32),
        "maxwell_offset": _raw(v.maxwell_offset, nb if v.n_maxwell_faces > 0 else 0, C.c_int64),
        "maxwell_specular": _raw(v.maxwell_specular, nb if v.n_maxwell_faces > 0 else 0, C.c_double),
        "weight_rel": _raw(v.weight_rel, nv, C.c_double),
        "grid_nodes": _raw(v.grid_nodes, sum(v.grid_n), C.c_double),
        "cell_new_to_old": setup.cell_new_to_old().tobytes(),
        "storage_to_grid": setup.storage_to_grid().tobytes(),
        "cell_centers": setup.cell_centers().tobytes(),
        "bface_geometry": b"".join(a.tobytes() for a in setup.bface_geometry()),
    }
    scalars = {name: (list(getattr(v, name)) if name in ("grid_n", "deq_lo", "deq_hi", "deq_d") else getattr(v, name))
               for name, ctype in type(v)._fields_ if not hasattr(ctype, "contents")}
    scalars["mass_flow_factor"] = setup.mass_flow_factor()
    desc = setup.describe()
    desc.pop("mesh", None)  # a path in the temporary directory
    out = {k: hashlib.sha256(b).hexdigest()[:32] for k, b in tables.items()}
    out["scalars"] = hashlib.sha256(json.dumps(scalars, sort_keys=True).encode()).hexdigest()[:32]
    out["describe"] = hashlib.sha256(json.dumps(desc, sort_keys=True).encode()).hexdigest()[:32]
    out["sizes"] = [nc, nv, nb, v.n_blocks, int(n_entries)]
    return out


def _box(kinetic, **kw):
    return lambda mg, d: conftest.box_case(mg, d, kinetic=kinetic, **kw)


# name -> case builder (meshgen, directory) -> toml path
CASES = {
    "slab200_24cubed": lambda mg, d: mg.slab_case(d, kinetic={"precision": "fp64"}),
    "slab_heat_nl_symmetry": lambda mg, d: conftest.slab_verification_case(mg, d, "heat_nl", 1.0, 50, sides="symmetry"),
    "slab_couette_periodic": lambda mg, d: conftest.slab_verification_case(mg, d, "couette_nl", 0.1, 100),
    "box_mixed": _box({"precision": "fp64"}),
    "box_mixed_rcm_dir13_3blocks": _box({"precision": "fp64", "cell_ordering": "rcm", "sweep_order": "dir13",
                                          "sweep_blocks": 3}),
    "box_vacuum_none_block_order": _box({"precision": "fp32", "cell_ordering": "none", "sweep_order": "block",
                                          "sweep_blocks": 2}, bcs="vacuum"),
    "box_closed_pseudo_time": _box({"precision": "fp64", "pseudo_cfl": 5.0, "limiter_eps": "absolute",
                                     "limiter_k": 2.0}, bcs="closed", n=(6, 4, 3), seed=3),
    "box_closed_velocity_pseudo_time": _box({"precision": "fp64", "pseudo_cfl": 2.0, "pseudo_time": "velocity"},
                                             bcs="closed"),
    "box_mixed_stretched_7blocks": _box({"precision": "fp64", "sweep_blocks": 7}, n=(8, 5, 4), stretch=1.5),
    "tet_box_mixed": _box({"precision": "fp64", "sweep_blocks": 3}, tet=True),
    "tube_curved_maxwell_stretched": lambda mg, d: mg.tube_case(d, n_core=4, n_ring=2, n_axial=3, n_vel=8, vrange=5.0,
                                                                 delta=5.0, accommodation=0.9, stretch=1.2,
                                                                 kinetic={"precision": "fp64"}),
    "tube_small_rcm_dir13": lambda mg, d: mg.tube_case(d, n_core=6, n_ring=3, n_axial=8, n_vel=10, delta=2.0,
                                                        kinetic={"precision": "fp64", "cell_ordering": "rcm",
                                                                 "sweep_order": "dir13"}),
    "tube40k_24cubed_tanh": lambda mg, d: mg.tube_case(d, n_core=16, n_ring=10, n_axial=45, n_vel=24, delta=10.0,
                                                        stretch=1.5, kinetic={"precision": "fp64"}),
    "tube135k_32cubed_tanh": lambda mg, d: mg.tube_case(d, n_core=24, n_ring=14, n_axial=70, n_vel=32, delta=1.0,
                                                         stretch=1.5, kinetic={"precision": "fp32",
                                                                               "moment_sums": "mfma"}),
}

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "setup_digests.json")


def build_digests(g, mg, directory, names=None) -> dict:
    out = {}
    for name, make in CASES.items():
        if names is not None and name not in names:
            continue
        d = os.path.join(directory, name)
        os.makedirs(d, exist_ok=True)
        out[name] = setup_digests(g, g.Setup(make(mg, d)))
    return out
