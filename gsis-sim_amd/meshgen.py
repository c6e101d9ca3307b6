"""Deterministic hexahedral mesh generators and case writers for the kinetic benchmarks.

The reference ships its meshes as .pmesh files made by its own Python tools; this package has to
make the same kinds of meshes on a machine that only has this repository, so it writes them itself:

  slab_mesh  -- 1D planar slab: nx cubic cells in a row, 1 cell thick, patches wall_left,
                wall_right, sym_y, sym_z (same layout as cases/verification/kinetic/mesh/slab_nx200.pmesh)
  tube_mesh  -- O-grid hexahedral cylinder (axis x, radius R, length L): a square core block and four
                ring blocks mapped to the circle; patches inlet (x = 0), outlet (x = L), wall

Output is the .pmesh v1 format (docs/pmesh_format.md §1-§3 of the reference): internal faces first in
(owner, neighbour) order with owner < neighbour and the normal pointing owner -> neighbour; boundary
faces grouped by patch, normals outward.
"""

from __future__ import annotations

import json
import struct

import numpy as np

HEX_FACES = np.array([[0, 2, 6, 4], [1, 3, 7, 5], [0, 1, 5, 4], [2, 3, 7, 6], [0, 1, 3, 2], [4, 5, 7, 6]])
TET_FACES = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]])


def hex_to_poly(points: np.ndarray, cells: np.ndarray, patch_of_face, patch_names, patch_kinds, faces=HEX_FACES):
    """Convert cells of one kind to the .pmesh face layout: hexahedra (8 corners, corner index bits =
    (i, j, k)) by default, or any cell whose faces `faces` lists as cyclic corner sequences of one
    length (TET_FACES for tetrahedra). Faces are oriented from owner to neighbour / outward.

    patch_of_face(centers, normals) -> int patch index per boundary face.
    Returns dict(points, face_nodes (n_faces, face size), owner, neighbour, patches).
    """
    points = np.asarray(points, np.float64)
    cells = np.asarray(cells, np.int64)
    nc = len(cells)
    cyc = np.asarray(faces)
    nfc, m = cyc.shape
    quads = cells[:, cyc].reshape(-1, m)  # (faces per cell x nc, face size)
    owner_cell = np.repeat(np.arange(nc), nfc)
    key = np.sort(quads, axis=1)
    uniq, inv, counts = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    inv = inv.ravel()
    assert counts.max() <= 2, "a face is shared by more than two cells"
    order = np.argsort(inv, kind="stable")
    sorted_inv = inv[order]
    first = np.ones(len(order), bool)
    first[1:] = sorted_inv[1:] != sorted_inv[:-1]
    # internal faces: pairs
    pair_start = np.nonzero(first)[0]
    n_per = np.diff(np.append(pair_start, len(order)))
    internal_groups = pair_start[n_per == 2]
    bnd_groups = pair_start[n_per == 1]
    a = order[internal_groups]
    b = order[internal_groups + 1]
    ca, cb = owner_cell[a], owner_cell[b]
    own = np.minimum(ca, cb)
    nbr = np.maximum(ca, cb)
    # take the quad as seen from the owner
    qi = np.where(ca == own, a, b)
    int_quads = quads[qi]
    # boundary faces
    bq_idx = order[bnd_groups]
    b_quads = quads[bq_idx]
    b_owner = owner_cell[bq_idx]

    cell_center = points[cells].mean(axis=1)

    def orient(q, c_from, toward_other=None):
        p = points[q]
        xbar = p.mean(axis=1)
        s = np.zeros_like(xbar)
        for k in range(m):
            s += np.cross(p[:, k] - xbar, p[:, (k + 1) % m] - xbar)
        fc = xbar
        ref = fc - cell_center[c_from] if toward_other is None else cell_center[toward_other] - cell_center[c_from]
        flip = np.einsum("ij,ij->i", s, ref) < 0
        q = q.copy()
        q[flip] = q[flip][:, ::-1]
        s[flip] = -s[flip]
        return q, s, fc

    int_quads, _, _ = orient(int_quads, own, nbr)
    b_quads, b_s, b_fc = orient(b_quads, b_owner)
    io = np.lexsort((nbr, own))
    int_quads, own, nbr = int_quads[io], own[io], nbr[io]
    normals = b_s / np.linalg.norm(b_s, axis=1)[:, None]
    patch = np.asarray(patch_of_face(b_fc, normals), np.int64)
    bo = np.lexsort((b_owner, patch))
    b_quads, b_owner, patch = b_quads[bo], b_owner[bo], patch[bo]
    n_int = len(own)
    patches = []
    start = n_int
    for pi, name in enumerate(patch_names):
        cnt = int(np.sum(patch == pi))
        patches.append({"name": name, "kind": patch_kinds[pi], "start": start, "count": cnt})
        start += cnt
    face_nodes = np.concatenate([int_quads, b_quads]).astype(np.int32)
    owner = np.concatenate([own, b_owner]).astype(np.int32)
    # drop unused points and renumber
    used = np.unique(face_nodes)
    remap = -np.ones(len(points), np.int64)
    remap[used] = np.arange(len(used))
    return {
        "points": points[used],
        "face_nodes": remap[face_nodes].astype(np.int32),
        "owner": owner,
        "neighbour": nbr.astype(np.int32),
        "n_cells": nc,
        "patches": patches,
    }


def write_pmesh(path: str, mesh: dict, metadata: dict | None = None) -> None:
    """Write a .pmesh v1 file (faces of one size: quads or triangles)."""
    pts = np.ascontiguousarray(mesh["points"], "<f8")
    fn = np.ascontiguousarray(mesh["face_nodes"], "<i4")
    nf, m = fn.shape
    fo = np.arange(0, m * nf + 1, m, dtype="<i8")
    owner = np.ascontiguousarray(mesh["owner"], "<i4")
    nbr = np.ascontiguousarray(mesh["neighbour"], "<i4")
    arrays = [("points", pts, "float64", [len(pts), 3]), ("face_offsets", fo, "int64", [nf + 1]),
              ("face_nodes", fn.ravel(), "int32", [fn.size]), ("owner", owner, "int32", [nf]),
              ("neighbour", nbr, "int32", [len(nbr)])]

    def header(base):
        off = (base + 63) // 64 * 64
        desc = {}
        for name, arr, dt, shape in arrays:
            desc[name] = {"dtype": dt, "shape": shape, "offset": off}
            off = (off + arr.nbytes + 63) // 64 * 64
        h = {"format": "pmesh", "version": 1, "length_unit": "m", "n_points": len(pts), "n_faces": nf,
             "n_internal_faces": len(nbr), "n_cells": int(mesh["n_cells"]), "patches": mesh["patches"],
             "arrays": desc, "metadata": {k: str(v) for k, v in (metadata or {}).items()}}
        return json.dumps(h).encode()

    hdr = header(16)
    for _ in range(8):
        nxt = header(16 + len(hdr))
        if len(nxt) == len(hdr):
            hdr = nxt
            break
        hdr = nxt
    h = json.loads(hdr)
    total = max(d["offset"] + a.nbytes for (n, a, _, _), d in zip(arrays, h["arrays"].values()))
    buf = bytearray(total)
    buf[0:8] = b"PMESH\0\0\0"
    buf[8:16] = struct.pack("<Q", len(hdr))
    buf[16:16 + len(hdr)] = hdr
    for name, arr, _, _ in arrays:
        o = h["arrays"][name]["offset"]
        buf[o:o + arr.nbytes] = arr.tobytes()
    with open(path, "wb") as f:
        f.write(bytes(buf))


def slab_mesh(nx: int = 200, length: float = 1.0e-3, thickness: float | None = None) -> dict:
    """nx cells on [0, length] in x, one cell of `thickness` (default length/nx) in y and z."""
    h = length / nx if thickness is None else thickness
    xs = np.linspace(0.0, length, nx + 1)
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(2), np.arange(2), indexing="ij")
    pts = np.stack([xs[I.ravel()], h * J.ravel(), h * K.ravel()], axis=1)
    nid = lambda i, j, k: (i * 2 + j) * 2 + k  # noqa: E731
    cells = np.array([[nid(i + (c & 1), (c >> 1) & 1, (c >> 2) & 1) for c in range(8)] for i in range(nx)])

    def patch_of(fc, n):
        p = np.full(len(fc), -1)
        p[n[:, 0] < -0.5] = 0
        p[n[:, 0] > 0.5] = 1
        p[np.abs(n[:, 1]) > 0.5] = 2
        p[np.abs(n[:, 2]) > 0.5] = 3
        assert (p >= 0).all()
        return p

    return hex_to_poly(pts, cells, patch_of, ["wall_left", "wall_right", "sym_y", "sym_z"],
                       ["wall", "wall", "symmetry", "symmetry"])


def box_mesh(nx: int, ny: int, nz: int, size=(1.0e-3, 0.6e-3, 0.4e-3), distort: float = 0.0,
             seed: int = 0, names=None) -> dict:
    """Structured box of hexahedra, optionally with interior nodes moved by up to `distort` x the
    local spacing (seeded numpy generator; boundary nodes stay on their planes so every boundary
    face keeps an axis-aligned normal). Patches xmin, xmax, ymin, ymax, zmin, zmax, or the six
    `names` in that order (sides with the same name form one patch)."""
    pts, cells, patch_of, unique = _box_parts(nx, ny, nz, size, distort, seed, names)
    return hex_to_poly(pts, cells, patch_of, unique, ["patch"] * len(unique))


def tet_box_mesh(nx: int, ny: int, nz: int, size=(1.0e-3, 0.6e-3, 0.4e-3), distort: float = 0.0, seed: int = 0,
                 names=None) -> dict:
    """The box of box_mesh with every hexahedron split into six tetrahedra around its (0,0,0)-(1,1,1)
    diagonal (one per ordering of the axes, so neighbouring hexahedra share their face diagonals)."""
    pts, hexes, patch_of, unique = _box_parts(nx, ny, nz, size, distort, seed, names)
    tets = []
    for a, b in ((0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1)):
        ba, bb = 1 << a, 1 << b
        tets.append(hexes[:, [0, ba, ba | bb, 7]])
    tets = np.stack(tets, axis=1).reshape(-1, 4)
    return hex_to_poly(pts, tets, patch_of, unique, ["patch"] * len(unique), faces=TET_FACES)


def _box_parts(nx, ny, nz, size, distort, seed, names):
    """Points, hexahedra and the boundary-patch rule of box_mesh."""
    n = (nx, ny, nz)
    axes = [np.linspace(0.0, size[a], n[a] + 1) for a in range(3)]
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    idx = np.stack([I.ravel(), J.ravel(), K.ravel()], 1)
    pts = np.stack([axes[a][idx[:, a]] for a in range(3)], 1)
    if distort > 0.0:
        rng = np.random.default_rng(seed)
        for a in range(3):
            h = size[a] / n[a]
            interior = (idx[:, a] > 0) & (idx[:, a] < n[a])
            pts[interior, a] += distort * h * rng.uniform(-1.0, 1.0, interior.sum())
    nid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k  # noqa: E731
    cells = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                cells.append([nid(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)) for c in range(8)])
    side_names = list(names) if names is not None else ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]
    assert len(side_names) == 6
    unique = list(dict.fromkeys(side_names))
    side_patch = [unique.index(n) for n in side_names]

    def patch_of(fc, nrm):
        p = np.full(len(fc), -1)
        for a in range(3):
            p[fc[:, a] < 1e-9 * size[a]] = side_patch[2 * a]
            p[fc[:, a] > size[a] * (1 - 1e-9)] = side_patch[2 * a + 1]
        assert (p >= 0).all()
        return p

    return pts, np.array(cells), patch_of, unique


def tube_mesh(n_core: int = 16, n_ring: int = 10, n_axial: int = 45, radius: float = 1.0e-3,
              length: float = 5.0e-3, core_frac: float = 0.5, wall_stretch: float = 1.5) -> dict:
    """O-grid hex mesh of a cylinder along x.

    Cross-section: an n_core x n_core square core of half-width core_frac*R, and four ring blocks of
    n_core (around) x n_ring (radial) cells joining the square's edges to the circle, tanh-clustered
    toward the wall. Cells: (n_core^2 + 4 n_core n_ring) * n_axial.
    """
    a = core_frac * radius
    s = np.linspace(-1.0, 1.0, n_core + 1)
    # square core nodes
    Y, Z = np.meshgrid(a * s, a * s, indexing="ij")
    sec_pts = [np.stack([Y.ravel(), Z.ravel()], 1)]
    sec_cells = []
    core_id = lambda i, j: i * (n_core + 1) + j  # noqa: E731
    for i in range(n_core):
        for j in range(n_core):
            sec_cells.append([core_id(i, j), core_id(i + 1, j), core_id(i, j + 1), core_id(i + 1, j + 1)])
    base = (n_core + 1) ** 2
    # radial clustering 0..1
    t = np.linspace(0.0, 1.0, n_ring + 1)
    if wall_stretch > 0:
        t = 1.0 - np.tanh(wall_stretch * (1.0 - t)) / np.tanh(wall_stretch)
    # four sides: (edge points on the square, angles on the circle)
    sides = []
    ang = s * np.pi / 4
    sides.append((np.stack([np.full_like(s, a), a * s], 1), np.stack([np.cos(ang), np.sin(ang)], 1)))        # +y
    sides.append((np.stack([-a * s, np.full_like(s, a)], 1), np.stack([-np.sin(ang), np.cos(ang)], 1)))      # +z
    sides.append((np.stack([np.full_like(s, -a), -a * s], 1), np.stack([-np.cos(ang), -np.sin(ang)], 1)))    # -y
    sides.append((np.stack([a * s, np.full_like(s, -a)], 1), np.stack([np.sin(ang), -np.cos(ang)], 1)))      # -z
    for edge, circ in sides:
        circ = radius * circ
        blk = np.concatenate([(1 - tt) * edge + tt * circ for tt in t])  # (n_ring+1)*(n_core+1)
        sec_pts.append(blk)
        rid = lambda r, i, b=base: b + r * (n_core + 1) + i  # noqa: E731
        for r in range(n_ring):
            for i in range(n_core):
                sec_cells.append([rid(r, i), rid(r, i + 1), rid(r + 1, i), rid(r + 1, i + 1)])
        base += (n_ring + 1) * (n_core + 1)
    sec = np.concatenate(sec_pts)
    # merge coincident section nodes (square edge nodes are shared with the ring blocks)
    key = np.round(sec / (radius * 1e-9)).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.ravel()
    sec_u = sec[first]
    sec_cells = inv[np.array(sec_cells)]
    ns = len(sec_u)
    xs = np.linspace(0.0, length, n_axial + 1)
    pts = np.concatenate([np.column_stack([np.full(ns, x), sec_u]) for x in xs])
    nq = len(sec_cells)
    cells = np.empty((nq * n_axial, 8), np.int64)
    for k in range(n_axial):
        lo = sec_cells + k * ns
        hi = sec_cells + (k + 1) * ns
        # corner bits: (i, j) in the section quad ordering, bit 2 = axial
        cells[k * nq:(k + 1) * nq] = np.concatenate([lo, hi], axis=1)

    def patch_of(fc, n):
        p = np.full(len(fc), 2)
        p[fc[:, 0] < 1e-9 * length] = 0
        p[fc[:, 0] > length * (1 - 1e-9)] = 1
        return p

    return hex_to_poly(pts, cells, patch_of, ["inlet", "outlet", "wall"], ["inlet", "outlet", "wall"])


HELIUM = dict(molecular_mass=6.65e-27, mu_ref=1.916e-5, T_ref=273.15, omega=0.66)


def write_case(path: str, mesh_file: str, name: str, boundaries: list, grid_n, grid_range, conditions: dict,
               kinetic: dict | None = None, gas: dict | None = None, stretch: float = 0.0) -> None:
    """Write a case TOML in the reference's schema (src/config/case_config.cpp:312-460); stretch > 0 adds
    velocity_grid.stretch, the tanh-stretched velocity grid (an extension with no reference counterpart)."""
    g = dict(HELIUM, **(gas or {}))
    lines = ["[case]", f'name = "{name}"', f'mesh = "{mesh_file}"', 'mode = "cis"', "", "[gas]"]
    lines += [f"{k} = {v!r}" for k, v in g.items()]
    lines += ["", "[conditions]"] + [f"{k} = {v!r}" for k, v in conditions.items()]
    for b in boundaries:
        lines += ["", "[[boundary]]"]
        for k, v in b.items():
            if isinstance(v, str):
                lines.append(f'{k} = "{v}"')
            elif isinstance(v, (list, tuple)):
                lines.append(f"{k} = [{', '.join(repr(float(x)) for x in v)}]")
            else:
                lines.append(f"{k} = {v!r}")
    lines += ["", "[velocity_grid]", f"n = [{', '.join(str(int(x)) for x in grid_n)}]",
              "range = [" + ", ".join(f"[{lo!r}, {hi!r}]" for lo, hi in grid_range) + "]"]
    if stretch > 0.0:
        lines.append(f"stretch = {float(stretch)!r}")
    if kinetic:
        lines += ["", "[kinetic]"]
        for k, v in kinetic.items():
            if isinstance(v, bool):
                lines.append(f"{k} = {'true' if v else 'false'}")
            elif isinstance(v, str):
                lines.append(f'{k} = "{v}"')
            else:
                lines.append(f"{k} = {v!r}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


T0 = 273.15


def slab_case(directory: str, nx: int = 200, n_vel: int = 24, vrange: float = 6.0, delta: float = 1.0,
              T_left: float = 0.95, T_right: float = 1.05, kinetic: dict | None = None, name: str = "slab") -> str:
    """1D slab heat transfer between diffuse walls (the 200-cell, 24^3 on [-6, 6] benchmark by default)."""
    import os
    mesh_file = f"{name}_nx{nx}.pmesh"
    write_pmesh(os.path.join(directory, mesh_file), slab_mesh(nx), {"generator": "gsis-sim_amd.meshgen.slab_mesh"})
    path = os.path.join(directory, f"{name}.toml")
    write_case(path, mesh_file, name,
               [dict(patch="wall_left", type="diffuse_wall", temperature=T_left * T0),
                dict(patch="wall_right", type="diffuse_wall", temperature=T_right * T0),
                dict(patch="sym_y", type="symmetry"), dict(patch="sym_z", type="symmetry")],
               [n_vel] * 3, [(-vrange, vrange)] * 3, dict(L_ref=1.0e-3, T0=T0, delta_rp=float(delta)), kinetic)
    return path


def tube_case(directory: str, n_core: int = 16, n_ring: int = 10, n_axial: int = 45, n_vel: int = 24,
              vrange: float = 6.0, delta: float = 10.0, accommodation: float = 1.0, kinetic: dict | None = None,
              name: str = "tube", stretch: float = 0.0) -> str:
    """Tube (R = L_ref, length 5R) from a pressure inlet (p_in, T0) to vacuum, diffuse (or Maxwell) wall."""
    import os
    mesh_file = f"{name}_{n_core}_{n_ring}_{n_axial}.pmesh"
    write_pmesh(os.path.join(directory, mesh_file), tube_mesh(n_core, n_ring, n_axial),
                {"generator": "gsis-sim_amd.meshgen.tube_mesh"})
    wall = dict(patch="wall", type="diffuse_wall", temperature=T0)
    if accommodation < 1.0:
        wall = dict(patch="wall", type="maxwell_wall", accommodation=float(accommodation), temperature=T0)
    path = os.path.join(directory, f"{name}.toml")
    write_case(path, mesh_file, name,
               [dict(patch="inlet", type="pressure_inlet", temperature=T0), dict(patch="outlet", type="vacuum"), wall],
               [n_vel] * 3, [(-vrange, vrange)] * 3, dict(L_ref=1.0e-3, T0=T0, delta_rp=float(delta)), kinetic,
               stretch=stretch)
    return path
