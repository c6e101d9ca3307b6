"""CPU tests of the boundary and the host setup: the C ABI loads and exports every symbol the header
declares; case/mesh input errors fail loudly; setup invariants the reference's own tests check
(LSQ exactness for linear fields, closed cells, sweep orders, mirror layout)."""

from __future__ import annotations

import os
import re

import numpy as np
import pytest

from conftest import ROOT, T0, box_case


def header_functions():
    text = open(os.path.join(ROOT, "include", "gsis_kinetic.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(gsk_\w+)\s*\(", text, re.M)))


def test_library_exports_every_header_symbol(gsk):
    L = gsk.lib()
    declared = header_functions()
    assert len(declared) >= 40
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(declared) == sorted(gsk.EXPORTS)
    assert b"gfx950" in L.gsk_version()


def test_setup_sizes_and_view(gsk, meshgen, tmp_path):
    s = gsk.Setup(box_case(meshgen, str(tmp_path)))
    assert (s.n_cells, s.nv, s.n_patches) == (30, 512, 6)
    v = s.view()
    assert v.n_cells == 30 and v.nv == 512 and v.orbit_size == 4  # symmetry z + Maxwell wall y mirrors
    assert v.n_maxwell_faces == 5 * 2
    assert v.n_sweep_dirs == 4 and v.block_passes == 2 and v.negative_check == 3 and v.discrete_eq == 1
    assert v.deq_tolerance == 1e-6  # fp32 default precision


def test_bad_inputs_fail_loudly(gsk, meshgen, tmp_path):
    path = box_case(meshgen, str(tmp_path))
    with pytest.raises(gsk.GskError, match="unknown key kinetic.no_such_key"):
        gsk.Setup(path, "no_such_key = 1")
    with pytest.raises(gsk.GskError, match="sweep_order"):
        gsk.Setup(path, {"sweep_order": "zigzag"})
    with pytest.raises(gsk.GskError, match="cannot open"):
        gsk.Setup(os.path.join(str(tmp_path), "missing.toml"))
    text = open(path).read().replace('patch = "zmax"', 'patch = "nope"')
    bad = os.path.join(str(tmp_path), "bad.toml")
    open(bad, "w").write(text)
    with pytest.raises(gsk.GskError, match="case and mesh do not match"):
        gsk.Setup(bad)


def test_backend_needs_a_gpu(gsk, meshgen, tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    s = gsk.Setup(box_case(meshgen, str(tmp_path)))
    with pytest.raises(gsk.GskError, match="no HIP device"):
        gsk.Backend(s)
    with pytest.raises(gsk.GskError, match="no HIP device"):
        gsk.Solver(s)


def _csr(setup):
    v = setup.view()
    off = np.ctypeslib.as_array(v.csr_offsets, (setup.n_cells + 1,)).copy()
    ent = np.ctypeslib.as_array(v.csr, (int(off[-1]),))
    return v, off, ent


def test_lsq_exact_for_linear_fields(gsk, meshgen, tmp_path):
    """sum_k c_k r_k^T = I per cell (LSQ reproduces a linear field exactly; reference
    tests/test_mesh_fv.cpp:122-194)."""
    s = gsk.Setup(box_case(meshgen, str(tmp_path), n=(4, 4, 3), distort=0.3))
    v, off, ent = _csr(s)
    for c in range(s.n_cells):
        M = np.zeros((3, 3))
        for k in range(off[c], off[c + 1]):
            e = ent[k]
            ds, dj, cc = np.array(e["ds"]), np.array(e["dother"]), np.array(e["c"])
            r = ds - dj if e["other"] >= 0 else ds
            M += np.outer(cc, r)
        assert np.allclose(M, np.eye(3), atol=1e-9), (c, M)


def test_cells_are_closed_and_volume_adds_up(gsk, meshgen, tmp_path):
    s = gsk.Setup(box_case(meshgen, str(tmp_path), n=(4, 4, 3), distort=0.3))
    v, off, ent = _csr(s)
    for c in range(s.n_cells):
        S = sum(np.array(ent[k]["s"]) for k in range(off[c], off[c + 1]))
        area = sum(np.linalg.norm(ent[k]["s"]) for k in range(off[c], off[c + 1]))
        assert np.linalg.norm(S) <= 1e-10 * area
    vol = np.ctypeslib.as_array(v.volume, (s.n_cells,))
    assert abs(vol.sum() - 1.0 * 0.6 * 0.4) < 1e-12  # nondimensional by L_ref = 1 mm


def test_sweep_orders_are_sorted_permutations(gsk, meshgen, tmp_path):
    s = gsk.Setup(box_case(meshgen, str(tmp_path), n=(4, 4, 3), distort=0.3), {"sweep_blocks": 3})
    v = s.view()
    nc, nd = s.n_cells, v.n_sweep_dirs
    order = np.ctypeslib.as_array(v.sweep_order, (nd * nc,)).reshape(nd, nc)
    pos = np.ctypeslib.as_array(v.sweep_pos, (nd * nc,)).reshape(nd, nc)
    blocks = np.ctypeslib.as_array(v.block_offsets, (v.n_blocks + 1,))
    assert v.n_blocks == 3 and blocks[0] == 0 and blocks[-1] == nc
    ctr = s.cell_centers()
    dirs = np.array([[1, 1, 1], [1, 1, -1], [1, -1, 1], [1, -1, -1]]) / np.sqrt(3.0)
    for g in range(nd):
        assert np.array_equal(pos[g][order[g]], np.arange(nc))
        key = ctr @ dirs[g]
        for b in range(v.n_blocks):
            seg = order[g][blocks[b]:blocks[b + 1]]
            assert sorted(seg) == list(range(blocks[b], blocks[b + 1]))
            assert np.all(np.diff(key[seg]) >= 0)


def test_velocity_mirror_is_storage_xor(gsk, meshgen, tmp_path):
    s = gsk.Setup(box_case(meshgen, str(tmp_path)))
    v = s.view()
    nv = v.nv
    x = np.ctypeslib.as_array(v.xi_x, (nv,))
    y = np.ctypeslib.as_array(v.xi_y, (nv,))
    z = np.ctypeslib.as_array(v.xi_z, (nv,))
    bf = np.ctypeslib.as_array(v.bfaces, (s.n_bfaces,))
    for b in bf:
        if b["kind"] == 3:  # symmetry (z faces)
            p = np.arange(nv)
            q = p ^ b["mirror_xor"]
            assert np.array_equal(z[q], -z[p]) and np.array_equal(x[q], x[p]) and np.array_equal(y[q], y[p])
    assert abs(v.weight - (10.0 / 8) ** 3) < 1e-15


def test_generated_slab_matches_reference_mesh(gsk, meshgen, tmp_path):
    """The slab generator reproduces the reference's slab_nx200.pmesh geometry (cell volumes and
    centres after the solver's renumbering)."""
    ref_mesh = "/root/reference/cases/verification/kinetic/mesh/slab_nx200.pmesh"
    if not os.path.exists(ref_mesh):
        pytest.skip("reference tree not present")
    a = gsk.Setup(meshgen.slab_case(str(tmp_path), nx=200, n_vel=4, vrange=6.0))
    text = open(os.path.join(str(tmp_path), "slab.toml")).read().replace("slab_nx200.pmesh", ref_mesh)
    open(os.path.join(str(tmp_path), "ref.toml"), "w").write(text)
    b = gsk.Setup(os.path.join(str(tmp_path), "ref.toml"))
    assert np.allclose(a.cell_centers(), b.cell_centers(), rtol=0, atol=1e-15)
    va = np.ctypeslib.as_array(a.view().volume, (200,))
    vb = np.ctypeslib.as_array(b.view().volume, (200,))
    assert np.allclose(va, vb, rtol=1e-12)


def test_solver_parameter_setters_validate(gsk, meshgen, oracle_mod, tmp_path):
    """The solver-level setters used by the sensitivity solve reject what they cannot honour, and the
    device-pointer entry points fail loudly on a backend without device buffers (the oracle)."""
    setup = gsk.Setup(box_case(meshgen, str(tmp_path), kinetic={"precision": "fp64"}))
    cb = oracle_mod.CpuBackend(setup, "oracle", ops_type=gsk.BackendOps)
    sv = gsk.Solver(setup, cb.ops)
    with pytest.raises(gsk.GskError, match="Maxwell walls only"):
        sv.set_accommodation(0, 0.9)  # patch 0 is the pressure inlet
    with pytest.raises(gsk.GskError, match="positive"):
        sv.set_accommodation(3, 0.0)  # patch 3 is the Maxwell wall
    with pytest.raises(gsk.GskError, match="boundary iterations"):
        sv.set_boundary_iterations(0)
    with pytest.raises(gsk.GskError, match="Newton tolerances"):
        sv.set_newton_tolerances(-1.0, 0.0)
    sv.set_boundary_iterations(3)
    assert sv.mu0() > 0.0
    # what needs the backend's device-resident operators fails loudly on a backend without them
    for call in (lambda: sv.set_accommodation(3, 0.7), lambda: sv.set_mu0(2.0 * sv.mu0())):
        with pytest.raises(gsk.GskError, match="no device-resident"):
            call()
    with pytest.raises(gsk.GskError, match="null argument"):
        sv.residual_device(0, setup.n_cells * setup.nv)
    sv.close()


def test_velocity_grid_stretch_is_validated(gsk, meshgen, tmp_path):
    path = box_case(meshgen, str(tmp_path), stretch=1.5)
    text = open(path).read()
    assert "stretch = 1.5" in text
    bad = os.path.join(str(tmp_path), "bad_stretch.toml")
    open(bad, "w").write(text.replace("stretch = 1.5", "stretch = -1.0"))
    with pytest.raises(gsk.GskError, match="stretch must be in"):
        gsk.Setup(bad)


@pytest.mark.parametrize("order,n_dirs", [("octant", 4), ("dir13", 13), ("block", 0)])
def test_sweep_directions_group_the_velocities(gsk, meshgen, tmp_path, order, n_dirs):
    """make_sweep_directions / sweep_direction_of / make_velocity_layout (the reference's
    tests/kinetic/test_kinetic_phase5.cpp:241-264): octant gives 4 directions, dir13 13, block none;
    xi and -xi share a direction while flipping one component changes it (octant); without mirror
    axes (walls only) the storage order runs through the directions in order."""
    from conftest import kinetic_case

    mesh = meshgen.box_mesh(2, 2, 2, size=(1.0, 1.0, 1.0), names=["wall"] * 6)
    path = kinetic_case(meshgen, str(tmp_path), mesh, [dict(patch="wall", type="diffuse_wall")], n_vel=6, vmax=3.0,
                        kinetic=dict(precision="fp64", sweep_order=order), name=f"dirs_{order}")
    s = gsk.Setup(path)
    v = s.view()
    assert v.n_sweep_dirs == n_dirs
    if n_dirs == 0:
        return
    nv = s.nv
    xi = np.column_stack([np.ctypeslib.as_array(p, (nv,)) for p in (v.xi_x, v.xi_y, v.xi_z)])
    d = np.ctypeslib.as_array(v.sweep_dir_of, (nv,)).astype(int)
    assert np.all(np.diff(d) >= 0) and len(np.unique(d)) == n_dirs  # grouped, every direction used
    index = {tuple(x): i for i, x in enumerate(xi)}
    for p in range(nv):
        assert d[index[tuple(-xi[p])]] == d[p]  # xi and -xi: one direction
        if order == "octant":
            flipped = xi[p] * np.array([1.0, -1.0, 1.0])
            assert d[index[tuple(flipped)]] != d[p]


def test_kinetic_parameters_defaults_values_and_errors(gsk, meshgen, tmp_path):
    """[kinetic] defaults, values and rejections as the reference's tests/kinetic/test_kinetic_params.cpp
    :44-86 state them, read back through the case description."""
    from conftest import kinetic_case

    mesh = meshgen.box_mesh(2, 1, 1, size=(1.0, 0.5, 0.5), names=["wall"] * 6)

    def params(kin, name):
        path = kinetic_case(meshgen, str(tmp_path), mesh, [dict(patch="wall", type="diffuse_wall")], kinetic=kin,
                            name=name)
        return gsk.Setup(path).describe()

    d = params(None, "defaults")
    assert d["params"]["pseudo_cfl"] == 0.0 and d["params"]["precision"] == "fp32"
    assert d["params"]["boundary_iterations"] == 1
    p = params(dict(pseudo_cfl=50.0, sweep_blocks=8, precision="fp64", backend="cpu"), "values")
    assert p["params"]["pseudo_cfl"] == 50.0 and p["params"]["precision"] == "fp64"
    assert p["params"]["backend"] == "cpu" and p["sweep_blocks"] == 2  # 8 requested, at most one block per cell
    assert params(dict(boundary_iterations=8), "bi")["params"]["boundary_iterations"] == 8
    q = params(dict(limiter="none", limiter_eps="relative", initial_pressure=0.5, initial_temperature=1.2,
                    convergence_u_floor=0.01, launch_cells=100, sweep_cells_per_launch=10), "more")["params"]
    assert q["limiter"] == "none" and q["limiter_eps"] == "relative" and q["initial_pressure"] == 0.5
    assert q["initial_temperature"] == 1.2 and q["convergence_u_floor"] == 0.01
    for kin, word in ((dict(boundary_iterations=0), "boundary_iterations"),
                      (dict(boundary_iterations=1001), "boundary_iterations"), (dict(sweep_block=8), "sweep_block"),
                      (dict(precision="fp16"), "kinetic.precision"), (dict(limiter="minmod"), "kinetic.limiter"),
                      (dict(initial_pressure=0.0), "initial_pressure")):
        with pytest.raises(gsk.GskError, match=word):
            params(kin, "bad")
