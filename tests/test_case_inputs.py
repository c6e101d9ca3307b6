"""Case input path: the automatic velocity grid and the reference's own case files.

* velocity_grid.auto = true (the reference's rule, src/config/case_config.cpp:589-610, and its test
  tests/test_config.cpp:77-100): with boundaries at 100 K and 273.15 K and a wall moving at 500 m/s,
  n = 2 ceil(L / (0.7 sqrt(R 100 K))) per axis on [-L, L] / sqrt(R T0), L = 500 + 5 sqrt(R 273.15);
  auto together with n is refused, and so is a grid of more than 1e8 points.
* Every case TOML under the reference's cases/ whose mesh file the reference tree holds builds a
  kinetic setup (skipped where the reference tree is absent).
"""

from __future__ import annotations

import glob
import math
import os
import re

import pytest

from conftest import T0

REFERENCE_CASES = "/root/reference/cases"


def _auto_case(mg, d, grid_lines, wall_velocity=(300.0, 0.0, 400.0)):
    mg.write_pmesh(os.path.join(d, "box.pmesh"), mg.box_mesh(3, 2, 2))
    path = os.path.join(d, "auto.toml")
    bnd = [dict(patch="xmin", type="pressure_inlet", temperature=T0), dict(patch="xmax", type="vacuum"),
           dict(patch="ymin", type="diffuse_wall", temperature=100.0, velocity=list(wall_velocity)),
           dict(patch="ymax", type="diffuse_wall", temperature=T0),
           dict(patch="zmin", type="symmetry"), dict(patch="zmax", type="symmetry")]
    mg.write_case(path, "box.pmesh", "auto", bnd, [8, 8, 8], [(-5.0, 5.0)] * 3, dict(L_ref=1.0e-3, T0=T0, delta_rp=1.0))
    text = open(path).read()
    text = re.sub(r"\[velocity_grid\]\nn = .*\nrange = .*\n", "[velocity_grid]\n" + grid_lines + "\n", text)
    open(path, "w").write(text)
    return path


def test_automatic_velocity_grid_follows_boundary_temperatures_and_speeds(gsk, meshgen, tmp_path):
    setup = gsk.Setup(_auto_case(meshgen, str(tmp_path), "auto = true"))
    R = 1.380649e-23 / 6.65e-27
    limit = 500.0 + 5.0 * math.sqrt(R * T0)
    expected = 2 * math.ceil(limit / (0.7 * math.sqrt(R * 100.0)))
    v = setup.view()
    assert list(v.grid_n) == [expected] * 3 and expected % 2 == 0
    assert setup.nv == expected ** 3
    for a in range(3):
        assert v.deq_hi[a] == pytest.approx(limit / math.sqrt(R * T0), rel=1e-14)
        assert v.deq_lo[a] == -v.deq_hi[a]
        assert (v.deq_hi[a] - v.deq_lo[a]) / expected <= 0.7 * math.sqrt(100.0 / T0)
    d = tmp_path / "both"
    d.mkdir()
    with pytest.raises(gsk.GskError, match="cannot be combined"):
        gsk.Setup(_auto_case(meshgen, str(d), "auto = true\nn = [8, 8, 8]"))
    d = tmp_path / "huge"
    d.mkdir()
    with pytest.raises(gsk.GskError, match="too many"):
        gsk.Setup(_auto_case(meshgen, str(d), "auto = true", wall_velocity=(1e6, 0.0, 0.0)))


def _reference_case_files():
    files = []
    for path in sorted(glob.glob(os.path.join(REFERENCE_CASES, "**", "*.toml"), recursive=True)):
        m = re.search(r'^\s*mesh\s*=\s*"([^"]+)"', open(path, encoding="utf-8").read(), re.M)
        if m and os.path.exists(os.path.join(os.path.dirname(path), m.group(1))):
            files.append(path)
    return files


@pytest.mark.skipif(not os.path.isdir(REFERENCE_CASES), reason="the reference tree is not present")
def test_reference_case_files_with_meshes_build_setups(gsk):
    files = _reference_case_files()
    assert len(files) >= 190
    failed = {}
    for path in files:
        try:
            s = gsk.Setup(path)
            assert s.n_cells > 0 and s.nv > 0
        except gsk.GskError as e:
            failed[os.path.relpath(path, REFERENCE_CASES)] = str(e)
    assert not failed, failed
