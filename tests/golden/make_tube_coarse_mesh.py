"""Writes tests/golden/tube_hex_coarse.npz: the reference's coarse tube mesh (14,464 hexahedra; the
mesh of its converged tube-to-vacuum CIS runs in cases/tube_vacuum/phase5/results/cis_study_coarse.csv
and cases/tube_vacuum/kinetic/results/param_study.csv), as generated by the reference's own
generator gsis_tools.geometry.tube.generate_tube_hex with the settings its case notes record
(cases/tube_vacuum/kinetic/mesh/tube_hex_coarse_quality.md: R = 1e-3 m, n_theta = 8,
target_cells = 15000, the remaining options at their defaults, RCM renumbering). The reference's
case file points at mesh/tube_hex_coarse.pmesh, which its tree does not hold.

The generator needs only numpy and runs in this container, reading the reference tree; what is kept
is data: points, face connectivity (all faces are quadrilaterals), owner / neighbour and patches.
tests/conftest.py writes it back as a .pmesh. Run from the repository root:
    python tests/golden/make_tube_coarse_mesh.py [reference tree, default /root/reference]
"""

from __future__ import annotations

import dataclasses
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main() -> None:
    reference = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    sys.path.insert(0, os.path.join(reference, "python", "src"))
    from gsis_tools.geometry.tube import TubeDims, TubeHexOptions, generate_tube_hex

    opt = dataclasses.replace(TubeHexOptions(), n_theta=8)
    mesh, info = generate_tube_hex(TubeDims(radius=1e-3), opt, target_cells=15000, renumber="rcm")
    sizes = np.diff(np.asarray(mesh.face_offsets))
    assert np.all(sizes == 4), "expected quadrilateral faces only"
    patches = [dict(name=p.name, kind=p.kind, start=int(p.start), count=int(p.count)) for p in mesh.patches]
    out = os.path.join(HERE, "tube_hex_coarse.npz")
    np.savez_compressed(out, points=np.asarray(mesh.points, np.float64),
                        face_nodes=np.asarray(mesh.face_nodes, np.int32).reshape(-1, 4),
                        owner=np.asarray(mesh.owner, np.int32), neighbour=np.asarray(mesh.neighbour, np.int32),
                        n_cells=np.int64(mesh.n_cells), patches=np.array(json.dumps(patches)),
                        options=np.array(json.dumps(info["options"])))
    print(f"{mesh.n_cells} cells, {len(mesh.owner)} faces, patches {patches}; options {info['options']}; wrote {out}")


if __name__ == "__main__":
    main()
