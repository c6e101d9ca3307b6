#pragma once
// Face-based polyhedral mesh (.pmesh v1), its finite-volume geometry, the cell->face CSR,
// least-squares gradient coefficients and the cell orderings used by the LU-SGS sweeps.
//
// Reference behaviour followed (file:line relative to the reference tree):
//   .pmesh layout and conventions       docs/pmesh_format.md §1-§4; src/mesh/pmesh_io.hpp:13-31
//   face / cell geometry                src/mesh/geometry.cpp:10-125
//   cell->face CSR (ascending face)     src/mesh/fv_geometry.cpp:53-91
//   LSQ coefficients c_k = w G^-1 r_k   src/mesh/fv_geometry.cpp:95-176
//   adjacency / renumbering             src/mesh/ordering.cpp:9-29, :122-236
//   RCM and RCB block orderings         src/mesh/ordering.cpp:30-120, src/kinetic/kinetic_setup.cpp:46-178
//                                       (ordering.cpp here)
//   contiguous partition                src/mesh/ordering.cpp:238-250

#include <array>
#include <string>
#include <vector>

#include "core.hpp"

namespace gsk {

struct Patch {
  std::string name;
  std::string kind = "patch";
  Index start = 0;
  Index count = 0;
  Index end() const { return start + count; }
};

struct PolyMesh {
  std::vector<Vec3> points;
  std::vector<Offset> face_offsets{0};
  std::vector<Index> face_nodes;
  std::vector<Index> owner;
  std::vector<Index> neighbour;
  std::vector<Patch> patches;
  Index n_cells = 0;

  Index n_points() const { return static_cast<Index>(points.size()); }
  Index n_faces() const { return static_cast<Index>(owner.size()); }
  Index n_internal_faces() const { return static_cast<Index>(neighbour.size()); }
  Index n_boundary_faces() const { return n_faces() - n_internal_faces(); }
  Index face_size(Index f) const {
    return static_cast<Index>(face_offsets[static_cast<std::size_t>(f) + 1] - face_offsets[static_cast<std::size_t>(f)]);
  }
  const Index* face(Index f) const { return face_nodes.data() + face_offsets[static_cast<std::size_t>(f)]; }
  int find_patch(const std::string& name) const;
  std::vector<int> boundary_face_patches() const;
};

// .pmesh v1 reader / writer (little endian, JSON header, 64-byte aligned arrays)
PolyMesh read_pmesh(const std::string& path);
void write_pmesh(const std::string& path, const PolyMesh& mesh);
void validate_topology(const PolyMesh& mesh);

struct MeshGeometry {
  std::vector<Vec3> face_area_vector;
  std::vector<double> face_area;
  std::vector<Vec3> face_normal;
  std::vector<Vec3> face_center;
  std::vector<double> cell_volume;
  std::vector<Vec3> cell_center;
};
MeshGeometry compute_geometry(const PolyMesh& mesh);

struct CellFaceCsr {
  std::vector<Offset> offsets;
  std::vector<Index> faces;
  std::vector<Index> other;
  std::vector<signed char> sign;
  Offset begin(Index c) const { return offsets[static_cast<std::size_t>(c)]; }
  Offset end(Index c) const { return offsets[static_cast<std::size_t>(c) + 1]; }
};
CellFaceCsr build_cell_faces(const PolyMesh& mesh);

// c_k per CSR entry; weight |r|^-p (p = 0: unweighted)
std::vector<Vec3> build_lsq_coefficients(const PolyMesh& mesh, const MeshGeometry& g, const CellFaceCsr& csr,
                                         double weight_power);

struct CellAdjacency {
  std::vector<Offset> offsets;
  std::vector<Index> cells;
  Index degree(Index c) const {
    return static_cast<Index>(offsets[static_cast<std::size_t>(c) + 1] - offsets[static_cast<std::size_t>(c)]);
  }
};
CellAdjacency build_cell_adjacency(const PolyMesh& mesh);

std::vector<Index> rcm_ordering(const PolyMesh& mesh);
// RCB into P blocks, reverse Cuthill-McKee inside each block; returns new_to_old and block offsets
std::vector<Index> rcb_ordering(const PolyMesh& mesh, const MeshGeometry& g, int P, std::vector<Index>& block_offsets);
std::vector<Index> invert_permutation(const std::vector<Index>& perm);
PolyMesh renumber_cells(const PolyMesh& mesh, const std::vector<Index>& new_to_old);
std::vector<Index> match_boundary_faces(const PolyMesh& original, const PolyMesh& renumbered);
std::vector<Index> partition_contiguous(Index n_cells, int n_blocks);

}  // namespace gsk
