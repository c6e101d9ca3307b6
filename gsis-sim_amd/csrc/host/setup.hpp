#pragma once
// Host-side kinetic setup: velocity grid and storage layout, boundary patches, LU-SGS blocks
// and per-direction sweep orders, and the FP64 backend tables.
//
// Reference behaviour followed:
//   uniform midpoint velocity grid            src/velocity/velocity_grid.cpp:10-53
//   storage layout (mirror orbits, direction  src/kinetic/kinetic_setup.cpp:218-291
//   grouping), sweep directions               src/kinetic/kinetic_setup.cpp:182-216
//   build_kinetic_setup                       src/kinetic/kinetic_setup.cpp:293-514
//   device tables                             src/kinetic/kinetic_setup.cpp:516-599

#include <array>
#include <string>
#include <vector>

#include "../../../include/gsis_kinetic.h"
#include "config.hpp"
#include "mesh.hpp"

namespace gsk {

enum BoundaryKind : std::int32_t {
  kBcDiffuseWall = GSK_BC_DIFFUSE_WALL,
  kBcPressure = GSK_BC_PRESSURE,
  kBcVacuum = GSK_BC_VACUUM,
  kBcSymmetry = GSK_BC_SYMMETRY,
  kBcPeriodic = GSK_BC_PERIODIC,
  kBcMaxwellWall = GSK_BC_MAXWELL_WALL,
};

class VelocityGrid {
 public:
  // stretch = 0: the reference's uniform midpoint grid. stretch = b > 0 (an extension): per axis,
  // points uniform in s = tanh(b (xi - c) / h) / tanh(b) over the range c +- h, i.e.
  // xi_i = c + (h / b) atanh(s_i tanh b) at the midpoints s_i of n equal cells of (-1, 1), each with
  // the weight (dxi/ds) ds = (2 / n) (h / b) tanh b / (1 - (s_i tanh b)^2): the midpoint rule in s,
  // clustering the points near the centre and widening them towards the range ends.
  VelocityGrid(std::array<int, 3> n, std::array<double, 3> lo, std::array<double, 3> hi, double stretch = 0.0);
  int size() const { return size_; }
  int n(int a) const { return n_[static_cast<std::size_t>(a)]; }
  double lo(int a) const { return lo_[static_cast<std::size_t>(a)]; }
  double hi(int a) const { return hi_[static_cast<std::size_t>(a)]; }
  double delta(int a) const { return delta_[static_cast<std::size_t>(a)]; }
  double weight() const { return weight_; }
  double stretch() const { return stretch_; }
  bool stretched() const { return stretch_ > 0.0; }
  // node i of axis a: lo + (i + 1/2) d on the uniform grid, written about the range centre so
  // mirrored points are exactly -xi (the stretched nodes are odd about the centre as well)
  double coordinate(int a, int i) const { return node_[static_cast<std::size_t>(a)][static_cast<std::size_t>(i)]; }
  // quadrature weight of point k relative to weight() (exactly 1 on the uniform grid)
  double relative_weight(int k) const {
    if (!stretched()) return 1.0;
    const std::array<int, 3> ijk = multi_index(k);
    double w = 1.0;
    for (std::size_t a = 0; a < 3; ++a) w *= axis_w_[a][static_cast<std::size_t>(ijk[a])];
    return w / weight_;
  }
  // largest node spacing of axis a (the spacing the discrete-equilibrium resolution test uses)
  double max_spacing(int a) const;
  int index(int i, int j, int l) const { return i + n_[0] * (j + n_[1] * l); }
  std::array<int, 3> multi_index(int k) const { return {k % n_[0], (k / n_[0]) % n_[1], k / (n_[0] * n_[1])}; }
  Vec3 xi(int k) const { return {xs_[static_cast<std::size_t>(k)], ys_[static_cast<std::size_t>(k)], zs_[static_cast<std::size_t>(k)]}; }
  bool symmetric(int a) const { return symmetric_[static_cast<std::size_t>(a)]; }
  int mirror(int a, int k) const {
    std::array<int, 3> ijk = multi_index(k);
    ijk[static_cast<std::size_t>(a)] = n_[static_cast<std::size_t>(a)] - 1 - ijk[static_cast<std::size_t>(a)];
    return index(ijk[0], ijk[1], ijk[2]);
  }

 private:
  std::array<int, 3> n_{};
  std::array<double, 3> lo_{}, hi_{}, delta_{}, center_{}, half_delta_{};
  double weight_ = 0.0, stretch_ = 0.0;
  std::array<std::vector<double>, 3> node_, axis_w_;
  int size_ = 0;
  std::vector<double> xs_, ys_, zs_;
  std::array<bool, 3> symmetric_{};
};

struct VelocityLayout {
  int nv = 0;
  int orbit_size = 1;
  std::array<int, 3> axis_xor{0, 0, 0};
  std::vector<int> storage_to_grid, grid_to_storage;
  std::vector<double> xi_x, xi_y, xi_z;
  double weight = 0.0;
  std::vector<double> weight_rel;  // storage order, w_p / weight (empty on the uniform grid)
  std::vector<double> nodes;       // stretched grid: node coordinates of axis 0, 1, 2 (empty: uniform)
  double xi_max = 0.0;
};


struct PatchCondition {
  std::string name;
  BoundaryType type = BoundaryType::DiffuseWall;
  BoundaryKind kind = kBcDiffuseWall;
  Index first_bface = 0, n_faces = 0;
  double accommodation = 1.0;
  double target_mass_flow = 0.0;
  double T = 1.0, p = 0.0;
  Vec3 u_wall{};
  int mirror_axis = -1;
  bool zero_tangential = false;
};

struct KineticSetup {
  KineticParams params;
  // case identity and the [gsis] convergence settings a cis-run driver reads (result files)
  std::string case_name, mesh_file;
  double gsis_tolerance = 1e-5, mass_balance_tolerance = 2e-3;
  bool mass_balance_required = false;
  std::vector<double> section_positions;  // [output].mass_flow_sections
  int section_axis = 0;
  int limiter_mode = 1;
  DerivedQuantities derived;
  double length_scale = 1.0;
  std::vector<Index> cell_new_to_old, cell_old_to_new, bface_new_to_old, bface_old_to_new;
  PolyMesh mesh;
  MeshGeometry geometry;
  CellFaceCsr csr;
  std::vector<Vec3> lsq;
  std::vector<double> volume, vdt;
  double diag_scale = 1.0;
  VelocityLayout velocity;
  std::vector<Index> block_offsets;
  bool discrete_eq = false;
  double deq_lo[3] = {0, 0, 0}, deq_hi[3] = {0, 0, 0}, deq_d[3] = {1, 1, 1};
  double deq_criterion = 27.0;
  int deq_max_iterations = 20;
  double deq_tolerance = 1e-6;
  int n_sweep_dirs = 0;
  std::vector<Vec3> sweep_dirs;
  std::vector<std::int32_t> sweep_order, sweep_pos;
  std::vector<std::uint8_t> sweep_dir_of;
  std::vector<PatchCondition> patches;
  std::vector<int> bface_patch;
  std::vector<std::int64_t> maxwell_offset;
  std::vector<double> maxwell_specular;
  int n_maxwell_faces = 0;
  std::array<int, 3> grid_n{};
  bool closed_domain = false;
  double mu0 = 0, omega = 0, Pr = 2.0 / 3.0, gamma = 5.0 / 3.0;

  // backend tables (internal numbering)
  std::vector<std::int32_t> csr_offsets;
  std::vector<gsk_csr_entry> csr_entries;
  std::vector<gsk_bface> bfaces;
  std::vector<double> eps2;

  int n_cells() const { return mesh.n_cells; }
  int n_bfaces() const { return mesh.n_boundary_faces(); }
  int nv() const { return velocity.nv; }
  int n_blocks() const { return static_cast<int>(block_offsets.size()) - 1; }
  gsk_setup_view view() const;
};

KineticSetup build_kinetic_setup(const CaseConfig& config, const PolyMesh& mesh, const KineticParams& params);
// load_case + read_pmesh + build_kinetic_setup with "key=value" [kinetic] overrides
KineticSetup setup_from_case(const std::string& toml_path, const std::string& overrides);

}  // namespace gsk
