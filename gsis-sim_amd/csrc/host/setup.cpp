// Kinetic setup (see setup.hpp for the reference locations followed).

#include "setup.hpp"

#include <algorithm>
#include <map>
#include <numeric>

namespace gsk {

VelocityGrid::VelocityGrid(std::array<int, 3> n, std::array<double, 3> lo, std::array<double, 3> hi, double stretch)
    : n_(n), lo_(lo), hi_(hi), stretch_(stretch) {
  GSK_CHECK(std::isfinite(stretch_) && stretch_ >= 0.0, "invalid velocity grid stretch");
  long long total = 1;
  for (std::size_t a = 0; a < 3; ++a) {
    GSK_CHECK(n_[a] >= 1, "velocity grid needs at least one point per axis");
    GSK_CHECK(std::isfinite(lo_[a]) && std::isfinite(hi_[a]) && lo_[a] < hi_[a], "invalid velocity grid range");
    delta_[a] = (hi_[a] - lo_[a]) / n_[a];
    half_delta_[a] = (hi_[a] - lo_[a]) / (2.0 * n_[a]);
    center_[a] = 0.5 * (lo_[a] + hi_[a]);
    total *= n_[a];
  }
  GSK_CHECK(total <= 100000000, "too many velocity points");
  size_ = static_cast<int>(total);
  weight_ = delta_[0] * delta_[1] * delta_[2];
  for (std::size_t a = 0; a < 3; ++a) {
    const int na = n_[a];
    node_[a].resize(static_cast<std::size_t>(na));
    axis_w_[a].resize(static_cast<std::size_t>(na));
    const double h = 0.5 * (hi_[a] - lo_[a]), tb = std::tanh(stretch_);
    for (int i = 0; i < na; ++i) {
      const int m = 2 * i + 1 - na;  // s_i = m / na; m and the mirror's -m are exact negatives
      if (!stretched()) {
        node_[a][static_cast<std::size_t>(i)] = center_[a] + static_cast<double>(m) * half_delta_[a];
        axis_w_[a][static_cast<std::size_t>(i)] = delta_[a];
        continue;
      }
      const double st = (static_cast<double>(std::abs(m)) / na) * tb;  // |s_i| tanh b
      const double off = (h / stretch_) * std::atanh(st);
      node_[a][static_cast<std::size_t>(i)] = center_[a] + (m < 0 ? -off : off);
      axis_w_[a][static_cast<std::size_t>(i)] = (2.0 / na) * (h / stretch_) * tb / (1.0 - st * st);
    }
  }
  xs_.resize(static_cast<std::size_t>(size_));
  ys_.resize(static_cast<std::size_t>(size_));
  zs_.resize(static_cast<std::size_t>(size_));
  for (int l = 0; l < n_[2]; ++l)
    for (int j = 0; j < n_[1]; ++j)
      for (int i = 0; i < n_[0]; ++i) {
        const auto k = static_cast<std::size_t>(index(i, j, l));
        xs_[k] = coordinate(0, i);
        ys_[k] = coordinate(1, j);
        zs_[k] = coordinate(2, l);
      }
  for (std::size_t a = 0; a < 3; ++a) symmetric_[a] = std::abs(lo_[a] + hi_[a]) <= 1e-12 * (hi_[a] - lo_[a]);
}

double VelocityGrid::max_spacing(int a) const {
  const auto s = static_cast<std::size_t>(a);
  double d = delta_[s];
  if (stretched())
    for (std::size_t i = 0; i < axis_w_[s].size(); ++i) d = std::max(d, axis_w_[s][i]);
  return d;
}


// ---------------------------------------------------------------------------------------------
// The setup is assembled in stages by SetupBuilder; each stage states the rule it applies. The
// rules are the reference's (src/kinetic/kinetic_setup.cpp:182-599), because the cell numbering,
// velocity storage order, sweep orders and face tables decide every iterate bit for bit.
// ---------------------------------------------------------------------------------------------
namespace {

// Sweep directions by [kinetic].sweep_order, as unit vectors in this order:
//   "octant": the four body diagonals (1,1,1) (1,1,-1) (1,-1,1) (1,-1,-1);
//   "dir13":  the axes x, y, z, the face diagonals (1,1,0) (1,-1,0) (1,0,1) (1,0,-1) (0,1,1) (0,1,-1),
//             then the four body diagonals;
//   "block":  none (every velocity sweeps the cells in internal numbering).
std::vector<Vec3> sweep_direction_table(const std::string& order) {
  static const int kAxes[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  static const int kFaceDiagonals[6][3] = {{1, 1, 0}, {1, -1, 0}, {1, 0, 1}, {1, 0, -1}, {0, 1, 1}, {0, 1, -1}};
  static const int kBodyDiagonals[4][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1}};
  GSK_CHECK(order == "octant" || order == "dir13" || order == "block", "unknown sweep_order \"" + order + "\"");
  std::vector<Vec3> dirs;
  auto add = [&dirs](const int(&v)[3]) {
    const Vec3 x{static_cast<double>(v[0]), static_cast<double>(v[1]), static_cast<double>(v[2])};
    dirs.push_back(x / norm(x));
  };
  if (order == "dir13") {
    for (const auto& v : kAxes) add(v);
    for (const auto& v : kFaceDiagonals) add(v);
  }
  if (order != "block")
    for (const auto& v : kBodyDiagonals) add(v);
  return dirs;
}

// The sweep direction a velocity follows: the one most nearly parallel or antiparallel to it,
// largest |d . xi|; a later direction must beat the current best by more than 1e-12 to replace it.
int nearest_direction(const std::vector<Vec3>& dirs, const Vec3& xi) {
  int pick = 0;
  double best = -1.0;
  for (int g = 0; g < static_cast<int>(dirs.size()); ++g) {
    const double align = std::abs(dot(dirs[static_cast<std::size_t>(g)], xi));
    if (align > best + 1e-12) {
      pick = g;
      best = align;
    }
  }
  return pick;
}

// The axis (0, 1, 2) every face of the patch is normal to, within tol of a unit axis vector; -1 if
// the faces are not all normal to one axis.
int patch_normal_axis(const MeshGeometry& g, const Patch& patch, double tol = 1e-8) {
  int common = -1;
  for (Index f = patch.start; f < patch.end(); ++f) {
    const Vec3& n = g.face_normal[static_cast<std::size_t>(f)];
    int axis = -1;
    for (int a = 0; a < 3; ++a)
      if (std::abs(std::abs(n[a]) - 1.0) <= tol) axis = a;
    if (axis < 0 || (common >= 0 && axis != common)) return -1;
    common = axis;
  }
  return common;
}

// Axis a face normal lies along within 1e-10 (Maxwell walls reflect through a storage-order xor
// there), or -1.
int aligned_axis(const Vec3& n) {
  int axis = -1;
  for (int a = 0; a < 3; ++a)
    if (std::abs(std::abs(n[a]) - 1.0) < 1e-10) axis = a;
  return axis;
}

// Storage order of the velocities.
//  * With mirror axes A (symmetry planes, axis-aligned Maxwell walls): velocities are stored in
//    mirror orbits. The orbit representatives are the grid points in the lower half of every axis
//    of A, taken in grid order; the r-th orbit occupies positions r 2^|A| + m, m = 0 .. 2^|A| - 1,
//    where position m is the representative reflected on the axes A[i] whose bit i is set in m. So
//    the reflection on A[i] is the storage xor 2^i.
//  * Without mirrors, with sweep directions: grid points grouped by their sweep direction (groups
//    in direction order, grid order inside a group), so that a velocity chunk mostly follows one
//    cell order.
//  * Otherwise grid order.
VelocityLayout layout_velocities(const VelocityGrid& grid, std::vector<int> mirror_axes, const std::vector<Vec3>& dirs) {
  VelocityLayout L;
  L.nv = grid.size();
  L.weight = grid.weight();
  std::sort(mirror_axes.begin(), mirror_axes.end());
  mirror_axes.erase(std::unique(mirror_axes.begin(), mirror_axes.end()), mirror_axes.end());
  for (const int a : mirror_axes) {
    GSK_CHECK(a >= 0 && a < 3, "invalid mirror axis");
    GSK_CHECK(grid.symmetric(a), strf("velocity range of axis %c is not symmetric about 0; mirrors unavailable", "xyz"[a]));
    GSK_CHECK(grid.n(a) % 2 == 0, strf("axis %c of a mirrored velocity grid needs an even point count", "xyz"[a]));
  }
  const int n_mirror = static_cast<int>(mirror_axes.size());
  L.orbit_size = 1 << n_mirror;
  for (int i = 0; i < n_mirror; ++i) L.axis_xor[static_cast<std::size_t>(mirror_axes[static_cast<std::size_t>(i)])] = 1 << i;

  L.storage_to_grid.assign(static_cast<std::size_t>(L.nv), -1);
  if (n_mirror > 0) {
    std::size_t next = 0;
    for (int k = 0; k < L.nv; ++k) {
      const std::array<int, 3> rep = grid.multi_index(k);
      bool lower = true;
      for (const int a : mirror_axes) lower = lower && rep[static_cast<std::size_t>(a)] < grid.n(a) / 2;
      if (!lower) continue;
      for (int m = 0; m < L.orbit_size; ++m) {
        std::array<int, 3> q = rep;
        for (int i = 0; i < n_mirror; ++i)
          if ((m >> i) & 1) {
            const auto a = static_cast<std::size_t>(mirror_axes[static_cast<std::size_t>(i)]);
            q[a] = grid.n(static_cast<int>(a)) - 1 - q[a];
          }
        L.storage_to_grid[next++] = grid.index(q[0], q[1], q[2]);
      }
    }
    GSK_CHECK(next == static_cast<std::size_t>(L.nv), "internal error: velocity orbits do not cover the grid");
  } else if (!dirs.empty()) {
    std::vector<std::vector<int>> bucket(dirs.size());
    for (int k = 0; k < L.nv; ++k) bucket[static_cast<std::size_t>(nearest_direction(dirs, grid.xi(k)))].push_back(k);
    std::size_t next = 0;
    for (const auto& b : bucket)
      for (const int k : b) L.storage_to_grid[next++] = k;
  } else {
    std::iota(L.storage_to_grid.begin(), L.storage_to_grid.end(), 0);
  }

  const auto nv = static_cast<std::size_t>(L.nv);
  L.grid_to_storage.assign(nv, -1);
  L.xi_x.resize(nv);
  L.xi_y.resize(nv);
  L.xi_z.resize(nv);
  for (std::size_t p = 0; p < nv; ++p) {
    const int k = L.storage_to_grid[p];
    L.grid_to_storage[static_cast<std::size_t>(k)] = static_cast<int>(p);
    const Vec3 xi = grid.xi(k);
    L.xi_x[p] = xi.x;
    L.xi_y[p] = xi.y;
    L.xi_z[p] = xi.z;
    L.xi_max = std::max(L.xi_max, norm(xi));
  }
  if (grid.stretched()) {
    L.weight_rel.resize(nv);
    for (std::size_t p = 0; p < nv; ++p) L.weight_rel[p] = grid.relative_weight(L.storage_to_grid[p]);
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < grid.n(a); ++i) L.nodes.push_back(grid.coordinate(a, i));
  }
  for (int i = 0; i < n_mirror; ++i) {  // the orbit layout's promise, checked
    const int a = mirror_axes[static_cast<std::size_t>(i)];
    for (std::size_t p = 0; p < nv; ++p)
      GSK_CHECK(L.storage_to_grid[p ^ static_cast<std::size_t>(1 << i)] == grid.mirror(a, L.storage_to_grid[p]),
                "internal error: a mirror is not a storage xor");
  }
  return L;
}

std::vector<std::string> mismatches(const CaseConfig& cfg, const PolyMesh& mesh, const MeshGeometry& g,
                                    const VelocityGrid& grid) {
  std::vector<std::string> errors;
  for (const Patch& p : mesh.patches)
    if (cfg.find_boundary(p.name) == nullptr) errors.push_back("mesh patch \"" + p.name + "\" has no boundary condition");
  for (const BoundarySpec& b : cfg.boundaries) {
    const int pi = mesh.find_patch(b.patch);
    if (pi < 0) {
      errors.push_back("boundary patch \"" + b.patch + "\" is not in the mesh");
      continue;
    }
    const Patch& patch = mesh.patches[static_cast<std::size_t>(pi)];
    if (patch.count == 0) continue;
    if (b.type == BoundaryType::PeriodicSlab) {
      // a periodic slab side pairs faces within each cell: their area vectors cancel per cell
      std::map<Index, std::pair<Vec3, double>> per_cell;
      for (Index f = patch.start; f < patch.end(); ++f) {
        auto& e = per_cell[mesh.owner[static_cast<std::size_t>(f)]];
        e.first += g.face_area_vector[static_cast<std::size_t>(f)];
        e.second += g.face_area[static_cast<std::size_t>(f)];
      }
      for (const auto& kv : per_cell)
        if (norm(kv.second.first) > 1e-9 * kv.second.second) {
          errors.push_back("periodic patch \"" + b.patch + "\" faces are not paired per cell");
          break;
        }
    } else if (b.type == BoundaryType::Symmetry) {
      const int axis = patch_normal_axis(g, patch);
      if (axis < 0)
        errors.push_back("symmetry patch \"" + b.patch + "\" normal is not aligned with one velocity axis");
      else if (!grid.symmetric(axis))
        errors.push_back("symmetry patch \"" + b.patch + "\": velocity range is not symmetric on its axis");
    }
  }
  return errors;
}

class SetupBuilder {
 public:
  SetupBuilder(const CaseConfig& cfg, const PolyMesh& mesh, const KineticParams& params)
      : cfg_(cfg), input_(mesh), grid_(cfg.vn, cfg.vlo, cfg.vhi, cfg.vstretch) {
    s_.params = params;
  }

  KineticSetup build() {
    case_facts();
    const MeshGeometry input_geometry = compute_geometry(input_);
    const std::vector<std::string> errors = mismatches(cfg_, input_, input_geometry, grid_);
    if (!errors.empty()) {
      std::string msg = "case and mesh do not match:";
      for (const std::string& e : errors) msg += "\n  " + e;
      throw Error(msg);
    }
    number_cells(input_geometry);
    boundary_conditions();
    maxwell_slots();
    velocities();
    cell_scalars();
    sweep_orders();
    face_tables();
    return std::move(s_);
  }

 private:
  const CaseConfig& cfg_;
  const PolyMesh& input_;
  const VelocityGrid grid_;
  KineticSetup s_;
  std::vector<int> mirror_axes_;

  // Case identity, scales and the per-run constants the backend reads.
  void case_facts() {
    const KineticParams& p = s_.params;
    s_.case_name = cfg_.name;
    s_.mesh_file = cfg_.mesh_path();
    s_.gsis_tolerance = cfg_.gsis_tolerance;
    s_.mass_balance_tolerance = cfg_.gsis_mass_balance_tolerance;
    s_.mass_balance_required = cfg_.gsis_mass_balance_required;
    s_.section_positions = cfg_.section_positions;
    s_.section_axis = cfg_.section_axis;
    s_.derived = derive_quantities(cfg_);
    s_.length_scale = cfg_.L_ref;
    s_.limiter_mode = p.limiter == "none" ? 0 : p.limiter_eps == "relative" ? 2 : 1;
    s_.discrete_eq = p.equilibrium == "discrete";
    s_.deq_tolerance = p.precision == "fp64" ? 1e-11 : 1e-6;
    for (int a = 0; a < 3; ++a) {
      s_.deq_lo[a] = grid_.lo(a);
      s_.deq_hi[a] = grid_.hi(a);
      s_.deq_d[a] = grid_.max_spacing(a);
    }
    s_.grid_n = cfg_.vn;
    s_.mu0 = s_.derived.mu_hat0;
    s_.omega = cfg_.omega;
    s_.Pr = cfg_.Pr;
    s_.gamma = cfg_.gamma;
  }

  // Block count: [kinetic].sweep_blocks, or n_cells / 2048 clamped to 1..64 when it is 0; never more
  // than the cells. "rcb" numbers cells block by block (bisection + block RCM); "rcm" and "none"
  // number the whole mesh and cut it into equal contiguous blocks. A renumbered mesh gets its
  // geometry recomputed (faces may be re-oriented), boundary faces keep their patch order.
  void number_cells(const MeshGeometry& input_geometry) {
    const std::string& ordering = s_.params.cell_ordering;
    const long long wanted = s_.params.sweep_blocks > 0 ? s_.params.sweep_blocks
                                                        : std::clamp<long long>(input_.n_cells / 2048, 1, 64);
    const int P = static_cast<int>(std::min<long long>(wanted, input_.n_cells));
    if (ordering == "none") {
      s_.mesh = input_;
      s_.geometry = input_geometry;
      s_.cell_new_to_old.resize(static_cast<std::size_t>(input_.n_cells));
      std::iota(s_.cell_new_to_old.begin(), s_.cell_new_to_old.end(), 0);
      s_.bface_new_to_old.resize(static_cast<std::size_t>(input_.n_boundary_faces()));
      std::iota(s_.bface_new_to_old.begin(), s_.bface_new_to_old.end(), 0);
    } else {
      s_.cell_new_to_old = ordering == "rcb" ? rcb_ordering(input_, input_geometry, P, s_.block_offsets)
                                             : rcm_ordering(input_);
      s_.mesh = renumber_cells(input_, s_.cell_new_to_old);
      s_.bface_new_to_old = match_boundary_faces(input_, s_.mesh);
      s_.geometry = compute_geometry(s_.mesh);
    }
    s_.cell_old_to_new = invert_permutation(s_.cell_new_to_old);
    s_.bface_old_to_new = invert_permutation(s_.bface_new_to_old);
    if (ordering != "rcb") s_.block_offsets = partition_contiguous(s_.mesh.n_cells, P);
    s_.csr = build_cell_faces(s_.mesh);
    s_.lsq = build_lsq_coefficients(s_.mesh, s_.geometry, s_.csr, s_.params.lsq_weight_power);
  }

  // One PatchCondition per mesh patch, in mesh order, from its [[boundary]] entry (values made
  // nondimensional by the reference scales). Inlets and outlets open the domain. A wall with
  // accommodation < 1 becomes a Maxwell wall; symmetry planes and the axis-aligned faces of
  // Maxwell walls add mirror axes to the velocity layout.
  void boundary_conditions() {
    const DerivedQuantities& ref = s_.derived;
    const Index first_bface = s_.mesh.n_internal_faces();
    s_.closed_domain = true;
    for (const Patch& patch : s_.mesh.patches) {
      const BoundarySpec& spec = *cfg_.find_boundary(patch.name);
      PatchCondition pc;
      pc.name = patch.name;
      pc.type = spec.type;
      pc.first_bface = patch.start - first_bface;
      pc.n_faces = patch.count;
      const bool wall = spec.type == BoundaryType::DiffuseWall || spec.type == BoundaryType::MaxwellWall;
      const bool inlet = spec.type == BoundaryType::PressureInlet || spec.type == BoundaryType::MassFlowInlet;
      if (wall) {
        pc.kind = kBcDiffuseWall;
        pc.accommodation = spec.accommodation;
        pc.T = *spec.temperature / ref.ref_temperature;
        if (spec.velocity) pc.u_wall = *spec.velocity / ref.ref_velocity;
      } else if (inlet) {
        pc.kind = kBcPressure;
        pc.T = *spec.temperature / ref.ref_temperature;
        pc.p = spec.pressure.value_or(ref.p_in) / ref.ref_pressure;
        pc.target_mass_flow = spec.mass_flow.value_or(0.0) / (ref.ref_mass_flux * ref.ref_length * ref.ref_length);
        pc.zero_tangential = spec.inlet_zero_tangential;
      } else if (spec.type == BoundaryType::PressureOutlet) {
        pc.p = *spec.pressure / ref.ref_pressure;
        pc.kind = pc.p > 0.0 ? kBcPressure : kBcVacuum;
      } else if (spec.type == BoundaryType::Vacuum) {
        pc.kind = kBcVacuum;
      } else if (spec.type == BoundaryType::Symmetry) {
        pc.kind = kBcSymmetry;
        if (patch.count > 0) {
          pc.mirror_axis = patch_normal_axis(s_.geometry, patch);
          GSK_CHECK(pc.mirror_axis >= 0, "symmetry patch \"" + patch.name + "\" normal is not axis aligned");
          mirror_axes_.push_back(pc.mirror_axis);
        }
      } else {
        pc.kind = kBcPeriodic;
      }
      if (pc.kind == kBcPressure || pc.kind == kBcVacuum) s_.closed_domain = false;
      if (wall && pc.accommodation < 1.0) {
        pc.kind = kBcMaxwellWall;
        for (int a = 0; a < 3; ++a) GSK_CHECK(grid_.n(a) >= 2, "a Maxwell wall needs at least 2 velocity points per axis");
        for (Index f = patch.start; f < patch.end(); ++f) {
          const int a = aligned_axis(s_.geometry.face_normal[static_cast<std::size_t>(f)]);
          if (a < 0) continue;
          GSK_CHECK(grid_.symmetric(a), "a Maxwell wall mirror axis needs a symmetric velocity range");
          mirror_axes_.push_back(a);
        }
      }
      if (wall && norm(pc.u_wall) > 0.0)
        for (Index f = patch.start; f < patch.end(); ++f)
          GSK_CHECK(std::abs(dot(pc.u_wall, s_.geometry.face_normal[static_cast<std::size_t>(f)])) <= 1e-9 * norm(pc.u_wall),
                    "wall \"" + patch.name + "\" velocity has a normal component");
      s_.patches.push_back(pc);
    }
    s_.bface_patch = s_.mesh.boundary_face_patches();
  }

  // Each Maxwell-wall face owns one row of nv reflected values in the backend's face cache; rows
  // are handed out in boundary-face order. Other faces: offset -1, specular fraction 0.
  void maxwell_slots() {
    for (const PatchCondition& pc : s_.patches)
      if (pc.kind == kBcMaxwellWall) s_.n_maxwell_faces += pc.n_faces;
    if (s_.n_maxwell_faces == 0) return;
    const auto nb = static_cast<std::size_t>(s_.n_bfaces());
    s_.maxwell_offset.assign(nb, -1);
    s_.maxwell_specular.assign(nb, 0.0);
    long long rows = 0;
    for (std::size_t b = 0; b < nb; ++b) {
      const PatchCondition& pc = s_.patches[static_cast<std::size_t>(s_.bface_patch[b])];
      if (pc.kind != kBcMaxwellWall) continue;
      s_.maxwell_offset[b] = rows * grid_.size();
      s_.maxwell_specular[b] = 1.0 - pc.accommodation;
      ++rows;
    }
  }

  void velocities() {
    s_.sweep_dirs = sweep_direction_table(s_.params.sweep_order);
    s_.n_sweep_dirs = static_cast<int>(s_.sweep_dirs.size());
    s_.velocity = layout_velocities(grid_, mirror_axes_, s_.sweep_dirs);
  }

  // Volumes in L_ref^3, the limiter's eps^2 = K^3 V, and the pseudo-time term: with
  // pseudo_cfl > 0, either V / dt = |xi|_max (sum of the cell's face areas) / cfl per cell
  // ("global"), or a diagonal scaled by 1 + 1 / cfl ("velocity").
  void cell_scalars() {
    const double L = s_.length_scale, K = s_.params.limiter_k;
    const auto nc = static_cast<std::size_t>(s_.n_cells());
    s_.volume.resize(nc);
    s_.eps2.resize(nc);
    s_.vdt.assign(nc, 0.0);
    for (std::size_t c = 0; c < nc; ++c) {
      s_.volume[c] = s_.geometry.cell_volume[c] / (L * L * L);
      s_.eps2[c] = K * K * K * s_.volume[c];
    }
    const double cfl = s_.params.pseudo_cfl;
    if (!(cfl > 0.0)) return;
    if (s_.params.pseudo_time == "velocity") {
      s_.diag_scale = 1.0 + 1.0 / cfl;
      return;
    }
    for (Index c = 0; c < s_.mesh.n_cells; ++c) {
      double surface = 0.0;
      for (Offset k = s_.csr.begin(c); k < s_.csr.end(c); ++k)
        surface += s_.geometry.face_area[static_cast<std::size_t>(s_.csr.faces[static_cast<std::size_t>(k)])];
      s_.vdt[static_cast<std::size_t>(c)] = s_.velocity.xi_max * surface / (L * L) / cfl;
    }
  }

  // Per sweep direction d, every block's cells in ascending d . x_centre (ties: lower internal
  // index first): sweep_order[d][i] is the i-th cell in sweep order (the cells of block b keep the
  // index range of block b) and sweep_pos its inverse. sweep_dir_of[p]: the direction velocity p
  // follows.
  void sweep_orders() {
    if (s_.n_sweep_dirs == 0) return;
    const auto nc = static_cast<std::size_t>(s_.n_cells());
    s_.sweep_order.resize(static_cast<std::size_t>(s_.n_sweep_dirs) * nc);
    s_.sweep_pos.resize(s_.sweep_order.size());
    std::vector<std::pair<double, std::int32_t>> ranked(nc);
    for (int d = 0; d < s_.n_sweep_dirs; ++d) {
      const Vec3& dir = s_.sweep_dirs[static_cast<std::size_t>(d)];
      for (std::size_t c = 0; c < nc; ++c) ranked[c] = {dot(dir, s_.geometry.cell_center[c]), static_cast<std::int32_t>(c)};
      for (int b = 0; b < s_.n_blocks(); ++b)  // pairs compare by (projection, index)
        std::sort(ranked.begin() + s_.block_offsets[static_cast<std::size_t>(b)],
                  ranked.begin() + s_.block_offsets[static_cast<std::size_t>(b) + 1]);
      std::int32_t* order = &s_.sweep_order[static_cast<std::size_t>(d) * nc];
      std::int32_t* pos = &s_.sweep_pos[static_cast<std::size_t>(d) * nc];
      for (std::size_t i = 0; i < nc; ++i) {
        order[i] = ranked[i].second;
        pos[static_cast<std::size_t>(ranked[i].second)] = static_cast<std::int32_t>(i);
      }
    }
    const VelocityLayout& v = s_.velocity;
    s_.sweep_dir_of.resize(static_cast<std::size_t>(v.nv));
    for (std::size_t p = 0; p < s_.sweep_dir_of.size(); ++p)
      s_.sweep_dir_of[p] =
          static_cast<std::uint8_t>(nearest_direction(s_.sweep_dirs, Vec3{v.xi_x[p], v.xi_y[p], v.xi_z[p]}));
  }

  // The backend's face tables in L_ref units: per CSR entry (cell i, face f) the outward area
  // vector, the offsets x_f - x_i and x_f - x_j, and the LSQ coefficient; per boundary face its
  // owner, kind, storage mirror (the symmetry plane's axis, or the axis an axis-aligned Maxwell
  // face is normal to), patch, normal, area and offset from the owner's centre.
  void face_tables() {
    const double L = s_.length_scale;
    const Index n_int = s_.mesh.n_internal_faces();
    const CellFaceCsr& csr = s_.csr;
    const MeshGeometry& g = s_.geometry;
    GSK_CHECK(csr.offsets.back() < (Offset{1} << 31), "too many cell->face entries");
    s_.csr_offsets.assign(csr.offsets.begin(), csr.offsets.end());
    s_.csr_entries.resize(csr.faces.size());
    for (Index c = 0; c < s_.mesh.n_cells; ++c)
      for (Offset k = csr.begin(c); k < csr.end(c); ++k) {
        const auto e = static_cast<std::size_t>(k);
        const auto f = static_cast<std::size_t>(csr.faces[e]);
        const Index nb = csr.other[e];
        const Vec3 area = static_cast<double>(csr.sign[e]) * g.face_area_vector[f] / (L * L);
        const Vec3 to_face = (g.face_center[f] - g.cell_center[static_cast<std::size_t>(c)]) / L;
        const Vec3 from_nb = nb == kInvalid ? Vec3{} : (g.face_center[f] - g.cell_center[static_cast<std::size_t>(nb)]) / L;
        const Vec3 lsq = s_.lsq[e] * L;
        gsk_csr_entry& out = s_.csr_entries[e];
        out.other = nb;
        out.bface = nb == kInvalid ? static_cast<int32_t>(csr.faces[e] - n_int) : -1;
        for (int a = 0; a < 3; ++a) {
          out.s[a] = area[a];
          out.ds[a] = to_face[a];
          out.dother[a] = from_nb[a];
          out.c[a] = lsq[a];
        }
      }
    s_.bfaces.resize(static_cast<std::size_t>(s_.n_bfaces()));
    for (std::size_t b = 0; b < s_.bfaces.size(); ++b) {
      const std::size_t f = b + static_cast<std::size_t>(n_int);
      const int patch = s_.bface_patch[b];
      const PatchCondition& pc = s_.patches[static_cast<std::size_t>(patch)];
      const Index owner = s_.mesh.owner[f];
      const Vec3& n = g.face_normal[f];
      const Vec3 d = (g.face_center[f] - g.cell_center[static_cast<std::size_t>(owner)]) / L;
      int mirror_axis = -1;
      if (pc.kind == kBcSymmetry) mirror_axis = pc.mirror_axis;
      if (pc.kind == kBcMaxwellWall) mirror_axis = aligned_axis(n);
      gsk_bface& out = s_.bfaces[b];
      out.owner = owner;
      out.kind = pc.kind;
      out.mirror_xor = mirror_axis >= 0 ? s_.velocity.axis_xor[static_cast<std::size_t>(mirror_axis)] : 0;
      out.patch = patch;
      for (int a = 0; a < 3; ++a) {
        out.n[a] = n[a];
        out.d[a] = d[a];
      }
      out.area = g.face_area[f] / (L * L);
    }
  }
};

}  // namespace

KineticSetup build_kinetic_setup(const CaseConfig& config, const PolyMesh& mesh, const KineticParams& params) {
  return SetupBuilder(config, mesh, params).build();
}

gsk_setup_view KineticSetup::view() const {
  gsk_setup_view v{};
  v.n_cells = n_cells();
  v.nv = nv();
  v.n_bfaces = n_bfaces();
  v.n_blocks = n_blocks();
  v.orbit_size = velocity.orbit_size;
  v.velocity_chunk = params.velocity_chunk;
  v.xi_x = velocity.xi_x.data();
  v.xi_y = velocity.xi_y.data();
  v.xi_z = velocity.xi_z.data();
  v.weight = velocity.weight;
  v.csr_offsets = csr_offsets.data();
  v.csr = csr_entries.data();
  v.eps2 = eps2.data();
  v.volume = volume.data();
  v.vdt = vdt.data();
  v.bfaces = bfaces.data();
  v.block_offsets = block_offsets.data();
  v.n_sweep_dirs = n_sweep_dirs;
  v.sweep_order = n_sweep_dirs > 0 ? sweep_order.data() : nullptr;
  v.sweep_pos = n_sweep_dirs > 0 ? sweep_pos.data() : nullptr;
  v.sweep_dir_of = n_sweep_dirs > 0 ? sweep_dir_of.data() : nullptr;
  v.grid_to_storage = velocity.grid_to_storage.data();
  for (int a = 0; a < 3; ++a) {
    v.grid_n[a] = grid_n[static_cast<std::size_t>(a)];
    v.deq_lo[a] = deq_lo[a];
    v.deq_hi[a] = deq_hi[a];
    v.deq_d[a] = deq_d[a];
  }
  v.n_maxwell_faces = n_maxwell_faces;
  v.maxwell_offset = n_maxwell_faces > 0 ? maxwell_offset.data() : nullptr;
  v.maxwell_specular = n_maxwell_faces > 0 ? maxwell_specular.data() : nullptr;
  v.deq_criterion = deq_criterion;
  v.deq_max_iterations = deq_max_iterations;
  v.deq_tolerance = deq_tolerance;
  v.discrete_eq = discrete_eq ? 1 : 0;
  v.deq_fallback_collisionless = params.deq_fallback == "collisionless" ? 1 : 0;
  v.limiter_mode = limiter_mode;
  v.positivity = params.positivity_limiter ? 1 : 0;
  v.diag_scale = diag_scale;
  v.interface_damping = params.interface_damping;
  v.cold_bgk_temperature = params.cold_bgk_temperature;
  v.mu0 = mu0;
  v.omega = omega;
  v.Pr = Pr;
  v.block_passes = static_cast<int32_t>(params.block_passes);
  const std::string& rule = params.gsis_negative_check;
  v.negative_check = rule == "created" ? 1 : rule == "limit" ? 2 : rule == "blend" ? 3 : 0;
  v.precision_fp64 = params.precision == "fp64" ? 1 : 0;
  v.weight_rel = velocity.weight_rel.empty() ? nullptr : velocity.weight_rel.data();
  v.grid_nodes = velocity.nodes.empty() ? nullptr : velocity.nodes.data();
  v.moments_mfma = params.moment_sums == "mfma" ? 1 : 0;
  return v;
}

KineticSetup setup_from_case(const std::string& toml_path, const std::string& overrides) {
  CaseConfig cfg = load_case(toml_path);
  if (!overrides.empty()) {
    const TomlDoc o = parse_toml("[kinetic]\n" + overrides + "\n", "<overrides>");
    for (const auto& kv : o.tables.at("kinetic")) cfg.kinetic[kv.first] = kv.second;
  }
  const KineticParams params = parse_kinetic_params(cfg.kinetic);
  const PolyMesh mesh = read_pmesh(cfg.mesh_path());
  return build_kinetic_setup(cfg, mesh, params);
}

}  // namespace gsk
