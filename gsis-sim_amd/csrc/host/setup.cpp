// Kinetic setup (see setup.hpp for the reference locations followed).

#include "setup.hpp"

#include <algorithm>
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

std::vector<Vec3> make_sweep_directions(const std::string& order) {
  std::vector<Vec3> d;
  if (order == "octant" || order == "dir13") {
    if (order == "dir13") {
      for (const Vec3& v : {Vec3{1, 0, 0}, Vec3{0, 1, 0}, Vec3{0, 0, 1}, Vec3{1, 1, 0}, Vec3{1, -1, 0}, Vec3{1, 0, 1},
                            Vec3{1, 0, -1}, Vec3{0, 1, 1}, Vec3{0, 1, -1}})
        d.push_back(v);
    }
    for (const Vec3& v : {Vec3{1, 1, 1}, Vec3{1, 1, -1}, Vec3{1, -1, 1}, Vec3{1, -1, -1}}) d.push_back(v);
  } else {
    GSK_CHECK(order == "block", "unknown sweep_order \"" + order + "\"");
  }
  for (Vec3& x : d) x = x / norm(x);
  return d;
}

int sweep_direction_of(const std::vector<Vec3>& dirs, const Vec3& xi) {
  int best = 0;
  double best_v = -1.0;
  for (std::size_t g = 0; g < dirs.size(); ++g) {
    const double v = std::abs(dot(dirs[g], xi));
    if (v > best_v + 1e-12) {
      best_v = v;
      best = static_cast<int>(g);
    }
  }
  return best;
}

VelocityLayout make_velocity_layout(const VelocityGrid& grid, const std::vector<int>& mirror_axes,
                                    const std::vector<Vec3>& sweep_dirs) {
  VelocityLayout L;
  L.nv = grid.size();
  L.weight = grid.weight();
  std::vector<int> axes = mirror_axes;
  std::sort(axes.begin(), axes.end());
  axes.erase(std::unique(axes.begin(), axes.end()), axes.end());
  for (const int a : axes) {
    GSK_CHECK(a >= 0 && a < 3, "invalid mirror axis");
    GSK_CHECK(grid.symmetric(a), strf("velocity range of axis %c is not symmetric about 0; mirrors unavailable", "xyz"[a]));
    GSK_CHECK(grid.n(a) % 2 == 0, strf("axis %c of a mirrored velocity grid needs an even point count", "xyz"[a]));
  }
  const int na = static_cast<int>(axes.size());
  L.orbit_size = 1 << na;
  for (int r = 0; r < na; ++r) L.axis_xor[static_cast<std::size_t>(axes[static_cast<std::size_t>(r)])] = 1 << r;
  L.storage_to_grid.reserve(static_cast<std::size_t>(L.nv));
  for (int k = 0; k < L.nv; ++k) {
    const std::array<int, 3> ijk = grid.multi_index(k);
    bool rep = true;
    for (const int a : axes)
      if (ijk[static_cast<std::size_t>(a)] >= grid.n(a) / 2) rep = false;
    if (!rep) continue;
    for (int m = 0; m < L.orbit_size; ++m) {
      std::array<int, 3> mm = ijk;
      for (int r = 0; r < na; ++r) {
        if (m & (1 << r)) {
          const auto a = static_cast<std::size_t>(axes[static_cast<std::size_t>(r)]);
          mm[a] = grid.n(static_cast<int>(a)) - 1 - mm[a];
        }
      }
      L.storage_to_grid.push_back(grid.index(mm[0], mm[1], mm[2]));
    }
  }
  GSK_CHECK(static_cast<int>(L.storage_to_grid.size()) == L.nv, "internal error: velocity layout size");
  if (na == 0 && !sweep_dirs.empty()) {
    std::vector<int> group(static_cast<std::size_t>(L.nv));
    for (int k = 0; k < L.nv; ++k) group[static_cast<std::size_t>(k)] = sweep_direction_of(sweep_dirs, grid.xi(k));
    std::stable_sort(L.storage_to_grid.begin(), L.storage_to_grid.end(),
                     [&](int a, int b) { return group[static_cast<std::size_t>(a)] < group[static_cast<std::size_t>(b)]; });
  }
  L.grid_to_storage.assign(static_cast<std::size_t>(L.nv), -1);
  L.xi_x.resize(static_cast<std::size_t>(L.nv));
  L.xi_y.resize(static_cast<std::size_t>(L.nv));
  L.xi_z.resize(static_cast<std::size_t>(L.nv));
  for (int p = 0; p < L.nv; ++p) {
    const int k = L.storage_to_grid[static_cast<std::size_t>(p)];
    L.grid_to_storage[static_cast<std::size_t>(k)] = p;
    const Vec3 xi = grid.xi(k);
    L.xi_x[static_cast<std::size_t>(p)] = xi.x;
    L.xi_y[static_cast<std::size_t>(p)] = xi.y;
    L.xi_z[static_cast<std::size_t>(p)] = xi.z;
    L.xi_max = std::max(L.xi_max, norm(xi));
  }
  if (grid.stretched()) {
    L.weight_rel.resize(static_cast<std::size_t>(L.nv));
    for (int p = 0; p < L.nv; ++p)
      L.weight_rel[static_cast<std::size_t>(p)] = grid.relative_weight(L.storage_to_grid[static_cast<std::size_t>(p)]);
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < grid.n(a); ++i) L.nodes.push_back(grid.coordinate(a, i));
  }
  for (const int a : axes)
    for (int p = 0; p < L.nv; ++p) {
      const int q = p ^ L.axis_xor[static_cast<std::size_t>(a)];
      GSK_CHECK(L.storage_to_grid[static_cast<std::size_t>(q)] == grid.mirror(a, L.storage_to_grid[static_cast<std::size_t>(p)]),
                "internal error: mirror is not a storage xor");
    }
  return L;
}

int symmetry_patch_axis(const MeshGeometry& g, const Patch& patch, double tol) {
  int axis = -1;
  for (Index f = patch.start; f < patch.end(); ++f) {
    const Vec3& n = g.face_normal[static_cast<std::size_t>(f)];
    int a = -1;
    for (int k = 0; k < 3; ++k)
      if (std::abs(std::abs(n[k]) - 1.0) <= tol) a = k;
    if (a < 0 || (axis >= 0 && a != axis)) return -1;
    axis = a;
  }
  return axis;
}

namespace {

std::vector<std::string> check_case_against_mesh(const CaseConfig& cfg, const PolyMesh& mesh, const MeshGeometry& g,
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
      continue;
    }
    if (b.type != BoundaryType::Symmetry) continue;
    const int axis = symmetry_patch_axis(g, patch);
    if (axis < 0)
      errors.push_back("symmetry patch \"" + b.patch + "\" normal is not aligned with one velocity axis");
    else if (!grid.symmetric(axis))
      errors.push_back("symmetry patch \"" + b.patch + "\": velocity range is not symmetric on its axis");
  }
  return errors;
}

}  // namespace

KineticSetup build_kinetic_setup(const CaseConfig& config, const PolyMesh& mesh, const KineticParams& params) {
  KineticSetup s;
  s.params = params;
  s.case_name = config.name;
  s.mesh_file = config.mesh_path();
  s.gsis_tolerance = config.gsis_tolerance;
  s.mass_balance_tolerance = config.gsis_mass_balance_tolerance;
  s.mass_balance_required = config.gsis_mass_balance_required;
  s.section_positions = config.section_positions;
  s.section_axis = config.section_axis;
  s.limiter_mode = params.limiter == "none" ? 0 : (params.limiter_eps == "relative" ? 2 : 1);
  s.derived = derive_quantities(config);
  s.length_scale = config.L_ref;
  const DerivedQuantities& ref = s.derived;
  const VelocityGrid grid(config.vn, config.vlo, config.vhi, config.vstretch);
  const MeshGeometry geometry0 = compute_geometry(mesh);
  {
    const std::vector<std::string> errors = check_case_against_mesh(config, mesh, geometry0, grid);
    if (!errors.empty()) {
      std::string msg = "case and mesh do not match:";
      for (const std::string& e : errors) msg += "\n  " + e;
      throw Error(msg);
    }
  }
  long long P_req = params.sweep_blocks;
  if (P_req == 0) P_req = std::clamp<long long>(mesh.n_cells / 2048, 1, 64);
  const int P = static_cast<int>(std::min<long long>(P_req, mesh.n_cells));
  if (params.cell_ordering == "rcm" || params.cell_ordering == "rcb") {
    if (params.cell_ordering == "rcb")
      s.cell_new_to_old = rcb_ordering(mesh, geometry0, P, s.block_offsets);
    else
      s.cell_new_to_old = rcm_ordering(mesh);
    s.mesh = renumber_cells(mesh, s.cell_new_to_old);
    s.cell_old_to_new = invert_permutation(s.cell_new_to_old);
    s.bface_new_to_old = match_boundary_faces(mesh, s.mesh);
    s.bface_old_to_new = invert_permutation(s.bface_new_to_old);
    s.geometry = compute_geometry(s.mesh);
  } else {
    s.mesh = mesh;
    s.geometry = geometry0;
    s.cell_new_to_old.resize(static_cast<std::size_t>(mesh.n_cells));
    std::iota(s.cell_new_to_old.begin(), s.cell_new_to_old.end(), 0);
    s.cell_old_to_new = s.cell_new_to_old;
    s.bface_new_to_old.resize(static_cast<std::size_t>(mesh.n_boundary_faces()));
    std::iota(s.bface_new_to_old.begin(), s.bface_new_to_old.end(), 0);
    s.bface_old_to_new = s.bface_new_to_old;
  }
  s.csr = build_cell_faces(s.mesh);
  s.lsq = build_lsq_coefficients(s.mesh, s.geometry, s.csr, params.lsq_weight_power);

  const Index n_int = s.mesh.n_internal_faces();
  std::vector<int> mirror_axes;
  s.closed_domain = true;
  for (const Patch& patch : s.mesh.patches) {
    const BoundarySpec* b = config.find_boundary(patch.name);
    PatchCondition pc;
    pc.name = patch.name;
    pc.type = b->type;
    pc.first_bface = patch.start - n_int;
    pc.n_faces = patch.count;
    switch (b->type) {
      case BoundaryType::MaxwellWall:
      case BoundaryType::DiffuseWall:
        pc.accommodation = b->accommodation;
        pc.kind = kBcDiffuseWall;
        pc.T = *b->temperature / ref.ref_temperature;
        if (b->velocity) pc.u_wall = *b->velocity / ref.ref_velocity;
        break;
      case BoundaryType::MassFlowInlet:
      case BoundaryType::PressureInlet:
        pc.target_mass_flow = b->mass_flow.value_or(0.0) / (ref.ref_mass_flux * ref.ref_length * ref.ref_length);
        pc.kind = kBcPressure;
        pc.T = *b->temperature / ref.ref_temperature;
        pc.p = b->pressure.value_or(ref.p_in) / ref.ref_pressure;
        pc.zero_tangential = b->inlet_zero_tangential;
        s.closed_domain = false;
        break;
      case BoundaryType::PressureOutlet:
        pc.p = *b->pressure / ref.ref_pressure;
        pc.kind = pc.p > 0.0 ? kBcPressure : kBcVacuum;
        s.closed_domain = false;
        break;
      case BoundaryType::Vacuum:
        pc.kind = kBcVacuum;
        s.closed_domain = false;
        break;
      case BoundaryType::Symmetry:
        pc.kind = kBcSymmetry;
        if (patch.count > 0) {
          pc.mirror_axis = symmetry_patch_axis(s.geometry, patch);
          GSK_CHECK(pc.mirror_axis >= 0, "symmetry patch \"" + patch.name + "\" normal is not axis aligned");
          mirror_axes.push_back(pc.mirror_axis);
        }
        break;
      case BoundaryType::PeriodicSlab:
        pc.kind = kBcPeriodic;
        break;
    }
    if (pc.accommodation < 1.0) {
      pc.kind = kBcMaxwellWall;
      for (int a = 0; a < 3; ++a) GSK_CHECK(grid.n(a) >= 2, "a Maxwell wall needs at least 2 velocity points per axis");
      for (Index f = patch.start; f < patch.end(); ++f) {
        const Vec3& n = s.geometry.face_normal[static_cast<std::size_t>(f)];
        for (int a = 0; a < 3; ++a)
          if (std::abs(std::abs(n[a]) - 1.0) < 1e-10) {
            GSK_CHECK(grid.symmetric(a), "a Maxwell wall mirror axis needs a symmetric velocity range");
            mirror_axes.push_back(a);
          }
      }
    }
    if ((pc.kind == kBcDiffuseWall || pc.kind == kBcMaxwellWall) && norm(pc.u_wall) > 0.0)
      for (Index f = patch.start; f < patch.end(); ++f) {
        const double un = dot(pc.u_wall, s.geometry.face_normal[static_cast<std::size_t>(f)]);
        GSK_CHECK(std::abs(un) <= 1e-9 * norm(pc.u_wall), "wall \"" + patch.name + "\" velocity has a normal component");
      }
    s.patches.push_back(pc);
  }
  s.bface_patch = s.mesh.boundary_face_patches();
  s.grid_n = config.vn;
  for (const auto& pc : s.patches)
    if (pc.kind == kBcMaxwellWall) s.n_maxwell_faces += pc.n_faces;
  if (s.n_maxwell_faces > 0) {
    s.maxwell_offset.assign(static_cast<std::size_t>(s.n_bfaces()), -1);
    s.maxwell_specular.assign(static_cast<std::size_t>(s.n_bfaces()), 0.0);
    long long slot = 0;
    for (int b = 0; b < s.n_bfaces(); ++b) {
      const auto& pc = s.patches[static_cast<std::size_t>(s.bface_patch[static_cast<std::size_t>(b)])];
      if (pc.kind == kBcMaxwellWall) {
        s.maxwell_offset[static_cast<std::size_t>(b)] = (slot++) * grid.size();
        s.maxwell_specular[static_cast<std::size_t>(b)] = 1.0 - pc.accommodation;
      }
    }
  }
  s.discrete_eq = params.equilibrium == "discrete";
  for (int a = 0; a < 3; ++a) {
    s.deq_lo[a] = grid.lo(a);
    s.deq_hi[a] = grid.hi(a);
    s.deq_d[a] = grid.max_spacing(a);
  }
  s.deq_tolerance = params.precision == "fp64" ? 1e-11 : 1e-6;
  s.sweep_dirs = make_sweep_directions(params.sweep_order);
  s.velocity = make_velocity_layout(grid, mirror_axes, s.sweep_dirs);

  const double L = s.length_scale;
  const auto nc = static_cast<std::size_t>(s.mesh.n_cells);
  s.volume.resize(nc);
  s.vdt.assign(nc, 0.0);
  for (std::size_t c = 0; c < nc; ++c) s.volume[c] = s.geometry.cell_volume[c] / (L * L * L);
  if (params.pseudo_cfl > 0.0 && params.pseudo_time == "velocity") {
    s.diag_scale = 1.0 + 1.0 / params.pseudo_cfl;
  } else if (params.pseudo_cfl > 0.0) {
    for (Index c = 0; c < s.mesh.n_cells; ++c) {
      double area = 0.0;
      for (Offset k = s.csr.begin(c); k < s.csr.end(c); ++k)
        area += s.geometry.face_area[static_cast<std::size_t>(s.csr.faces[static_cast<std::size_t>(k)])];
      s.vdt[static_cast<std::size_t>(c)] = s.velocity.xi_max * area / (L * L) / params.pseudo_cfl;
    }
  }
  if (params.cell_ordering != "rcb") s.block_offsets = partition_contiguous(s.mesh.n_cells, P);

  s.n_sweep_dirs = static_cast<int>(s.sweep_dirs.size());
  if (s.n_sweep_dirs > 0) {
    s.sweep_order.resize(static_cast<std::size_t>(s.n_sweep_dirs) * nc);
    s.sweep_pos.resize(static_cast<std::size_t>(s.n_sweep_dirs) * nc);
    std::vector<double> key(nc);
    for (int g = 0; g < s.n_sweep_dirs; ++g) {
      const Vec3& d = s.sweep_dirs[static_cast<std::size_t>(g)];
      for (std::size_t c = 0; c < nc; ++c) key[c] = dot(d, s.geometry.cell_center[c]);
      std::int32_t* order = s.sweep_order.data() + static_cast<std::size_t>(g) * nc;
      std::int32_t* pos = s.sweep_pos.data() + static_cast<std::size_t>(g) * nc;
      for (std::size_t b = 0; b + 1 < s.block_offsets.size(); ++b) {
        const Index c0 = s.block_offsets[b], c1 = s.block_offsets[b + 1];
        for (Index c = c0; c < c1; ++c) order[c] = c;
        std::sort(order + c0, order + c1, [&](std::int32_t a, std::int32_t bb) {
          const double ka = key[static_cast<std::size_t>(a)], kb = key[static_cast<std::size_t>(bb)];
          return ka != kb ? ka < kb : a < bb;
        });
      }
      for (std::size_t k = 0; k < nc; ++k) pos[static_cast<std::size_t>(order[k])] = static_cast<std::int32_t>(k);
    }
    s.sweep_dir_of.resize(static_cast<std::size_t>(s.velocity.nv));
    for (int p = 0; p < s.velocity.nv; ++p) {
      const auto pi = static_cast<std::size_t>(p);
      s.sweep_dir_of[pi] = static_cast<std::uint8_t>(
          sweep_direction_of(s.sweep_dirs, Vec3{s.velocity.xi_x[pi], s.velocity.xi_y[pi], s.velocity.xi_z[pi]}));
    }
  }
  s.mu0 = s.derived.mu_hat0;
  s.omega = config.omega;
  s.Pr = config.Pr;
  s.gamma = config.gamma;

  // backend tables (make_csr_entries / make_bface_infos / make_eps2)
  GSK_CHECK(s.csr.offsets.back() < (Offset{1} << 31), "too many cell->face entries");
  s.csr_offsets.assign(s.csr.offsets.begin(), s.csr.offsets.end());
  s.csr_entries.resize(s.csr.faces.size());
  for (Index c = 0; c < s.mesh.n_cells; ++c) {
    const Vec3& xc = s.geometry.cell_center[static_cast<std::size_t>(c)];
    for (Offset k = s.csr.begin(c); k < s.csr.end(c); ++k) {
      const auto ki = static_cast<std::size_t>(k);
      const Index f = s.csr.faces[ki];
      const Index other = s.csr.other[ki];
      const Vec3& xf = s.geometry.face_center[static_cast<std::size_t>(f)];
      const Vec3 S = static_cast<double>(s.csr.sign[ki]) * s.geometry.face_area_vector[static_cast<std::size_t>(f)] / (L * L);
      const Vec3 ds = (xf - xc) / L;
      Vec3 dother{};
      if (other != kInvalid) dother = (xf - s.geometry.cell_center[static_cast<std::size_t>(other)]) / L;
      const Vec3 coef = s.lsq[ki] * L;
      gsk_csr_entry& e = s.csr_entries[ki];
      e.other = other;
      e.bface = other == kInvalid ? f - n_int : -1;
      for (int a = 0; a < 3; ++a) {
        e.s[a] = S[a];
        e.ds[a] = ds[a];
        e.dother[a] = dother[a];
        e.c[a] = coef[a];
      }
    }
  }
  s.bfaces.resize(static_cast<std::size_t>(s.n_bfaces()));
  for (Index b = 0; b < s.n_bfaces(); ++b) {
    const auto f = static_cast<std::size_t>(b + n_int);
    const int pi = s.bface_patch[static_cast<std::size_t>(b)];
    const PatchCondition& pc = s.patches[static_cast<std::size_t>(pi)];
    const Index owner = s.mesh.owner[f];
    const Vec3 dx = (s.geometry.face_center[f] - s.geometry.cell_center[static_cast<std::size_t>(owner)]) / L;
    const Vec3& n = s.geometry.face_normal[f];
    gsk_bface& o = s.bfaces[static_cast<std::size_t>(b)];
    o.owner = owner;
    o.kind = pc.kind;
    o.mirror_xor = pc.kind == kBcSymmetry ? s.velocity.axis_xor[static_cast<std::size_t>(pc.mirror_axis)] : 0;
    if (pc.kind == kBcMaxwellWall)
      for (int a = 0; a < 3; ++a)
        if (std::abs(std::abs(n[a]) - 1.0) < 1e-10) o.mirror_xor = s.velocity.axis_xor[static_cast<std::size_t>(a)];
    o.patch = pi;
    for (int a = 0; a < 3; ++a) {
      o.n[a] = n[a];
      o.d[a] = dx[a];
    }
    o.area = s.geometry.face_area[f] / (L * L);
  }
  const double K = params.limiter_k;
  s.eps2.resize(nc);
  for (std::size_t c = 0; c < nc; ++c) s.eps2[c] = K * K * K * s.volume[c];
  return s;
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
