// Boundary state of the CIS iteration (see boundary.hpp for the reference locations followed).

#include "boundary.hpp"

#include <algorithm>
#include <cmath>

#include "../kernels/kinetic_math.hpp"

namespace gsk {

namespace {

dev::DeqGrid deq_grid(const KineticSetup& s, double tolerance) {
  dev::DeqGrid g{};
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = s.deq_lo[a];
    g.hi[a] = s.deq_hi[a];
    g.d[a] = s.deq_d[a];
  }
  g.criterion = s.deq_criterion;
  g.max_iterations = s.deq_max_iterations;
  g.tolerance = tolerance;
  return g;
}

// Newton on the five exponents beta of M_d = exp(beta_0 + beta . c + beta_4 |c|^2) (c = xi - u)
// until its discrete moments are (rho, 0, 3 rho T / 2), with the same step control and accept rule
// as the cells' discrete equilibrium (kinetic_math.hpp). Velocity sums run in storage order.
// Returns true with the accepted Maxwellian in m.
bool discrete_maxwellian(const KineticSetup& s, const dev::DeqGrid& g, const gsk_inflow_state& st,
                         dev::Maxw<double>& m) {
  double beta[5];
  dev::deq_initial_beta(st.rho, st.T, beta);
  double previous = 1e300;
  const VelocityLayout& v = s.velocity;
  const bool stretched = !v.weight_rel.empty();
  for (int it = 0; it < g.max_iterations; ++it) {
    const dev::Maxw<double> trial = dev::deq_params<double>(beta, st.ux, st.uy, st.uz);
    double sums[dev::kDeqSums] = {};
    double terms[dev::kDeqSums];
    for (std::size_t p = 0; p < v.xi_x.size(); ++p) {
      const double x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
      double e = dev::maxw_eval(trial, x, y, z);
      if (stretched) e *= v.weight_rel[p];
      dev::deq_terms<double>(e, 0.0, x - st.ux, y - st.uy, z - st.uz, terms);
      for (int k = 0; k < dev::kDeqSums; ++k) sums[k] += terms[k];
    }
    double next[5] = {beta[0], beta[1], beta[2], beta[3], beta[4]};
    double rel = 0.0;
    const int step = dev::deq_newton_step(sums, v.weight, st.rho, st.T, g.tolerance, next, rel);
    if (dev::deq_accept(step, rel, previous, it + 1 == g.max_iterations)) {
      m = trial;
      return true;
    }
    if (step < 0) return false;
    std::copy(next, next + 5, beta);
    previous = rel;
  }
  return false;
}

}  // namespace

void discrete_inflow_state(const KineticSetup& s, double tolerance, gsk_inflow_state& st) {
  if (!s.discrete_eq || !(st.rho > 0.0) || !(st.T > 0.0)) return;
  const dev::DeqGrid g = deq_grid(s, tolerance);
  if (dev::deq_resolved(g, st.ux, st.uy, st.uz, st.T)) return;
  dev::Maxw<double> m;
  if (!discrete_maxwellian(s, g, st, m)) return;
  // the Maxwellian's (A, B, u) back to (rho, u, T)
  st.T = 1.0 / (2.0 * m.B);
  const double tpt = 2.0 * kPi * st.T;
  st.rho = m.A * tpt * std::sqrt(tpt);
  st.ux = m.ux;
  st.uy = m.uy;
  st.uz = m.uz;
}

BoundaryModel::BoundaryModel(const KineticSetup& s) : s_(s), patches_(s.patches) {
  const auto nb = static_cast<std::size_t>(s.n_bfaces());
  inflow.assign(nb, gsk_inflow_state{0, 0, 0, 0, 0, 1.0});
  flux_out.assign(nb, {});
  flux_in.assign(nb, {});
  face_state.assign(nb, {});
  patch_flow.assign(patches_.size(), 0.0);
}

void BoundaryModel::reset(double inflow_tolerance) {
  has_face_state_ = false;
  // the unit-density wall Maxwellian is the same on every face of a wall patch: build it per patch
  std::vector<gsk_inflow_state> wall_state(patches_.size());
  for (std::size_t pi = 0; pi < patches_.size(); ++pi) {
    const PatchCondition& pc = patches_[pi];
    if (pc.kind != kBcDiffuseWall && pc.kind != kBcMaxwellWall) continue;
    wall_state[pi] = gsk_inflow_state{1.0, 1.0, pc.u_wall.x, pc.u_wall.y, pc.u_wall.z, pc.T};
    discrete_inflow_state(s_, inflow_tolerance, wall_state[pi]);
  }
  for (Index b = 0; b < s_.n_bfaces(); ++b) {
    const auto pi = static_cast<std::size_t>(s_.bface_patch[static_cast<std::size_t>(b)]);
    const PatchCondition& pc = patches_[pi];
    gsk_inflow_state& st = inflow[static_cast<std::size_t>(b)];
    if (pc.kind == kBcDiffuseWall || pc.kind == kBcMaxwellWall)
      st = wall_state[pi];
    else if (pc.kind == kBcPressure)
      st = gsk_inflow_state{1.0, pc.p / pc.T, 0.0, 0.0, 0.0, pc.T};
    else
      st = gsk_inflow_state{0, 0, 0, 0, 0, 1.0};
  }
  if (s_.n_maxwell_faces == 0) return;
  wall_unit.assign(static_cast<std::size_t>(s_.n_bfaces()), {});
  for (Index b = 0; b < s_.n_bfaces(); ++b)
    if (patch_of(b).kind == kBcMaxwellWall) wall_unit[static_cast<std::size_t>(b)] = unit_wall_sums(b);
}

void BoundaryModel::seed_wall_densities(const std::vector<double>& moments) {
  const Index n_int = s_.mesh.n_internal_faces();
  for (Index b = 0; b < s_.n_bfaces(); ++b) {
    const PatchCondition& pc = patch_of(b);
    if (pc.kind != kBcDiffuseWall && pc.kind != kBcMaxwellWall) continue;
    const double* cell = &moments[static_cast<std::size_t>(s_.mesh.owner[static_cast<std::size_t>(b + n_int)]) * 15];
    inflow[static_cast<std::size_t>(b)].scale = pc.accommodation * cell[0] * std::sqrt(cell[4] / pc.T);
  }
}

// (p, u, T) at the face: the owner's value plus the limited LSQ gradient along x_face - x_owner.
// Neighbours are the adjacent cells, and across boundary faces the face state of the previous round
// (when it has a positive density and temperature; else the owner's own value). The limiter is
// Venkatakrishnan's with the absolute eps^2 = K^3 V and the periodic faces left out. A reconstruction
// that is not positive, or [kinetic].pressure_boundary_reconstruction = "cell", keeps the owner's
// value.
std::array<double, 5> BoundaryModel::reconstruct_face_state(Index b, const std::vector<double>& moments) const {
  using W = std::array<double, 5>;
  const Index n_int = s_.mesh.n_internal_faces();
  const auto face = static_cast<std::size_t>(b + n_int);
  const Index owner = s_.mesh.owner[face];
  auto cell_state = [&moments](Index c) {
    const double* m = &moments[static_cast<std::size_t>(c) * 15];
    return W{m[5], m[1], m[2], m[3], m[4]};
  };
  const W own = cell_state(owner);
  if (s_.params.pressure_boundary_reconstruction == "cell") return own;
  const Vec3& x_own = s_.geometry.cell_center[static_cast<std::size_t>(owner)];
  const Offset k0 = s_.csr.begin(owner), k1 = s_.csr.end(owner);
  std::vector<W> nb_state;
  std::vector<char> periodic;
  for (Offset k = k0; k < k1; ++k) {
    const auto e = static_cast<std::size_t>(k);
    const Index nb = s_.csr.other[e];
    if (nb != kInvalid) {
      nb_state.push_back(cell_state(nb));
      periodic.push_back(0);
      continue;
    }
    const auto bk = static_cast<std::size_t>(s_.csr.faces[e] - n_int);
    const bool is_periodic = patches_[static_cast<std::size_t>(s_.bface_patch[bk])].kind == kBcPeriodic;
    const auto& fs = face_state[bk];
    const bool usable = has_face_state_ && !is_periodic && fs[0] > 0.0 && fs[4] > 0.0;
    nb_state.push_back(usable ? W{fs[0] * fs[4], fs[1], fs[2], fs[3], fs[4]} : own);
    periodic.push_back(is_periodic ? 1 : 0);
  }
  const Vec3 to_face = s_.geometry.face_center[face] - x_own;
  const double K = s_.params.limiter_k;
  const double eps2 = K * K * K * s_.volume[static_cast<std::size_t>(owner)];
  const bool record = branches_ == Branches::Record, follow = branches_ == Branches::Follow;
  FaceBranch* fb = (record || follow) ? &face_branch(b) : nullptr;
  W out{};
  for (std::size_t q = 0; q < 5; ++q) {
    Vec3 grad{};
    double hi = own[q], lo = own[q];
    int hi_slot = -1, lo_slot = -1;
    for (std::size_t j = 0; j < nb_state.size(); ++j) {
      grad += (nb_state[j][q] - own[q]) * s_.lsq[static_cast<std::size_t>(k0) + j];
      if (hi < nb_state[j][q]) {
        hi = nb_state[j][q];
        hi_slot = static_cast<int>(j);
      }
      if (nb_state[j][q] < lo) {
        lo = nb_state[j][q];
        lo_slot = static_cast<int>(j);
      }
    }
    auto face_offset = [&](std::size_t j) {
      return s_.geometry.face_center[static_cast<std::size_t>(s_.csr.faces[static_cast<std::size_t>(k0) + j])] - x_own;
    };
    double phi = 1.0;
    if (follow) {
      hi = fb->hi_slot[q] < 0 ? own[q] : nb_state[static_cast<std::size_t>(fb->hi_slot[q])][q];
      lo = fb->lo_slot[q] < 0 ? own[q] : nb_state[static_cast<std::size_t>(fb->lo_slot[q])][q];
      const int j = fb->phi_slot[q];
      if (j >= 0)
        phi = dev::venkat_frozen<double>(dot(grad, face_offset(static_cast<std::size_t>(j))),
                                         fb->positive_side[q] ? hi - own[q] : lo - own[q], eps2, fb->clamp[q]);
    } else if (s_.limiter_mode != 0) {
      int phi_slot = -1;
      unsigned phi_clamp = 0;
      bool phi_positive = false;
      for (std::size_t j = 0; j < nb_state.size(); ++j) {
        if (periodic[j]) continue;
        const double d2 = dot(grad, face_offset(j));
        unsigned clamp = 0;
        const double v = dev::venkat_tracked<double>(d2, hi - own[q], lo - own[q], eps2, clamp);
        if (v < phi) {  // std::min(phi, v) keeps phi on ties
          phi = v;
          phi_slot = static_cast<int>(j);
          phi_clamp = clamp;
          phi_positive = d2 > 0.0;
        }
      }
      if (record) {
        fb->phi_slot[q] = phi_slot;
        fb->clamp[q] = phi_clamp;
        fb->positive_side[q] = phi_positive;
        fb->hi_slot[q] = hi_slot;
        fb->lo_slot[q] = lo_slot;
      }
    } else if (record) {
      fb->phi_slot[q] = -1;
      fb->hi_slot[q] = hi_slot;
      fb->lo_slot[q] = lo_slot;
    }
    out[q] = own[q] + phi * dot(grad, to_face);
  }
  bool keep_owner = !(out[0] > 0.0 && out[4] > 0.0);
  if (record) fb->owner_state = keep_owner;
  if (follow) keep_owner = fb->owner_state;
  return keep_owner ? own : out;
}

BoundaryModel::FaceBranch& BoundaryModel::face_branch(Index b) const {
  if (branch_.size() != static_cast<std::size_t>(s_.n_bfaces())) branch_.assign(static_cast<std::size_t>(s_.n_bfaces()), {});
  return branch_[static_cast<std::size_t>(b)];
}

// Pressure faces (inlets, and outlets with p > 0): from the reconstructed interior state (p_i, u_i,
// T_i) the face takes the patch's pressure (and, at inlets, temperature; outlets keep T_i), and the
// normal velocity moves by the acoustic characteristic, du = (p_i - p_b) / (rho_i c_s) with
// c_s = sqrt(gamma T_i). Inlets with tangential_velocity = "zero" keep only the normal component.
std::vector<Index> BoundaryModel::update_pressure_faces(const std::vector<double>& moments) {
  std::vector<Index> set;
  const Index n_int = s_.mesh.n_internal_faces();
  for (const PatchCondition& pc : patches_) {
    if (pc.kind != kBcPressure) continue;
    const bool inlet = pc.type == BoundaryType::PressureInlet || pc.type == BoundaryType::MassFlowInlet;
    for (Index b = pc.first_bface; b < pc.first_bface + pc.n_faces; ++b) {
      const std::array<double, 5> w = reconstruct_face_state(b, moments);
      const double p_i = w[0], T_i = w[4];
      const Vec3 u_i{w[1], w[2], w[3]};
      const double rho_i = p_i / T_i;
      const double c_s = std::sqrt(s_.gamma * T_i);
      const double p_b = pc.p, T_b = inlet ? pc.T : T_i;
      const double du = (p_i - p_b) / (rho_i * c_s);
      const Vec3& n = s_.geometry.face_normal[static_cast<std::size_t>(b + n_int)];
      const Vec3 u = pc.zero_tangential ? (dot(u_i, n) + du) * n : u_i + du * n;
      inflow[static_cast<std::size_t>(b)] = gsk_inflow_state{1.0, p_b / T_b, u.x, u.y, u.z, T_b};
      set.push_back(b);
    }
  }
  return set;
}

// Half-range sums of the unit-density wall Maxwellian over the incoming velocities (xi . n < 0) of
// Maxwell-wall face b: [0..5) the flux w M (xi . n) psi, [5..10) the moments w M psi, psi = (1, xi,
// |xi|^2 / 2); storage order.
std::array<double, 10> BoundaryModel::unit_wall_sums(Index b) const {
  std::array<double, 10> out{};
  const Vec3& n = s_.geometry.face_normal[static_cast<std::size_t>(b + s_.mesh.n_internal_faces())];
  const gsk_inflow_state& st = inflow[static_cast<std::size_t>(b)];
  const dev::Maxw<double> m = dev::maxw_from_state<double>(st.rho, st.ux, st.uy, st.uz, st.T);
  const VelocityLayout& v = s_.velocity;
  for (std::size_t p = 0; p < v.xi_x.size(); ++p) {
    const Vec3 xi{v.xi_x[p], v.xi_y[p], v.xi_z[p]};
    const double un = dot(xi, n);
    if (un >= 0.0) continue;
    double wm = v.weight * dev::maxw_eval(m, xi.x, xi.y, xi.z);
    if (!v.weight_rel.empty()) wm *= v.weight_rel[p];
    const double psi[5] = {1.0, xi.x, xi.y, xi.z, 0.5 * norm2(xi)};
    for (std::size_t k = 0; k < 5; ++k) {
      out[k] += wm * un * psi[k];
      out[5 + k] += wm * psi[k];
    }
  }
  return out;
}

// Diffuse-wall density: the incoming half-range Maxwellian carries back what leaves, rho = -j+ / j-
// (j+ the outgoing mass flux, j- that of the unit-density wall Maxwellian), never negative.
double BoundaryModel::diffuse_wall_density(double out_flux, double unit_in_flux) const {
  return unit_in_flux < 0.0 ? std::max(0.0, -out_flux / unit_in_flux) : 0.0;
}

namespace {
// Maxwell-wall density after a round: the diffuse part grows by the mass imbalance of the face,
// -(j+ + j-) / j-unit; a clearly negative result means the grid cannot carry the wall.
double maxwell_wall_density(double scale, double delta) {
  const double next = scale + delta;
  GSK_CHECK(next >= -1e-8, "Maxwell wall diffuse density became negative; refine or widen the velocity grid");
  return std::max(0.0, next);
}
}  // namespace

// Sums per face (20): [0..5) outgoing flux, [5..10) incoming flux of the current inflow state,
// [10..15) outgoing and [15..20) incoming moments (psi = 1, xi, |xi|^2 / 2), nondimensional.
void BoundaryModel::absorb(const std::vector<double>& sums) {
  const double L = s_.length_scale;
  const Index n_int = s_.mesh.n_internal_faces();
  std::fill(patch_flow.begin(), patch_flow.end(), 0.0);
  for (Index b = 0; b < s_.n_bfaces(); ++b) {
    const auto bi = static_cast<std::size_t>(b);
    const double* sm = &sums[bi * 20];
    const PatchCondition& pc = patch_of(b);
    gsk_inflow_state& st = inflow[bi];
    const double area = s_.geometry.face_area[static_cast<std::size_t>(b + n_int)] / (L * L);
    // incoming part: in_scale x the sums of the current state, plus (Maxwell walls) delta x the
    // unit-density sums
    double in_scale = 1.0, delta = 0.0;
    const std::array<double, 10> none{};
    const std::array<double, 10>& unit = pc.kind == kBcMaxwellWall ? wall_unit[bi] : none;
    if (pc.kind == kBcDiffuseWall) {
      const double rho = diffuse_wall_density(sm[0], sm[5]);
      const double w = s_.params.wall_relaxation;
      st.scale = w == 1.0 ? rho : w * rho + (1.0 - w) * st.scale;
      in_scale = rho;
    } else if (pc.kind == kBcVacuum) {
      in_scale = 0.0;
    } else if (pc.kind == kBcMaxwellWall) {
      GSK_CHECK(unit[0] < 0.0, "Maxwell wall has no incoming velocity points");
      delta = -(sm[0] + sm[5]) / unit[0];
      st.scale = maxwell_wall_density(st.scale, delta);
    }
    std::array<double, 5> mom;
    for (std::size_t k = 0; k < 5; ++k) {
      flux_out[bi][k] = area * sm[k];
      if (pc.kind == kBcMaxwellWall) {
        flux_in[bi][k] = area * (sm[5 + k] + delta * unit[k]);
        mom[k] = sm[10 + k] + sm[15 + k] + delta * unit[5 + k];
      } else {
        flux_in[bi][k] = area * in_scale * sm[5 + k];
        mom[k] = sm[10 + k] + in_scale * sm[15 + k];
      }
    }
    std::array<double, 5>& fs = face_state[bi];
    fs = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (mom[0] > 0.0) {
      const Vec3 u{mom[1] / mom[0], mom[2] / mom[0], mom[3] / mom[0]};
      fs = {mom[0], u.x, u.y, u.z, (2.0 * mom[4] / mom[0] - norm2(u)) / 3.0};
    }
    patch_flow[static_cast<std::size_t>(s_.bface_patch[bi])] += flux_out[bi][0] + flux_in[bi][0];
  }
  has_face_state_ = true;
}

void BoundaryModel::refresh_walls(const std::vector<double>& sums, const std::vector<int>& wall_faces) {
  for (const int b : wall_faces) {
    const auto bi = static_cast<std::size_t>(b);
    const double* sm = &sums[bi * 20];
    gsk_inflow_state& st = inflow[bi];
    if (patch_of(b).kind == kBcMaxwellWall) {
      st.scale = maxwell_wall_density(st.scale, -(sm[0] + sm[5]) / wall_unit[bi][0]);
      continue;
    }
    const double rho = diffuse_wall_density(sm[0], sm[5]);
    const double w = s_.params.wall_relaxation;
    st.scale = w == 1.0 ? rho : w * rho + (1.0 - w) * st.scale;
  }
}

}  // namespace gsk
