// CIS orchestration (restates src/kinetic/kinetic_solver.cpp:27-722; line references below are
// to that file). All host arithmetic is FP64.

#include "solver.hpp"

#include <algorithm>
#include <chrono>

#include "../kernels/kinetic_math.hpp"

namespace gsk {

namespace {

double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Sequential host form of the discrete conservative Maxwellian in FP64 (discrete_maxwell_seq<double>,
// kinetic_kernels.hpp:499-540), used for boundary inflow states. Returns iterations / 0 / -1.
int discrete_maxwell_host(const KineticSetup& s, double tol, double rho, double ux, double uy, double uz, double T,
                          dev::Maxw<double>& out) {
  dev::DeqGrid g{};
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = s.deq_lo[a];
    g.hi[a] = s.deq_hi[a];
    g.d[a] = s.deq_d[a];
  }
  g.criterion = s.deq_criterion;
  g.max_iterations = s.deq_max_iterations;
  g.tolerance = tol;
  out = dev::maxw_from_state<double>(rho, ux, uy, uz, T);
  if (dev::deq_resolved(g, ux, uy, uz, T)) return 0;
  double beta[5];
  dev::deq_initial_beta(rho, T, beta);
  double rel_prev = 1e300;
  const int nv = s.nv();
  for (int it = 0; it < g.max_iterations; ++it) {
    const dev::Maxw<double> m = dev::deq_params<double>(beta, ux, uy, uz);
    double sums[dev::kDeqSums] = {};
    double t[dev::kDeqSums];
    for (int p = 0; p < nv; ++p) {
      const double x = s.velocity.xi_x[static_cast<std::size_t>(p)], y = s.velocity.xi_y[static_cast<std::size_t>(p)],
                   z = s.velocity.xi_z[static_cast<std::size_t>(p)];
      double e = dev::maxw_eval(m, x, y, z);
      if (!s.velocity.weight_rel.empty()) e *= s.velocity.weight_rel[static_cast<std::size_t>(p)];  // stretched grid
      dev::deq_terms<double>(e, 0.0, x - ux, y - uy, z - uz, t);
      for (int k = 0; k < dev::kDeqSums; ++k) sums[k] += t[k];
    }
    double b[5] = {beta[0], beta[1], beta[2], beta[3], beta[4]};
    double rel = 0.0;
    const int r = dev::deq_newton_step(sums, s.velocity.weight, rho, T, g.tolerance, b, rel);
    if (dev::deq_accept(r, rel, rel_prev, it + 1 == g.max_iterations)) {
      out = m;
      return it + 1;
    }
    if (r < 0) break;
    for (int k = 0; k < 5; ++k) beta[k] = b[k];
    rel_prev = rel;
  }
  out = dev::maxw_from_state<double>(rho, ux, uy, uz, T);
  return -1;
}

}  // namespace

KineticSolver::KineticSolver(const KineticSetup& setup, std::unique_ptr<Backend> backend)
    : s_(setup), patches_(setup.patches), backend_(std::move(backend)) {
  mu0_ = setup.mu0;
  boundary_iterations_ = setup.params.boundary_iterations;
  const auto nc = static_cast<std::size_t>(s_.n_cells());
  const auto nb = static_cast<std::size_t>(s_.n_bfaces());
  moments_.assign(nc * 15, 0.0);
  prev_moments_.assign(nc * 15, 0.0);
  bsums_.assign(nb * 20, 0.0);
  inflow_.assign(nb, gsk_inflow_state{0, 0, 0, 0, 0, 1.0});
  F_plus_.assign(nb, {});
  F_minus_.assign(nb, {});
  face_state_.assign(nb, {});
  residual_cell_.assign(nc, 0.0);
  init_static_inflow();
}

// discrete-equilibrium inflow: replace (rho, u, T) by the Maxwellian whose discrete moments match (:103-114)
void KineticSolver::correct_inflow(gsk_inflow_state& in) const {
  if (!s_.discrete_eq || !(in.rho > 0.0) || !(in.T > 0.0)) return;
  dev::DeqGrid g{};
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = s_.deq_lo[a];
    g.hi[a] = s_.deq_hi[a];
    g.d[a] = s_.deq_d[a];
  }
  g.criterion = s_.deq_criterion;
  if (dev::deq_resolved(g, in.ux, in.uy, in.uz, in.T)) return;
  dev::Maxw<double> m;
  const int r = discrete_maxwell_host(s_, inflow_tol_, in.rho, in.ux, in.uy, in.uz, in.T, m);
  if (r > 0) {
    // maxwell_params_to_state (kinetic_kernels.hpp:543-552)
    in.T = 1.0 / (2.0 * m.B);
    const double tpt = 2.0 * kPi * in.T;
    in.rho = m.A * tpt * std::sqrt(tpt);
    in.ux = m.ux;
    in.uy = m.uy;
    in.uz = m.uz;
  }
}

void KineticSolver::init_static_inflow() {  // (:117-142)
  wall_cache_patch_ = -1;
  for (Index b = 0; b < s_.n_bfaces(); ++b) {
    const PatchCondition& pc = patches_[static_cast<std::size_t>(s_.bface_patch[static_cast<std::size_t>(b)])];
    gsk_inflow_state& in = inflow_[static_cast<std::size_t>(b)];
    if (pc.kind == kBcDiffuseWall || pc.kind == kBcMaxwellWall) {
      in.rho = 1.0;
      in.ux = pc.u_wall.x;
      in.uy = pc.u_wall.y;
      in.uz = pc.u_wall.z;
      in.T = pc.T;
      in.scale = 1.0;
      // every face of a wall patch has the same unit-density state: correct it once per patch
      const int pi = s_.bface_patch[static_cast<std::size_t>(b)];
      if (wall_cache_patch_ != pi) {
        wall_cache_ = in;
        correct_inflow(wall_cache_);
        wall_cache_patch_ = pi;
      }
      in = wall_cache_;
    } else if (pc.kind == kBcPressure) {
      in.scale = 1.0;
      in.rho = pc.p / pc.T;
      in.T = pc.T;
    } else {
      in = gsk_inflow_state{0, 0, 0, 0, 0, 1.0};
    }
  }
  if (s_.n_maxwell_faces > 0) {
    wall_unit_moments_.assign(static_cast<std::size_t>(s_.n_bfaces()), {});
    for (Index b = 0; b < s_.n_bfaces(); ++b)
      if (patches_[static_cast<std::size_t>(s_.bface_patch[static_cast<std::size_t>(b)])].kind == kBcMaxwellWall)
        wall_unit_moments_[static_cast<std::size_t>(b)] = compute_wall_unit_sums(b);
  }
}

void KineticSolver::reset_after_initialization() {
  ++version_;
  iteration_ = 0;
  has_prev_ = false;
  has_face_state_ = false;
  has_mass_target_ = false;
  init_static_inflow();
}

void KineticSolver::init_wall_scale_from_cells() {  // (:156-164)
  const Index n_int = s_.mesh.n_internal_faces();
  for (Index b = 0; b < s_.n_bfaces(); ++b) {
    const PatchCondition& pc = patches_[static_cast<std::size_t>(s_.bface_patch[static_cast<std::size_t>(b)])];
    if (pc.kind != kBcDiffuseWall && pc.kind != kBcMaxwellWall) continue;
    const Index c = s_.mesh.owner[static_cast<std::size_t>(b + n_int)];
    inflow_[static_cast<std::size_t>(b)].scale =
        pc.accommodation * cell_value(moments_, c, 0) * std::sqrt(cell_value(moments_, c, 4) / pc.T);
  }
}

void KineticSolver::update_pressure_inflow() {  // (:172-260)
  const Index n_int = s_.mesh.n_internal_faces();
  const double K = s_.params.limiter_k;
  const bool limit = s_.limiter_mode != 0;
  std::vector<Index> pending;
  for (const PatchCondition& pc : patches_) {
    if (pc.kind != kBcPressure) continue;
    for (Index b = pc.first_bface; b < pc.first_bface + pc.n_faces; ++b) {
      const auto f = static_cast<std::size_t>(b + n_int);
      const Index ci = s_.mesh.owner[f];
      auto state_of = [&](Index c) {
        return std::array<double, 5>{cell_value(moments_, c, 5), cell_value(moments_, c, 1), cell_value(moments_, c, 2),
                                     cell_value(moments_, c, 3), cell_value(moments_, c, 4)};
      };
      const std::array<double, 5> wi = state_of(ci);
      std::array<double, 5> wmax = wi, wmin = wi;
      std::array<Vec3, 5> grad{};
      for (Offset k = s_.csr.begin(ci); k < s_.csr.end(ci); ++k) {
        const auto ki = static_cast<std::size_t>(k);
        std::array<double, 5> wk;
        if (s_.csr.other[ki] != kInvalid) {
          wk = state_of(s_.csr.other[ki]);
        } else {
          const auto bk = static_cast<std::size_t>(s_.csr.faces[ki] - n_int);
          const int kind = patches_[static_cast<std::size_t>(s_.bface_patch[bk])].kind;
          const auto& fs = face_state_[bk];
          if (has_face_state_ && kind != kBcPeriodic && fs[0] > 0.0 && fs[4] > 0.0)
            wk = {fs[0] * fs[4], fs[1], fs[2], fs[3], fs[4]};
          else
            wk = wi;
        }
        for (int q = 0; q < 5; ++q) {
          const auto qi = static_cast<std::size_t>(q);
          const double d = wk[qi] - wi[qi];
          grad[qi] += d * s_.lsq[ki];
          wmax[qi] = std::max(wmax[qi], wk[qi]);
          wmin[qi] = std::min(wmin[qi], wk[qi]);
        }
      }
      const Vec3 dxb = s_.geometry.face_center[f] - s_.geometry.cell_center[static_cast<std::size_t>(ci)];
      const double eps2 = K * K * K * s_.volume[static_cast<std::size_t>(ci)];
      std::array<double, 5> wib{};
      for (int q = 0; q < 5; ++q) {
        const auto qi = static_cast<std::size_t>(q);
        double phi = 1.0;
        if (limit) {
          const double dmx = wmax[qi] - wi[qi], dmn = wmin[qi] - wi[qi];
          for (Offset k = s_.csr.begin(ci); k < s_.csr.end(ci); ++k) {
            const auto ki = static_cast<std::size_t>(k);
            const Index fk = s_.csr.faces[ki];
            if (s_.csr.other[ki] == kInvalid &&
                patches_[static_cast<std::size_t>(s_.bface_patch[static_cast<std::size_t>(fk - n_int)])].kind == kBcPeriodic)
              continue;
            const double d2 = dot(grad[qi], s_.geometry.face_center[static_cast<std::size_t>(fk)] -
                                                s_.geometry.cell_center[static_cast<std::size_t>(ci)]);
            phi = std::min(phi, dev::venkat_face<double>(d2, dmx, dmn, eps2));
          }
        }
        wib[qi] = wi[qi] + phi * dot(grad[qi], dxb);
      }
      if (!(wib[0] > 0.0 && wib[4] > 0.0)) wib = wi;
      if (s_.params.pressure_boundary_reconstruction == "cell") wib = wi;
      const double p_ib = wib[0];
      const Vec3 u_ib{wib[1], wib[2], wib[3]};
      const double T_ib = wib[4];
      const double rho_ib = p_ib / T_ib;
      const double cs = std::sqrt(s_.gamma * T_ib);
      const Vec3& n = s_.geometry.face_normal[f];
      double p_b, T_b;
      if (pc.type == BoundaryType::PressureInlet || pc.type == BoundaryType::MassFlowInlet) {
        p_b = pc.p;
        T_b = pc.T;
      } else {
        p_b = pc.p;
        T_b = T_ib;
      }
      const double du = (p_ib - p_b) / (rho_ib * cs);
      gsk_inflow_state& in = inflow_[static_cast<std::size_t>(b)];
      in.scale = 1.0;
      in.rho = p_b / T_b;
      const Vec3 u = pc.zero_tangential ? (dot(u_ib, n) + du) * n : u_ib + du * n;
      in.ux = u.x;
      in.uy = u.y;
      in.uz = u.z;
      in.T = T_b;
      pending.push_back(b);
    }
  }
  // discrete-equilibrium correction of the new inflow states: independent per face, on the
  // backend's device when it provides it, else on the host cores
  if (!s_.discrete_eq || pending.empty()) return;
  std::vector<gsk_inflow_state> batch(pending.size());
  for (std::size_t i = 0; i < pending.size(); ++i) batch[i] = inflow_[static_cast<std::size_t>(pending[i])];
  if (!backend_->correct_inflow(batch.data(), batch.size())) {
#pragma omp parallel for schedule(dynamic, 4)
    for (std::size_t i = 0; i < batch.size(); ++i) correct_inflow(batch[i]);
  }
  for (std::size_t i = 0; i < pending.size(); ++i) inflow_[static_cast<std::size_t>(pending[i])] = batch[i];
}

std::array<double, 10> KineticSolver::compute_wall_unit_sums(Index b) const {  // (:263-277)
  std::array<double, 10> out{};
  const Vec3& n = s_.geometry.face_normal[static_cast<std::size_t>(b + s_.mesh.n_internal_faces())];
  const gsk_inflow_state& in = inflow_[static_cast<std::size_t>(b)];
  const dev::Maxw<double> m = dev::maxw_from_state<double>(in.rho, in.ux, in.uy, in.uz, in.T);
  for (int p = 0; p < s_.nv(); ++p) {
    const Vec3 xi{s_.velocity.xi_x[static_cast<std::size_t>(p)], s_.velocity.xi_y[static_cast<std::size_t>(p)],
                  s_.velocity.xi_z[static_cast<std::size_t>(p)]};
    const double un = dot(xi, n);
    if (un >= 0.0) continue;
    double wf = s_.velocity.weight * dev::maxw_eval(m, xi.x, xi.y, xi.z);
    if (!s_.velocity.weight_rel.empty()) wf *= s_.velocity.weight_rel[static_cast<std::size_t>(p)];  // stretched grid
    const double psi[5] = {1.0, xi.x, xi.y, xi.z, 0.5 * norm2(xi)};
    for (int k = 0; k < 5; ++k) {
      out[static_cast<std::size_t>(k)] += wf * un * psi[k];
      out[static_cast<std::size_t>(5 + k)] += wf * psi[k];
    }
  }
  return out;
}

void KineticSolver::process_boundary_sums() {  // (:280-335)
  const Index nb = s_.n_bfaces();
  const double L = s_.length_scale;
  const Index n_int = s_.mesh.n_internal_faces();
  patch_mass_flow_.assign(patches_.size(), 0.0);
  for (Index b = 0; b < nb; ++b) {
    const auto bi = static_cast<std::size_t>(b);
    const double* sm = &bsums_[bi * 20];
    const int pi = s_.bface_patch[bi];
    const PatchCondition& pc = patches_[static_cast<std::size_t>(pi)];
    const double area = s_.geometry.face_area[static_cast<std::size_t>(b + n_int)] / (L * L);
    double in_scale = 1.0;
    if (pc.kind == kBcDiffuseWall) {
      const double jp = sm[0], jm = sm[5];
      const double rho_g = jm < 0.0 ? std::max(0.0, -jp / jm) : 0.0;
      const double w = s_.params.wall_relaxation;
      inflow_[bi].scale = w == 1.0 ? rho_g : w * rho_g + (1.0 - w) * inflow_[bi].scale;
      in_scale = rho_g;
    } else if (pc.kind == kBcVacuum) {
      in_scale = 0.0;
    }
    std::array<double, 10> unit{};
    double delta = 0.0;
    if (pc.kind == kBcMaxwellWall) {
      unit = wall_unit_moments_[bi];
      GSK_CHECK(unit[0] < 0.0, "Maxwell wall has no incoming velocity points");
      delta = -(sm[0] + sm[5]) / unit[0];
      const double next = inflow_[bi].scale + delta;
      GSK_CHECK(next >= -1e-8, "Maxwell wall diffuse density became negative; refine or widen the velocity grid");
      inflow_[bi].scale = std::max(0.0, next);
    }
    std::array<double, 5> mom{};
    for (int k = 0; k < 5; ++k) {
      const auto ks = static_cast<std::size_t>(k);
      F_plus_[bi][ks] = area * sm[k];
      if (pc.kind == kBcMaxwellWall) {
        F_minus_[bi][ks] = area * (sm[5 + k] + delta * unit[ks]);
        mom[ks] = sm[10 + k] + sm[15 + k] + delta * unit[5 + ks];
      } else {
        F_minus_[bi][ks] = area * in_scale * sm[5 + k];
        mom[ks] = sm[10 + k] + in_scale * sm[15 + k];
      }
    }
    auto& fs = face_state_[bi];
    if (mom[0] > 0.0) {
      const Vec3 u{mom[1] / mom[0], mom[2] / mom[0], mom[3] / mom[0]};
      fs = {mom[0], u.x, u.y, u.z, (2.0 * mom[4] / mom[0] - norm2(u)) / 3.0};
    } else {
      fs = {0.0, 0.0, 0.0, 0.0, 0.0};
    }
    patch_mass_flow_[static_cast<std::size_t>(pi)] += F_plus_[bi][0] + F_minus_[bi][0];
  }
  has_face_state_ = true;
}

void KineticSolver::compute_epsilon() {  // eq. (59) (:338-363)
  double num[3] = {0, 0, 0}, den[3] = {0, 0, 0}, vol = 0.0;
  for (Index c = 0; c < s_.n_cells(); ++c) {
    const double V = s_.volume[static_cast<std::size_t>(c)];
    const double* a = &moments_[static_cast<std::size_t>(c) * 15];
    const double* b = &prev_moments_[static_cast<std::size_t>(c) * 15];
    num[0] += (a[0] - b[0]) * (a[0] - b[0]) * V;
    den[0] += b[0] * b[0] * V;
    const double du2 = (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]) + (a[3] - b[3]) * (a[3] - b[3]);
    num[1] += du2 * V;
    den[1] += (b[1] * b[1] + b[2] * b[2] + b[3] * b[3]) * V;
    num[2] += (a[4] - b[4]) * (a[4] - b[4]) * V;
    den[2] += b[4] * b[4] * V;
    vol += V;
  }
  const double floor = s_.params.convergence_u_floor;
  den[1] = std::max(den[1], floor * floor * vol);
  auto ratio = [](double n, double d) { return d > 0.0 ? std::sqrt(n / d) : (n > 0.0 ? 1.0 : 0.0); };
  info_.eps_rho = ratio(num[0], den[0]);
  info_.eps_u = ratio(num[1], den[1]);
  info_.eps_T = ratio(num[2], den[2]);
  info_.epsilon = std::max({info_.eps_rho, info_.eps_u, info_.eps_T});
  info_.has_epsilon = 1;
}

const gsk_iteration_info& KineticSolver::prepare() {  // (:365-452)
  if (prepared_version_ == version_) return info_;
  const double t0 = now_seconds();
  info_ = gsk_iteration_info{};
  info_.iteration = iteration_;
  info_.mass_scale = 1.0;
  backend_->compute_moments(moments_.data(), moments_.size());
  Index bad = 0, first_bad = -1;
  for (Index c = 0; c < s_.n_cells(); ++c) {
    const double rho = cell_value(moments_, c, 0), T = cell_value(moments_, c, 4);
    if (!(rho > 0.0) || !(T > 0.0) || !std::isfinite(rho) || !std::isfinite(T)) {
      if (bad == 0) first_bad = c;
      ++bad;
    }
  }
  GSK_CHECK(bad == 0, strf("iteration %lld: %d cells have non-positive density or temperature (e.g. original cell %d, "
                           "rho = %.4g, T = %.4g)",
                           iteration_, bad, bad ? s_.cell_new_to_old[static_cast<std::size_t>(first_bad)] : -1,
                           bad ? cell_value(moments_, first_bad, 0) : 0.0, bad ? cell_value(moments_, first_bad, 4) : 0.0));
  double mass = 0.0;
  for (Index c = 0; c < s_.n_cells(); ++c) mass += cell_value(moments_, c, 0) * s_.volume[static_cast<std::size_t>(c)];
  info_.total_mass = mass;
  if (s_.closed_domain) {
    if (!has_mass_target_) {
      mass_target_ = mass;
      has_mass_target_ = true;
    }
    const double scale = mass_target_ / mass;
    if (scale != 1.0) {
      backend_->scale_distribution(scale);
      for (Index c = 0; c < s_.n_cells(); ++c) {
        double* m = &moments_[static_cast<std::size_t>(c) * 15];
        m[0] *= scale;
        for (int k = 5; k < 15; ++k) m[k] *= scale;
      }
    }
    info_.mass_scale = scale;
  }
  if (has_prev_) compute_epsilon();
  {
    double vol = 0.0;
    for (Index c = 0; c < s_.n_cells(); ++c) {
      const double V = s_.volume[static_cast<std::size_t>(c)];
      const double* m = &moments_[static_cast<std::size_t>(c) * 15];
      for (int k = 0; k < 3; ++k) info_.q_mean[k] += V * m[12 + k];
      for (int k = 0; k < 6; ++k) info_.sigma_mean[k] += V * m[6 + k];
      vol += V;
    }
    for (double& x : info_.q_mean) x /= vol;
    for (double& x : info_.sigma_mean) x /= vol;
  }
  backend_->update_cell_params();
  if (s_.discrete_eq) {
    const gsk_deq_counts dc = backend_->deq_counts();
    info_.deq_corrected = dc.corrected;
    info_.deq_failed = dc.failed;
    info_.deq_max_iterations = dc.max_iterations;
  }
  if (!has_face_state_) init_wall_scale_from_cells();
  update_pressure_inflow();
  backend_->set_inflow(inflow_.data(), inflow_.size());
  backend_->boundary_sums(bsums_.data(), bsums_.size());
  process_boundary_sums();
  backend_->set_inflow(inflow_.data(), inflow_.size());
  for (long long b = 1; b < boundary_iterations_; ++b) {
    update_pressure_inflow();
    backend_->set_inflow(inflow_.data(), inflow_.size());
    backend_->boundary_sums(bsums_.data(), bsums_.size());
    process_boundary_sums();
    backend_->set_inflow(inflow_.data(), inflow_.size());
  }
  prepared_version_ = version_;
  info_.seconds = now_seconds() - t0;
  return info_;
}

void KineticSolver::refresh_wall_inflow() {  // (:456-480)
  // only the wall faces' sums are read here
  if (wall_faces_.empty()) {
    for (Index b = 0; b < s_.n_bfaces(); ++b) {
      const int k = patches_[static_cast<std::size_t>(s_.bface_patch[static_cast<std::size_t>(b)])].kind;
      if (k == kBcDiffuseWall || k == kBcMaxwellWall) wall_faces_.push_back(b);
    }
  }
  if (wall_faces_.empty()) return;
  // on the device when the backend keeps the scales there (no host round trip per inner sweep)
  wall_maxwell_.resize(wall_faces_.size());
  wall_unit0_.resize(wall_faces_.size());
  for (std::size_t i = 0; i < wall_faces_.size(); ++i) {
    const auto bi = static_cast<std::size_t>(wall_faces_[i]);
    const bool mw = patches_[static_cast<std::size_t>(s_.bface_patch[bi])].kind == kBcMaxwellWall;
    wall_maxwell_[i] = mw ? 1 : 0;
    wall_unit0_[i] = mw ? wall_unit_moments_[bi][0] : 0.0;
  }
  if (backend_->refresh_wall_inflow(wall_faces_.data(), wall_maxwell_.data(), wall_unit0_.data(), wall_faces_.size(),
                                    s_.params.wall_relaxation)) {
    wall_scales_on_device_ = true;
    return;
  }
  backend_->boundary_sums_subset(wall_faces_.data(), wall_faces_.size(), bsums_.data(), bsums_.size());
  bool any = false;
  for (Index b = 0; b < s_.n_bfaces(); ++b) {
    const auto bi = static_cast<std::size_t>(b);
    const PatchCondition& pc = patches_[static_cast<std::size_t>(s_.bface_patch[bi])];
    if (pc.kind != kBcDiffuseWall && pc.kind != kBcMaxwellWall) continue;
    const double* sm = &bsums_[bi * 20];
    if (pc.kind == kBcMaxwellWall) {
      const auto& unit = wall_unit_moments_[bi];
      const double next = inflow_[bi].scale - (sm[0] + sm[5]) / unit[0];
      GSK_CHECK(next >= -1e-8, "Maxwell wall diffuse density became negative; refine the velocity grid");
      inflow_[bi].scale = std::max(0.0, next);
      any = true;
      continue;
    }
    const double jp = sm[0], jm = sm[5];
    const double rho_g = jm < 0.0 ? std::max(0.0, -jp / jm) : 0.0;
    const double w = s_.params.wall_relaxation;
    inflow_[bi].scale = w == 1.0 ? rho_g : w * rho_g + (1.0 - w) * inflow_[bi].scale;
    any = true;
  }
  if (any) backend_->set_inflow(inflow_.data(), inflow_.size());
}

void KineticSolver::update_mass_flow_pressure() {  // (:482-501)
  prepare();
  bool changed = false;
  for (std::size_t pi = 0; pi < patches_.size(); ++pi) {
    PatchCondition& pc = patches_[pi];
    if (pc.type != BoundaryType::MassFlowInlet) continue;
    double D = 0.0;
    for (Index b = pc.first_bface; b < pc.first_bface + pc.n_faces; ++b) {
      const auto f = static_cast<std::size_t>(b + s_.mesh.n_internal_faces());
      const double T = cell_value(moments_, s_.mesh.owner[f], 4);
      D += s_.geometry.face_area[f] / (s_.length_scale * s_.length_scale) *
           (1.0 / std::sqrt(2.0 * kPi * pc.T) + 0.5 / std::sqrt(s_.gamma * T));
    }
    GSK_CHECK(D > 0.0, "mass-flow inlet has zero area");
    const double dp = 0.2 * (pc.target_mass_flow + patch_mass_flow_[pi]) / D;
    pc.p = std::max(1e-12, pc.p + std::clamp(dp, -0.1 * pc.p, 0.1 * pc.p));
    changed = true;
  }
  if (changed) {
    ++version_;
    prepare();
  }
}

const gsk_iteration_info& KineticSolver::sweep(bool update_pressure) {  // (:503-531)
  prepare();
  if (update_pressure) update_mass_flow_pressure();
  const double t0 = now_seconds();
  backend_->sweep(residual_cell_.data(), residual_cell_.size());
  for (long long m = 1; m < s_.params.inner_sweeps; ++m) {
    if (s_.params.inner_wall_update) refresh_wall_inflow();
    inner_residual_.resize(residual_cell_.size());
    backend_->sweep_unreported(inner_residual_.data(), inner_residual_.size());
  }
  if (wall_scales_on_device_) {  // the host copy of the wall scales catches up once per iteration
    wall_scales_on_device_ = false;
    std::vector<double> sc(wall_faces_.size());
    backend_->wall_inflow_scales(wall_faces_.data(), wall_faces_.size(), sc.data());
    for (std::size_t i = 0; i < wall_faces_.size(); ++i) inflow_[static_cast<std::size_t>(wall_faces_[i])].scale = sc[i];
  }
  double sum = 0.0, vol = 0.0;
  for (Index c = 0; c < s_.n_cells(); ++c) {
    const double V = s_.volume[static_cast<std::size_t>(c)];
    sum += residual_cell_[static_cast<std::size_t>(c)] / V;
    vol += V;
  }
  info_.residual = std::sqrt(sum * s_.velocity.weight / vol);
  info_.has_residual = 1;
  prev_moments_ = moments_;
  has_prev_ = true;
  ++version_;
  ++iteration_;
  info_.timings = backend_->take_timings();  // waits for the backend's queued work
  info_.seconds += now_seconds() - t0;
  return info_;
}

void KineticSolver::initialize_uniform(double rho, const Vec3& u, double T) {
  GSK_CHECK(rho > 0.0 && T > 0.0, "initial density and temperature must be positive");
  const auto nc = static_cast<std::size_t>(s_.n_cells());
  std::vector<double> st(nc * 5);
  for (std::size_t c = 0; c < nc; ++c) {
    st[c * 5 + 0] = rho;
    st[c * 5 + 1] = u.x;
    st[c * 5 + 2] = u.y;
    st[c * 5 + 3] = u.z;
    st[c * 5 + 4] = T;
  }
  backend_->set_equilibrium(st.data(), st.size());
  reset_after_initialization();
}

void KineticSolver::initialize_equilibrium(const double* states, std::size_t n) {
  const auto nc = static_cast<std::size_t>(s_.n_cells());
  GSK_CHECK(n == nc * 5, "initial states length differs from 5 x cells");
  std::vector<double> st(nc * 5);
  for (std::size_t c = 0; c < nc; ++c) {
    const auto o = static_cast<std::size_t>(s_.cell_new_to_old[c]);
    GSK_CHECK(states[o * 5] > 0.0 && states[o * 5 + 4] > 0.0, strf("cell %zu: initial density/temperature not positive", o));
    for (std::size_t k = 0; k < 5; ++k) st[c * 5 + k] = states[o * 5 + k];
  }
  backend_->set_equilibrium(st.data(), st.size());
  reset_after_initialization();
}

void KineticSolver::moments(double* out, std::size_t n) {
  prepare();
  const auto nc = static_cast<std::size_t>(s_.n_cells());
  GSK_CHECK(n == nc * 15, "moments output length differs from 15 x cells");
  for (std::size_t c = 0; c < nc; ++c) {
    const auto o = static_cast<std::size_t>(s_.cell_new_to_old[c]);
    for (std::size_t k = 0; k < 15; ++k) out[o * 15 + k] = moments_[c * 15 + k];
  }
}

void KineticSolver::boundary_fluxes(double* fp, double* fm, double* fs, std::size_t nb) {
  prepare();
  GSK_CHECK(nb == static_cast<std::size_t>(s_.n_bfaces()), "boundary flux output length differs from the face count");
  for (std::size_t b = 0; b < nb; ++b) {
    const auto o = static_cast<std::size_t>(s_.bface_new_to_old[b]);
    for (std::size_t k = 0; k < 5; ++k) {
      fp[o * 5 + k] = F_plus_[b][k];
      fm[o * 5 + k] = F_minus_[b][k];
      fs[o * 5 + k] = face_state_[b][k];
    }
  }
}

long long KineticSolver::apply_gsis_correction(const double* w_half, const double* w_new, std::size_t n) {
  const auto nc = static_cast<std::size_t>(s_.n_cells());
  GSK_CHECK(n == nc * 5, "GSIS correction states length differs from 5 x cells");
  std::vector<double> wh(nc * 5), wn(nc * 5);
  for (std::size_t c = 0; c < nc; ++c) {
    const auto o = static_cast<std::size_t>(s_.cell_new_to_old[c]);
    GSK_CHECK(w_new[o * 5] > 0.0 && w_new[o * 5 + 4] > 0.0 && w_half[o * 5] > 0.0 && w_half[o * 5 + 4] > 0.0,
              strf("cell %zu: correction density/temperature not positive", o));
    for (std::size_t k = 0; k < 5; ++k) {
      wh[c * 5 + k] = w_half[o * 5 + k];
      wn[c * 5 + k] = w_new[o * 5 + k];
    }
  }
  std::vector<int> flags(nc, 0);
  const long long replaced = backend_->gsis_correction(wh.data(), wn.data(), wh.size(), flags.data(), flags.size());
  ++version_;
  return replaced;
}

void KineticSolver::distribution(double* f, std::size_t n) {
  const auto nc = static_cast<std::size_t>(s_.n_cells()), nv = static_cast<std::size_t>(s_.nv());
  GSK_CHECK(n == nc * nv, "distribution length differs from cells x velocities");
  std::vector<double> internal(nc * nv);
  backend_->get_distribution(internal.data(), internal.size());
  for (std::size_t c = 0; c < nc; ++c) {
    const auto o = static_cast<std::size_t>(s_.cell_new_to_old[c]);
    for (std::size_t p = 0; p < nv; ++p)
      f[o * nv + static_cast<std::size_t>(s_.velocity.storage_to_grid[p])] = internal[c * nv + p];
  }
}

void KineticSolver::set_distribution(const double* f, std::size_t n) {
  const auto nc = static_cast<std::size_t>(s_.n_cells()), nv = static_cast<std::size_t>(s_.nv());
  GSK_CHECK(n == nc * nv, "distribution length differs from cells x velocities");
  std::vector<double> internal(nc * nv);
  for (std::size_t c = 0; c < nc; ++c) {
    const auto o = static_cast<std::size_t>(s_.cell_new_to_old[c]);
    for (std::size_t p = 0; p < nv; ++p)
      internal[c * nv + p] = f[o * nv + static_cast<std::size_t>(s_.velocity.storage_to_grid[p])];
  }
  backend_->set_distribution(internal.data(), internal.size());
  reset_after_initialization();
}

void KineticSolver::distribution_device(void* out, std::size_t n) {
  GSK_CHECK(n == static_cast<std::size_t>(s_.n_cells()) * static_cast<std::size_t>(s_.nv()),
            "distribution length differs from cells x velocities");
  backend_->distribution_device(out);
}

void KineticSolver::set_distribution_device(const void* in, std::size_t n) {
  GSK_CHECK(n == static_cast<std::size_t>(s_.n_cells()) * static_cast<std::size_t>(s_.nv()),
            "distribution length differs from cells x velocities");
  backend_->set_distribution_device(in);
  ++version_;
}

void KineticSolver::residual_device(void* out, std::size_t n) {
  GSK_CHECK(n == static_cast<std::size_t>(s_.n_cells()) * static_cast<std::size_t>(s_.nv()),
            "residual length differs from cells x velocities");
  ++version_;
  prepare();
  backend_->residual_field_device(out);
}

void KineticSolver::precondition_device(const void* r, void* out, std::size_t n) {
  GSK_CHECK(n == static_cast<std::size_t>(s_.n_cells()) * static_cast<std::size_t>(s_.nv()),
            "vector length differs from cells x velocities");
  prepare();
  backend_->precondition_device(r, out);
}

void KineticSolver::set_mu0(double mu0) {
  GSK_CHECK(mu0 > 0.0, "viscosity scale must be positive");
  backend_->set_mu0(mu0);
  mu0_ = mu0;
  ++version_;
}

void KineticSolver::set_accommodation(int patch, double alpha) {
  GSK_CHECK(patch >= 0 && patch < static_cast<int>(patches_.size()), "patch index out of range");
  PatchCondition& pc = patches_[static_cast<std::size_t>(patch)];
  GSK_CHECK(pc.kind == kBcMaxwellWall, "accommodation can be changed on Maxwell walls only");
  GSK_CHECK(alpha > 0.0, "accommodation must be positive");
  pc.accommodation = alpha;
  if (specular_.empty()) {
    specular_.assign(static_cast<std::size_t>(s_.n_bfaces()), 0.0);
    if (!s_.maxwell_specular.empty()) specular_ = s_.maxwell_specular;
  }
  for (Index b = pc.first_bface; b < pc.first_bface + pc.n_faces; ++b) specular_[static_cast<std::size_t>(b)] = 1.0 - alpha;
  backend_->set_specular_fraction(specular_.data(), specular_.size());
  ++version_;
}

void KineticSolver::set_boundary_iterations(long long n) {
  GSK_CHECK(n >= 1 && n <= 100000, "boundary iterations must be 1..100000");
  boundary_iterations_ = n;
  ++version_;
}

void KineticSolver::set_newton_tolerances(double cell_tol, double inflow_tol) {
  GSK_CHECK(cell_tol >= 0.0 && inflow_tol >= 0.0, "Newton tolerances must be >= 0");
  backend_->set_newton_tolerances(cell_tol, inflow_tol);
  inflow_tol_ = inflow_tol;
  ++version_;
}

std::vector<SectionCut> build_section_cuts(const PolyMesh& mesh, const MeshGeometry& geometry, double length_scale,
                                           int axis, const std::vector<double>& positions) {  // (section_cut.cpp:48-85)
  GSK_CHECK(axis >= 0 && axis <= 2, "section axis must be 0 (x), 1 (y) or 2 (z)");
  GSK_CHECK(length_scale > 0.0, "section length scale must be positive");
  const double L = length_scale;
  auto coord = [axis](const Vec3& v) { return axis == 0 ? v.x : axis == 1 ? v.y : v.z; };
  std::vector<SectionCut> cuts;
  cuts.reserve(positions.size());
  for (const double X : positions) {
    GSK_CHECK(std::isfinite(X), "section position is not a finite number");
    SectionCut cut;
    cut.position = X;
    cut.axis = axis;
    double weighted = 0.0;
    for (Index f = 0; f < mesh.n_internal_faces(); ++f) {
      const auto fi = static_cast<std::size_t>(f);
      const bool o_before = coord(geometry.cell_center[static_cast<std::size_t>(mesh.owner[fi])]) / L < X;
      const bool n_before = coord(geometry.cell_center[static_cast<std::size_t>(mesh.neighbour[fi])]) / L < X;
      if (o_before == n_before) continue;
      const signed char sg = o_before ? 1 : -1;
      const double A = geometry.face_area[fi] / (L * L);
      cut.faces.push_back(f);
      cut.sign.push_back(sg);
      cut.area += A;
      cut.projected_area += static_cast<double>(sg) * coord(geometry.face_area_vector[fi]) / (L * L);
      weighted += A * coord(geometry.face_center[fi]) / L;
    }
    cut.face_position = cut.area > 0.0 ? weighted / cut.area : X;
    cuts.push_back(std::move(cut));
  }
  return cuts;
}

void KineticSolver::build_sections(int axis, const std::vector<double>& positions) {  // (kinetic_solver.cpp:56-77)
  if (axis == section_axis_ && positions == section_positions_) return;
  section_cuts_ = build_section_cuts(s_.mesh, s_.geometry, s_.length_scale, axis, positions);
  section_cells_.clear();
  section_entries_.clear();
  for (const SectionCut& cut : section_cuts_) {
    for (const Index f : cut.faces) {
      const Index o = s_.mesh.owner[static_cast<std::size_t>(f)];
      Offset e = -1;
      for (Offset k = s_.csr.begin(o); k < s_.csr.end(o); ++k)
        if (s_.csr.faces[static_cast<std::size_t>(k)] == f) {
          e = k;
          break;
        }
      GSK_CHECK(e >= 0 && s_.csr.sign[static_cast<std::size_t>(e)] > 0,
                strf("section face %d is not among the faces of its owner %d", static_cast<int>(f), static_cast<int>(o)));
      section_cells_.push_back(static_cast<int>(o));
      section_entries_.push_back(static_cast<int>(e));
    }
  }
  section_axis_ = axis;
  section_positions_ = positions;
}

const std::vector<SectionCut>& KineticSolver::section_cuts(int axis, const std::vector<double>& positions) {
  build_sections(axis, positions);
  return section_cuts_;
}

void KineticSolver::section_mass_flow(int axis, const std::vector<double>& positions, double* out) {
  if (positions.empty()) return;
  prepare();
  build_sections(axis, positions);
  std::vector<double> face(section_cells_.size(), 0.0);
  if (!face.empty())
    backend_->internal_face_mass_flux(section_cells_.data(), section_entries_.data(), face.size(), face.data());
  std::size_t k = 0;
  for (std::size_t c = 0; c < section_cuts_.size(); ++c) {
    double sum = 0.0;
    for (std::size_t j = 0; j < section_cuts_[c].faces.size(); ++j, ++k)
      sum += static_cast<double>(section_cuts_[c].sign[j]) * face[k];
    out[c] = sum;
  }
}

void KineticSolver::residual_per_cell(double* out, std::size_t n) const {
  const auto nc = static_cast<std::size_t>(s_.n_cells());
  GSK_CHECK(n == nc, "residual output length differs from the cell count");
  for (std::size_t c = 0; c < nc; ++c) {
    const auto o = static_cast<std::size_t>(s_.cell_new_to_old[c]);
    out[o] = std::sqrt(s_.velocity.weight * residual_cell_[c]) / s_.volume[c];
  }
}

}  // namespace gsk
