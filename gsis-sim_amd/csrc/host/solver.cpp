// CIS orchestration over a kinetic backend, host FP64 (the reference's KineticSolver,
// src/kinetic/kinetic_solver.cpp:27-722; the boundary state lives in boundary.cpp).
//
// One CIS iteration is prepare() then sweep():
//   prepare  moments of f -> positivity check -> closed-domain mass rescaling -> eq. (59) epsilon
//            -> cell parameters (Shakhov target, discrete equilibrium) -> boundary rounds: pressure
//            face states -> inflow table -> boundary sums -> fluxes, face states, wall densities
//   sweep    inner_sweeps LU-SGS sweeps (walls refreshed between them), the residual norm of the
//            first one, and the moments kept for the next epsilon.

#include "solver.hpp"

#include <algorithm>
#include <chrono>

namespace gsk {

namespace {
double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

KineticSolver::KineticSolver(const KineticSetup& setup, std::unique_ptr<Backend> backend)
    : s_(setup), backend_(std::move(backend)), bnd_(setup) {
  mu0_ = setup.mu0;
  boundary_iterations_ = setup.params.boundary_iterations;
  const auto nc = static_cast<std::size_t>(s_.n_cells());
  moments_.assign(nc * 15, 0.0);
  prev_moments_.assign(nc * 15, 0.0);
  bsums_.assign(static_cast<std::size_t>(s_.n_bfaces()) * 20, 0.0);
  residual_cell_.assign(nc, 0.0);
  bnd_.reset(inflow_tol_);
  build_device_control();
}

// The tables of the device prepare: per boundary face its owner, patch, area and normal; per pressure
// face the stencil of its owner's reconstruction (BoundaryModel::reconstruct_face_state: the owner's
// CSR faces in order, with their LSQ coefficients and face-centre offsets); volumes; the Maxwell
// walls' unit-density sums.
void KineticSolver::build_device_control() {
  ControlTables t;
  const Index n_int = s_.mesh.n_internal_faces(), nb = s_.n_bfaces();
  const double L = s_.length_scale;
  for (Index b = 0; b < nb; ++b) {
    const auto f = static_cast<std::size_t>(b + n_int);
    t.face_owner.push_back(static_cast<int>(s_.mesh.owner[f]));
    t.face_patch.push_back(static_cast<int>(s_.bface_patch[static_cast<std::size_t>(b)]));
    t.face_area.push_back(s_.geometry.face_area[f] / (L * L));
    const Vec3& n = s_.geometry.face_normal[f];
    t.face_normal.insert(t.face_normal.end(), {n.x, n.y, n.z});
  }
  for (const PatchCondition& pc : bnd_.patches()) {
    if (pc.kind != kBcPressure) continue;
    for (Index b = pc.first_bface; b < pc.first_bface + pc.n_faces; ++b) {
      const auto face = static_cast<std::size_t>(b + n_int);
      const Index owner = s_.mesh.owner[face];
      const Vec3& x_own = s_.geometry.cell_center[static_cast<std::size_t>(owner)];
      const Offset k0 = s_.csr.begin(owner), k1 = s_.csr.end(owner);
      t.press_face.push_back(static_cast<int>(b));
      t.press_owner.push_back(static_cast<int>(owner));
      t.press_slot0.push_back(static_cast<int>(t.slot_cell.size()));
      t.press_nslots.push_back(static_cast<int>(k1 - k0));
      const double K = s_.params.limiter_k;
      t.press_eps2.push_back(K * K * K * s_.volume[static_cast<std::size_t>(owner)]);
      const Vec3 to_face = s_.geometry.face_center[face] - x_own;
      t.press_to_face.insert(t.press_to_face.end(), {to_face.x, to_face.y, to_face.z});
      for (Offset k = k0; k < k1; ++k) {
        const auto e = static_cast<std::size_t>(k);
        const Index other = s_.csr.other[e];
        int bface = -1, periodic = 0;
        if (other == kInvalid) {
          bface = static_cast<int>(s_.csr.faces[e] - n_int);
          periodic = patches_kind(bface) == kBcPeriodic ? 1 : 0;
        }
        t.slot_cell.push_back(other == kInvalid ? -1 : static_cast<int>(other));
        t.slot_bface.push_back(bface);
        t.slot_periodic.push_back(periodic);
        const Vec3& l = s_.lsq[e];
        t.slot_lsq.insert(t.slot_lsq.end(), {l.x, l.y, l.z});
        const Vec3 off = s_.geometry.face_center[static_cast<std::size_t>(s_.csr.faces[e])] - x_own;
        t.slot_offset.insert(t.slot_offset.end(), {off.x, off.y, off.z});
      }
    }
  }
  t.volume = s_.volume;
  if (!bnd_.wall_unit.empty()) {
    t.wall_unit.resize(static_cast<std::size_t>(nb) * 10);
    for (std::size_t b = 0; b < static_cast<std::size_t>(nb); ++b)
      for (std::size_t k = 0; k < 10; ++k) t.wall_unit[b * 10 + k] = bnd_.wall_unit[b][k];
  }
  t.gamma = s_.gamma;
  t.u_floor = s_.params.convergence_u_floor;
  t.relaxation = s_.params.wall_relaxation;
  t.limiter = s_.limiter_mode != 0;
  t.cell_reconstruction = s_.params.pressure_boundary_reconstruction == "cell";
  t.closed_domain = s_.closed_domain;
  control_ = backend_->control_setup(t);
}

BoundaryKind KineticSolver::patches_kind(int bface) const {
  return bnd_.patches()[static_cast<std::size_t>(s_.bface_patch[static_cast<std::size_t>(bface)])].kind;
}

const std::vector<double>& KineticSolver::host_moments() {
  if (!moments_on_host_) {
    backend_->download_moments(moments_.data(), moments_.size());
    moments_on_host_ = true;
  }
  return moments_;
}

// prepare on the device: the same stages, one exchange of the boundary state and the step's scalars
void KineticSolver::prepare_on_device() {
  static_assert(sizeof(std::array<double, 5>) == 5 * sizeof(double), "face arrays must be flat");
  std::vector<ControlPatch> patches;
  for (const PatchCondition& pc : bnd_.patches()) {
    ControlPatch c;
    c.kind = pc.kind;
    c.inlet = pc.type == BoundaryType::PressureInlet || pc.type == BoundaryType::MassFlowInlet;
    c.zero_tangential = pc.zero_tangential ? 1 : 0;
    c.first = static_cast<int>(pc.first_bface);
    c.n = static_cast<int>(pc.n_faces);
    c.p = pc.p;
    c.T = pc.T;
    c.accommodation = pc.accommodation;
    patches.push_back(c);
  }
  if (has_prev_ && !prev_on_device_) backend_->set_previous_moments(prev_moments_.data(), prev_moments_.size());
  ControlStep st;
  st.patches = patches.data();
  st.n_patches = patches.size();
  st.inflow = bnd_.inflow.data();
  st.face_state = bnd_.face_state.empty() ? nullptr : bnd_.face_state[0].data();
  st.flux_out = bnd_.flux_out.empty() ? nullptr : bnd_.flux_out[0].data();
  st.flux_in = bnd_.flux_in.empty() ? nullptr : bnd_.flux_in[0].data();
  st.n_faces = bnd_.inflow.size();
  st.patch_flow = bnd_.patch_flow.data();
  st.has_face_state = bnd_.has_face_states();
  st.seed_walls = !bnd_.has_face_states();
  st.has_prev = has_prev_;
  st.rounds = boundary_iterations_;
  st.has_mass_target = has_mass_target_;
  st.mass_target = mass_target_;
  backend_->control_prepare(st);
  moments_on_host_ = false;
  prepared_on_device_ = true;
  if (st.bad > 0)
    throw Error(strf("iteration %lld: %lld cells have non-positive density or temperature (e.g. original cell %d, rho = "
                     "%.4g, T = %.4g)",
                     iteration_, st.bad, s_.cell_new_to_old[static_cast<std::size_t>(st.first_bad)], st.bad_rho,
                     st.bad_T));
  info_.total_mass = st.mass;
  if (s_.closed_domain) {
    has_mass_target_ = st.has_mass_target;
    mass_target_ = st.mass_target;
    info_.mass_scale = st.scale;
  }
  if (has_prev_) {
    info_.eps_rho = st.eps[0];
    info_.eps_u = st.eps[1];
    info_.eps_T = st.eps[2];
    info_.epsilon = std::max({info_.eps_rho, info_.eps_u, info_.eps_T});
    info_.has_epsilon = 1;
  }
  for (int k = 0; k < 3; ++k) info_.q_mean[k] = st.q_mean[k];
  for (int k = 0; k < 6; ++k) info_.sigma_mean[k] = st.sigma_mean[k];
  if (s_.discrete_eq) {
    const gsk_deq_counts dc = backend_->deq_counts();
    info_.deq_corrected = dc.corrected;
    info_.deq_failed = dc.failed;
    info_.deq_max_iterations = dc.max_iterations;
  }
  GSK_CHECK(!st.maxwell_no_incoming, "Maxwell wall has no incoming velocity points");
  GSK_CHECK(!st.maxwell_negative, "Maxwell wall diffuse density became negative; refine or widen the velocity grid");
  bnd_.mark_face_states();
}

void KineticSolver::reset_after_initialization() {
  ++version_;
  iteration_ = 0;
  has_prev_ = false;
  has_mass_target_ = false;
  bnd_.reset(inflow_tol_);
}

// Moments of the current f; every cell must keep a positive, finite density and temperature.
void KineticSolver::read_moments() {
  backend_->compute_moments(moments_.data(), moments_.size());
  moments_on_host_ = true;
  Index bad = 0, first = -1;
  for (Index c = 0; c < s_.n_cells(); ++c) {
    const double* m = &moments_[static_cast<std::size_t>(c) * 15];
    const bool ok = m[0] > 0.0 && m[4] > 0.0 && std::isfinite(m[0]) && std::isfinite(m[4]);
    if (ok) continue;
    if (bad++ == 0) first = c;
  }
  if (bad == 0) return;
  const double* m = &moments_[static_cast<std::size_t>(first) * 15];
  throw Error(strf("iteration %lld: %d cells have non-positive density or temperature (e.g. original cell %d, rho = "
                   "%.4g, T = %.4g)",
                   iteration_, bad, s_.cell_new_to_old[static_cast<std::size_t>(first)], m[0], m[4]));
}

// A closed domain keeps the mass of its first prepare: f and the density-proportional moments
// (rho and p, stress, heat flux) are scaled by target / current mass.
void KineticSolver::rescale_closed_domain() {
  double mass = 0.0;
  for (Index c = 0; c < s_.n_cells(); ++c)
    mass += moments_[static_cast<std::size_t>(c) * 15] * s_.volume[static_cast<std::size_t>(c)];
  info_.total_mass = mass;
  if (!s_.closed_domain) return;
  if (!has_mass_target_) {
    mass_target_ = mass;
    has_mass_target_ = true;
  }
  const double scale = mass_target_ / mass;
  info_.mass_scale = scale;
  if (scale == 1.0) return;
  backend_->scale_distribution(scale);
  for (std::size_t c = 0; c < static_cast<std::size_t>(s_.n_cells()); ++c) {
    double* m = &moments_[c * 15];
    m[0] *= scale;
    for (int k = 5; k < 15; ++k) m[k] *= scale;
  }
}

// eq. (59): epsilon = max over (rho, u, T) of the volume-weighted relative L2 change since the last
// iteration; the velocity change is measured against max(|u|, convergence_u_floor).
void KineticSolver::measure_convergence() {
  if (prev_on_device_) {
    backend_->previous_moments(prev_moments_.data(), prev_moments_.size());
    prev_on_device_ = false;
  }
  double change[3] = {0, 0, 0}, size[3] = {0, 0, 0}, vol = 0.0;
  for (Index c = 0; c < s_.n_cells(); ++c) {
    const double V = s_.volume[static_cast<std::size_t>(c)];
    const double* now = &moments_[static_cast<std::size_t>(c) * 15];
    const double* was = &prev_moments_[static_cast<std::size_t>(c) * 15];
    const double dr = now[0] - was[0], dT = now[4] - was[4];
    const double du[3] = {now[1] - was[1], now[2] - was[2], now[3] - was[3]};
    change[0] += dr * dr * V;
    size[0] += was[0] * was[0] * V;
    change[1] += (du[0] * du[0] + du[1] * du[1] + du[2] * du[2]) * V;
    size[1] += (was[1] * was[1] + was[2] * was[2] + was[3] * was[3]) * V;
    change[2] += dT * dT * V;
    size[2] += was[4] * was[4] * V;
    vol += V;
  }
  const double floor = s_.params.convergence_u_floor;
  size[1] = std::max(size[1], floor * floor * vol);
  auto rel = [](double d, double n) { return n > 0.0 ? std::sqrt(d / n) : (d > 0.0 ? 1.0 : 0.0); };
  info_.eps_rho = rel(change[0], size[0]);
  info_.eps_u = rel(change[1], size[1]);
  info_.eps_T = rel(change[2], size[2]);
  info_.epsilon = std::max({info_.eps_rho, info_.eps_u, info_.eps_T});
  info_.has_epsilon = 1;
}

// Volume means of the heat flux and the stress (reported per iteration).
void KineticSolver::iteration_statistics() {
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

// Discrete-equilibrium correction of new pressure-face states: on the backend's device when it
// offers it, else per face on the host cores.
void KineticSolver::correct_pressure_states(const std::vector<Index>& faces) {
  if (!s_.discrete_eq || faces.empty()) return;
  std::vector<gsk_inflow_state> batch(faces.size());
  for (std::size_t i = 0; i < faces.size(); ++i) batch[i] = bnd_.inflow[static_cast<std::size_t>(faces[i])];
  if (!backend_->correct_inflow(batch.data(), batch.size())) {
#pragma omp parallel for schedule(dynamic, 4)
    for (std::size_t i = 0; i < batch.size(); ++i) discrete_inflow_state(s_, inflow_tol_, batch[i]);
  }
  for (std::size_t i = 0; i < faces.size(); ++i) bnd_.inflow[static_cast<std::size_t>(faces[i])] = batch[i];
}

// boundary_iterations rounds: pressure-face states from the moments (and the previous round's face
// states), then the boundary sums with those states, then what the sums imply (wall densities,
// fluxes, face states); the backend always holds the latest inflow table.
void KineticSolver::boundary_rounds() {
  if (!bnd_.has_face_states()) bnd_.seed_wall_densities(moments_);
  for (long long round = 0; round < boundary_iterations_; ++round) {
    correct_pressure_states(bnd_.update_pressure_faces(moments_));
    backend_->set_inflow(bnd_.inflow.data(), bnd_.inflow.size());
    backend_->boundary_sums(bsums_.data(), bsums_.size());
    bnd_.absorb(bsums_);
    backend_->set_inflow(bnd_.inflow.data(), bnd_.inflow.size());
  }
}

const gsk_iteration_info& KineticSolver::prepare() {  // (:365-452)
  if (prepared_version_ == version_) return info_;
  const double t0 = now_seconds();
  info_ = gsk_iteration_info{};
  info_.iteration = iteration_;
  info_.mass_scale = 1.0;
  if (device_control()) {
    prepare_on_device();
  } else {
    prepared_on_device_ = false;
    read_moments();
    rescale_closed_domain();
    if (has_prev_) measure_convergence();
    iteration_statistics();
    backend_->update_cell_params();
    if (s_.discrete_eq) {
      const gsk_deq_counts dc = backend_->deq_counts();
      info_.deq_corrected = dc.corrected;
      info_.deq_failed = dc.failed;
      info_.deq_max_iterations = dc.max_iterations;
    }
    boundary_rounds();
  }
  prepared_version_ = version_;
  info_.seconds = now_seconds() - t0;
  return info_;
}

// Between inner sweeps only the wall densities follow f (kinetic_solver.cpp:456-480): on the device
// when the backend keeps them there, else from the wall faces' boundary sums on the host.
void KineticSolver::refresh_wall_inflow() {
  if (wall_faces_.empty()) {
    for (Index b = 0; b < s_.n_bfaces(); ++b) {
      const BoundaryKind k = bnd_.patch_of(b).kind;
      if (k != kBcDiffuseWall && k != kBcMaxwellWall) continue;
      const bool maxwell = k == kBcMaxwellWall;
      wall_faces_.push_back(b);
      wall_maxwell_.push_back(maxwell ? 1 : 0);
      wall_unit0_.push_back(maxwell ? bnd_.wall_unit[static_cast<std::size_t>(b)][0] : 0.0);
    }
  }
  if (wall_faces_.empty()) return;
  if (backend_->refresh_wall_inflow(wall_faces_.data(), wall_maxwell_.data(), wall_unit0_.data(), wall_faces_.size(),
                                    s_.params.wall_relaxation)) {
    wall_scales_on_device_ = true;
    return;
  }
  backend_->boundary_sums_subset(wall_faces_.data(), wall_faces_.size(), bsums_.data(), bsums_.size());
  bnd_.refresh_walls(bsums_, wall_faces_);
  backend_->set_inflow(bnd_.inflow.data(), bnd_.inflow.size());
}

// Mass-flow inlets: the patch pressure moves towards the target flow by a relaxed Newton step on
// the free-molecular flow sensitivity D = sum_faces A (1 / sqrt(2 pi T_in) + 1 / (2 sqrt(gamma T_c))),
// limited to 10% of p per iteration (kinetic_solver.cpp:482-501).
void KineticSolver::update_mass_flow_pressure() {
  prepare();
  bool changed = false;
  const double L2 = s_.length_scale * s_.length_scale;
  std::vector<PatchCondition>& patches = bnd_.patches();
  for (std::size_t pi = 0; pi < patches.size(); ++pi) {
    PatchCondition& pc = patches[pi];
    if (pc.type != BoundaryType::MassFlowInlet) continue;
    double D = 0.0;
    for (Index b = pc.first_bface; b < pc.first_bface + pc.n_faces; ++b) {
      const auto f = static_cast<std::size_t>(b + s_.mesh.n_internal_faces());
      const double T_cell = host_moments()[static_cast<std::size_t>(s_.mesh.owner[f]) * 15 + 4];
      D += s_.geometry.face_area[f] / L2 * (1.0 / std::sqrt(2.0 * kPi * pc.T) + 0.5 / std::sqrt(s_.gamma * T_cell));
    }
    GSK_CHECK(D > 0.0, "mass-flow inlet has zero area");
    const double step = 0.2 * (pc.target_mass_flow + bnd_.patch_flow[pi]) / D;
    pc.p = std::max(1e-12, pc.p + std::clamp(step, -0.1 * pc.p, 0.1 * pc.p));
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
  for (long long inner = 1; inner < s_.params.inner_sweeps; ++inner) {
    if (s_.params.inner_wall_update) refresh_wall_inflow();
    inner_residual_.resize(residual_cell_.size());
    backend_->sweep_unreported(inner_residual_.data(), inner_residual_.size());
  }
  if (wall_scales_on_device_) {  // the host copy of the wall densities catches up once per iteration
    wall_scales_on_device_ = false;
    std::vector<double> scales(wall_faces_.size());
    backend_->wall_inflow_scales(wall_faces_.data(), wall_faces_.size(), scales.data());
    for (std::size_t i = 0; i < wall_faces_.size(); ++i)
      bnd_.inflow[static_cast<std::size_t>(wall_faces_[i])].scale = scales[i];
  }
  // residual norm of the first inner sweep: sqrt(w sum_c (sum R^2)_c / V_c / sum_c V_c)
  double sum = 0.0, vol = 0.0;
  for (Index c = 0; c < s_.n_cells(); ++c) {
    const double V = s_.volume[static_cast<std::size_t>(c)];
    sum += residual_cell_[static_cast<std::size_t>(c)] / V;
    vol += V;
  }
  info_.residual = std::sqrt(sum * s_.velocity.weight / vol);
  info_.has_residual = 1;
  if (prepared_on_device_) {
    backend_->keep_previous_moments();
    prev_on_device_ = true;
  } else {
    prev_moments_ = moments_;
    prev_on_device_ = false;
  }
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
  const std::vector<double>& m = host_moments();
  for (std::size_t c = 0; c < nc; ++c) {
    const auto o = static_cast<std::size_t>(s_.cell_new_to_old[c]);
    for (std::size_t k = 0; k < 15; ++k) out[o * 15 + k] = m[c * 15 + k];
  }
}

void KineticSolver::boundary_fluxes(double* fp, double* fm, double* fs, std::size_t nb) {
  prepare();
  GSK_CHECK(nb == static_cast<std::size_t>(s_.n_bfaces()), "boundary flux output length differs from the face count");
  for (std::size_t b = 0; b < nb; ++b) {
    const auto o = static_cast<std::size_t>(s_.bface_new_to_old[b]);
    for (std::size_t k = 0; k < 5; ++k) {
      fp[o * 5 + k] = bnd_.flux_out[b][k];
      fm[o * 5 + k] = bnd_.flux_in[b][k];
      fs[o * 5 + k] = bnd_.face_state[b][k];
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
  if (freeze_ == Freeze::Record) {  // the backend follows its record from now on, and so does the host
    freeze_ = Freeze::Follow;
    bnd_.set_branches(BoundaryModel::Branches::Follow);
  }
}

void KineticSolver::freeze_limiters(bool on) {
  backend_->set_limiter_branches(on ? 1 : 0);
  freeze_ = on ? Freeze::Record : Freeze::Off;
  bnd_.set_branches(on ? BoundaryModel::Branches::Record : BoundaryModel::Branches::Off);
  ++version_;
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
  GSK_CHECK(patch >= 0 && patch < static_cast<int>(bnd_.patches().size()), "patch index out of range");
  PatchCondition& pc = bnd_.patches()[static_cast<std::size_t>(patch)];
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
