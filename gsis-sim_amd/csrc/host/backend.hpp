#pragma once
// The kinetic backend method set (the reference's class KineticBackend,
// src/kinetic/kinetic_backend.hpp:59-100), host-side and free of HIP types.

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/gsis_kinetic.h"

namespace gsk {

// The prepare of one CIS iteration on the device (Backend::control_prepare): the tables it needs,
// built once by the solver from the setup (FP64), and what one call takes and returns.
struct ControlTables {
  // per boundary face: owner cell, patch, area / L_ref^2, unit normal (3)
  std::vector<int> face_owner, face_patch;
  std::vector<double> face_area, face_normal;
  // pressure faces and their owners' reconstruction stencils: face index, owner, first slot and
  // slot count, K^3 V of the owner, face centre - owner centre (3)
  std::vector<int> press_face, press_owner, press_slot0, press_nslots;
  std::vector<double> press_eps2, press_to_face;
  // stencil slots in the owner's CSR order: neighbour cell (or -1) and boundary face (or -1),
  // whether the face is periodic, LSQ coefficients (3), face centre - owner centre (3)
  std::vector<int> slot_cell, slot_bface, slot_periodic;
  std::vector<double> slot_lsq, slot_offset;
  std::vector<double> volume;     // per cell
  std::vector<double> wall_unit;  // 10 per boundary face: unit-density half-range sums of Maxwell walls
  double gamma = 5.0 / 3.0, u_floor = 0.0, relaxation = 1.0;
  bool limiter = true, cell_reconstruction = false, closed_domain = false;
};

struct ControlPatch {
  int kind = 0, inlet = 0, zero_tangential = 0, first = 0, n = 0;
  double p = 0.0, T = 1.0, accommodation = 1.0;
};

struct ControlStep {
  // in: the host's boundary state (inflow: one per boundary face, face_state: 5 per face) and the
  // iteration's settings; out: the same state after the boundary rounds
  const ControlPatch* patches = nullptr;
  std::size_t n_patches = 0;
  gsk_inflow_state* inflow = nullptr;
  double* face_state = nullptr;
  std::size_t n_faces = 0;
  bool has_face_state = false, has_prev = false, seed_walls = false;
  long long rounds = 1;
  // in/out: the closed-domain target mass (set by the first prepare after initialisation)
  bool has_mass_target = false;
  double mass_target = 0.0;
  // out: fluxes F+ / F- (5 per face), patch mass flows, positivity, mass, closed-domain scale,
  // epsilon parts and volume means
  double* flux_out = nullptr;
  double* flux_in = nullptr;
  double* patch_flow = nullptr;
  long long bad = 0;
  int first_bad = -1;
  double bad_rho = 0.0, bad_T = 0.0;
  double mass = 0.0, scale = 1.0, eps[3] = {0.0, 0.0, 0.0}, q_mean[3] = {0.0, 0.0, 0.0};
  double sigma_mean[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  bool maxwell_negative = false, maxwell_no_incoming = false;
};

class Backend {
 public:
  virtual ~Backend() = default;
  virtual const char* name() const = 0;
  virtual int chunk_width() const = 0;
  virtual std::size_t memory_bytes() const = 0;
  virtual void set_equilibrium(const double* states, std::size_t n) = 0;
  virtual void set_distribution(const double* f, std::size_t n) = 0;
  virtual void get_distribution(double* f, std::size_t n) = 0;
  virtual void compute_moments(double* out, std::size_t n) = 0;
  virtual void scale_distribution(double s) = 0;
  virtual void update_cell_params() = 0;
  virtual void set_inflow(const gsk_inflow_state* inflow, std::size_t n) = 0;
  virtual void boundary_sums(double* out, std::size_t n) = 0;
  virtual void sweep(double* residual_cell, std::size_t n) = 0;
  // A sweep whose per-cell residual sums the caller does not read (the inner sweeps after the
  // first); a backend may then skip returning them. Default: sweep() into the scratch span.
  virtual void sweep_unreported(double* scratch, std::size_t n) { sweep(scratch, n); }
  virtual long long gsis_correction(const double* w_half, const double* w_new, std::size_t n_w, int* flags,
                                    std::size_t n_flags) = 0;
  virtual void internal_face_mass_flux(const int* cells, const int* entries, std::size_t n, double* out) = 0;
  // gsis_correction without any device allocation (no reference counterpart; for callers that cannot
  // allocate): the HIP backend stages w_half / w_new in a fixed per-device scratch array
  virtual long long gsis_correction_noalloc(const double* w_half, const double* w_new, std::size_t n_w, int* flags,
                                            std::size_t n_flags) {
    (void)w_half;
    (void)w_new;
    (void)n_w;
    (void)flags;
    (void)n_flags;
    throw std::runtime_error(std::string("backend \"") + name() + "\" has no non-allocating GSIS correction");
  }
  virtual gsk_timings take_timings() = 0;
  // boundary sums of at least the listed faces, written at their rows of out (20 x n_bfaces); the
  // other rows hold unspecified values. Default: all faces.
  virtual void boundary_sums_subset(const int* faces, std::size_t n_faces, double* out, std::size_t n) {
    (void)faces;
    (void)n_faces;
    boundary_sums(out, n);
  }
  virtual gsk_deq_counts deq_counts() = 0;
  // Discrete-equilibrium correction of boundary inflow states in place (the FP64 Newton of the
  // reference's KineticSolver::correct_inflow, kinetic_solver.cpp:103-114). Returns false when the
  // backend does not provide it; the solver then corrects the states on the host.
  virtual bool correct_inflow(gsk_inflow_state* states, std::size_t n) {
    (void)states;
    (void)n;
    return false;
  }
  // The wall part of the refresh between inner sweeps (KineticSolver::refresh_wall_inflow,
  // kinetic_solver.cpp:456-480) kept on the device: boundary sums of the listed wall faces, then
  // their inflow scales from those sums (maxwell[i] != 0: Maxwell wall with unit-density sum
  // unit0[i]; else diffuse wall with the given relaxation), starting from the scales of the last
  // set_inflow. Returns false when the backend does not provide it; the solver then refreshes on
  // the host. wall_inflow_scales brings the scales of the listed faces back and raises the error a
  // negative Maxwell-wall density would have raised during the refresh.
  virtual bool refresh_wall_inflow(const int* faces, const int* maxwell, const double* unit0, std::size_t n,
                                   double relaxation) {
    (void)faces;
    (void)maxwell;
    (void)unit0;
    (void)n;
    (void)relaxation;
    return false;
  }
  virtual void wall_inflow_scales(const int* faces, std::size_t n, double* out) {
    (void)faces;
    (void)n;
    (void)out;
  }

  // The prepare of a CIS iteration kept on the device (the reference's KineticSolver::prepare,
  // kinetic_solver.cpp:365-452, without its per-step host round trips): control_setup hands over the
  // tables once and returns false when the backend does not provide it. control_prepare then takes
  // the moments, the positivity check, mass and closed-domain rescaling, epsilon against the moments
  // kept by keep_previous_moments, the volume means, the cell parameters and `rounds` boundary rounds
  // (pressure-face states, their discrete correction, boundary sums, fluxes, face states, wall
  // densities, patch flows), and brings back only the boundary state and the step's scalars. The
  // moments stay on the device until download_moments; previous_moments brings back the kept ones,
  // set_previous_moments hands over moments the host computed.
  virtual bool control_setup(const ControlTables& t) {
    (void)t;
    return false;
  }
  virtual void control_prepare(ControlStep& step) { (void)step; }
  virtual void download_moments(double* out, std::size_t n) {
    (void)out;
    (void)n;
  }
  virtual void keep_previous_moments() {}
  virtual void previous_moments(double* out, std::size_t n) {
    (void)out;
    (void)n;
  }
  virtual void set_previous_moments(const double* m, std::size_t n) {
    (void)m;
    (void)n;
  }

  // Device-resident operators for sensitivity solves (no counterpart in the reference backend).
  // Arrays are device pointers to [cell][storage position] values of the backend's precision.
  virtual int element_size() const { return 0; }  // sizeof(f); 0 when the operators are absent
  virtual void residual_field_device(void* out) { unsupported(); }  // steady residual R(f), all velocities
  virtual void precondition_device(const void* r, void* out) { unsupported(); }  // LU-SGS passes on r, f untouched
  virtual void distribution_device(void* out) { unsupported(); }
  virtual void set_distribution_device(const void* in) { unsupported(); }
  virtual void set_mu0(double mu0) { unsupported(); }                               // viscosity scale
  // discrete-equilibrium Newton tolerances of cells and of inflow states (defaults 1e-11 / 1e-6 by
  // precision, and 1e-10)
  virtual void set_newton_tolerances(double cell_tol, double inflow_tol) { unsupported(); }
  virtual void set_specular_fraction(const double* frac, std::size_t n) { unsupported(); }  // 1 - alpha per face
  // limiter branches of every limited gradient: 0 off, 1 record at the next residual evaluation and
  // follow the record afterwards, 2 follow the last record
  virtual void set_limiter_branches(int mode) { unsupported(); }
  // Adjoint operators (FP64 form), at the current f with the limiter branches followed.
  // adjoint_prepare tabulates the cell parameters' Jacobians. adjoint_stage1 writes to out the
  // transposed residual applied to lam (device arrays; lam may be null for zero) and returns, per
  // boundary face, the adjoint of its inflow parameters (scale, A, B, ux, uy, uz) from the residual
  // (inflow_adj, 6 per face). adjoint_stage2 then adds to out the transposed boundary sums
  // (sums_adj, 20 per face), moments (moments_adj, 15 per cell, may be null) and what stage 1 left
  // in the backend (gradients, reflections, cell parameters).
  virtual void adjoint_prepare() { unsupported(); }
  virtual void adjoint_stage1(const void* lam, void* out, double* inflow_adj, std::size_t n) { unsupported(); }
  virtual void adjoint_stage2(const double* sums_adj, std::size_t n_sums, const double* moments_adj, std::size_t n_m,
                              void* out) {
    unsupported();
  }

 private:
  [[noreturn]] void unsupported() const {
    throw std::runtime_error(std::string("backend \"") + name() + "\" has no device-resident sensitivity operators");
  }
};

// HIP backend on the current device (FP64 or FP32 f per view.precision_fp64)
std::unique_ptr<Backend> make_hip_backend(const gsk_setup_view& view);

}  // namespace gsk
