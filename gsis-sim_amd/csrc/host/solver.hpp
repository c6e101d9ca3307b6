#pragma once
// CIS orchestration over a kinetic backend: moments -> closed-domain scaling -> epsilon ->
// cell params -> pressure-inflow / wall-density updates -> boundary sums -> sweeps.
// Restates class KineticSolver (src/kinetic/kinetic_solver.hpp:61-115,
// src/kinetic/kinetic_solver.cpp:27-722) on top of gsk::Backend.

#include <array>
#include <memory>
#include <vector>

#include "backend.hpp"
#include "boundary.hpp"
#include "setup.hpp"

namespace gsk {

// A section normal to an axis at position X (L_ref units): the internal faces between a cell whose
// centre lies before X and one whose centre does not (src/mesh/section_cut.hpp:20-55). A staircase on
// a general mesh, a plane on an aligned one; on a steady conservative solution the discrete flux
// through it equals the inlet / outlet flow. sign = +1 when the face's area vector points forward.
struct SectionCut {
  double position = 0.0;
  int axis = 0;
  std::vector<Index> faces;  // internal faces of the setup's (renumbered) mesh
  std::vector<signed char> sign;
  double area = 0.0, projected_area = 0.0;  // L_ref^2 units; projected = sum of sign * S_axis
  double face_position = 0.0;               // area-weighted face-centre coordinate (X when empty)
};

std::vector<SectionCut> build_section_cuts(const PolyMesh& mesh, const MeshGeometry& geometry, double length_scale,
                                           int axis, const std::vector<double>& positions);

class KineticSolver {
 public:
  // takes ownership of the backend; the setup must outlive the solver
  KineticSolver(const KineticSetup& setup, std::unique_ptr<Backend> backend);

  const KineticSetup& setup() const { return s_; }
  Backend& backend() { return *backend_; }

  void initialize_uniform(double rho, const Vec3& u, double T);
  void initialize_equilibrium(const double* states_original, std::size_t n);  // 5 per cell, original numbering
  const gsk_iteration_info& prepare();
  void update_mass_flow_pressure();
  const gsk_iteration_info& sweep(bool update_pressure);
  long long iteration() const { return iteration_; }
  const std::vector<double>& patch_mass_flow() const { return bnd_.patch_flow; }

  void moments(double* out, std::size_t n);  // original numbering
  void boundary_fluxes(double* fp, double* fm, double* fs, std::size_t nb);
  long long apply_gsis_correction(const double* w_half, const double* w_new, std::size_t n);
  void distribution(double* f, std::size_t n);
  void set_distribution(const double* f, std::size_t n);
  void residual_per_cell(double* out, std::size_t n) const;
  // nondimensional mass flow through each section, along +axis (KineticSolver::section_mass_flow,
  // kinetic_solver.cpp:669-686); prepares first
  void section_mass_flow(int axis, const std::vector<double>& positions, double* out);
  const std::vector<SectionCut>& section_cuts(int axis, const std::vector<double>& positions);

  // Device-resident operators for sensitivity solves: arrays are device pointers to
  // [cell][storage position] values in the f precision (internal numbering). Setting f keeps the
  // boundary state of the last prepare; the residual prepares for this f first (a function of f
  // alone once the prepare's inflow / boundary-sum rounds have converged, see
  // set_boundary_iterations).
  int element_size() const { return backend_->element_size(); }
  void distribution_device(void* out, std::size_t n);
  void set_distribution_device(const void* in, std::size_t n);
  void residual_device(void* out, std::size_t n);
  void precondition_device(const void* r, void* out, std::size_t n);
  double mu0() const { return mu0_; }
  void set_mu0(double mu0);
  void set_accommodation(int patch, double alpha);
  void set_newton_tolerances(double cell_tol, double inflow_tol);
  // inflow / boundary-sum rounds per prepare ([kinetic].boundary_iterations): the pressure-inlet
  // reconstruction reads the boundary face states of the previous round, so R(f) is a function of f
  // alone once these rounds have converged
  void set_boundary_iterations(long long n);
  // the prepare on the backend's device (the default where it offers one) or on the host
  void set_device_prepare(bool on) {
    use_device_ = on;
    ++version_;
  }
  bool prepared_on_device() const { return prepared_on_device_; }
  // Limiter branches for sensitivity solves: on = the next residual_device records the branch of
  // every limiter (the kinetic limited gradients and the pressure-face reconstruction) at the
  // current f, and later evaluations follow that record, so R is smooth around the recorded state;
  // off = the normal limiter.
  void freeze_limiters(bool on);
  // Adjoint operators (FP64 form; no counterpart in the reference). adjoint_prepare, at the current f
  // whose limiter branches freeze_limiters has recorded: tabulates what the transposed Jacobian of
  // R(f) needs there (Jacobians of the boundary model's face maps and of the cell parameters).
  // adjoint_apply_device: out = J^T lam + (d/df sum_k w_k patch_mass_flow_k)^T at that state, w one
  // weight per patch, lam and out device [cell][storage position] arrays (lam may be null: 0); valid
  // until the state changes (a new f, a residual evaluation, a parameter).
  void adjoint_prepare();
  void adjoint_apply_device(const void* lam, const double* patch_weights, std::size_t n_patches, void* out,
                            std::size_t n);

 private:
  void reset_after_initialization();
  // the prepare on the backend's device (Backend::control_prepare) when it offers one and no
  // limiter branches are being recorded or followed (those stay with the host's boundary model)
  void build_device_control();
  bool device_control() const { return control_ && use_device_ && freeze_ == Freeze::Off; }
  void prepare_on_device();
  // the moments of the last prepare on the host (the device prepare leaves them on the device)
  const std::vector<double>& host_moments();
  BoundaryKind patches_kind(int bface) const;
  // prepare, in stages
  void read_moments();
  void rescale_closed_domain();
  void iteration_statistics();
  void boundary_rounds();
  void measure_convergence();  // eq. (59) epsilon against the previous iteration's moments
  void refresh_wall_inflow();
  void correct_pressure_states(const std::vector<Index>& faces);

  const KineticSetup& s_;
  std::unique_ptr<Backend> backend_;
  BoundaryModel bnd_;
  long long iteration_ = 0, version_ = 0, prepared_version_ = -1;
  double mu0_ = 0.0;
  double inflow_tol_ = 1e-10;
  long long boundary_iterations_ = 1;
  enum class Freeze { Off, Record, Follow } freeze_ = Freeze::Off;
  struct AdjointTables;  // adjoint.cpp
  std::shared_ptr<AdjointTables> adj_;
  long long adj_version_ = -1;
  std::vector<double> specular_;  // 1 - alpha per boundary face (Maxwell walls)
  gsk_iteration_info info_{};
  std::vector<double> moments_, prev_moments_;  // 15 per cell, internal numbering
  bool has_prev_ = false;
  bool control_ = false;            // the backend can run the prepare on its device
  bool use_device_ = true;          // and is to
  bool prepared_on_device_ = false;  // the last prepare ran there
  bool moments_on_host_ = true;      // moments_ holds the last prepare's moments
  bool prev_on_device_ = false;      // the previous iteration's moments are the backend's, not prev_moments_
  std::vector<double> bsums_;
  std::vector<double> residual_cell_, inner_residual_;
  double mass_target_ = 0.0;
  bool has_mass_target_ = false;
  // wall faces and, per wall face, whether it is a Maxwell wall and its unit-density incoming mass
  // flux (the inner-sweep wall refresh)
  std::vector<int> wall_faces_, wall_maxwell_;
  std::vector<double> wall_unit0_;
  bool wall_scales_on_device_ = false;  // the wall scales in bnd_.inflow are behind the backend's
  // section cuts of the last requested (axis, positions) and, per cut face, its owner and CSR entry
  int section_axis_ = -1;
  std::vector<double> section_positions_;
  std::vector<SectionCut> section_cuts_;
  std::vector<int> section_cells_, section_entries_;
  void build_sections(int axis, const std::vector<double>& positions);
};

}  // namespace gsk
