// extern "C" entry points declared in include/gsis_kinetic.h. Every call catches exceptions and
// turns them into a status code plus a thread-local message (the reference throws gsis::Error,
// src/core/error.hpp:16-47; a C++ host rethrows the message).

#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "../../include/gsis_kinetic.h"
#include "host/backend.hpp"
#include "host/setup.hpp"
#include "host/solver.hpp"

struct gsk_setup {
  gsk::KineticSetup s;
};
struct gsk_backend {
  std::unique_ptr<gsk::Backend> owner;  // empty when the handle aliases a solver's backend
  gsk::Backend* b = nullptr;
};
struct gsk_solver {
  std::unique_ptr<gsk::KineticSolver> sv;
  gsk_backend handle;  // non-owning alias of the solver's backend
};

namespace {
thread_local std::string g_last_error;

template <class Fn>
int guarded(Fn&& fn) {
  try {
    fn();
    return GSK_OK;
  } catch (const gsk::Error& e) {
    g_last_error = e.what();
    const std::string m = e.what();
    if (m.find("HIP error") != std::string::npos) return GSK_ERR_DEVICE;
    if (m.find("device allocation") != std::string::npos || m.find("device memory") != std::string::npos)
      return GSK_ERR_MEMORY;
    return GSK_ERR_INVALID;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return GSK_ERR_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return GSK_ERR_INVALID;
  }
}

void need(const void* p, const char* what) {
  if (p == nullptr) throw gsk::Error(std::string("null argument: ") + what);
}
}  // namespace

extern "C" {

const char* gsk_last_error(void) { return g_last_error.c_str(); }
const char* gsk_version(void) { return "gsis-sim_amd 0.1 (gfx950)"; }

int gsk_setup_from_case(const char* toml_path, const char* overrides, gsk_setup** out) {
  return guarded([&] {
    need(toml_path, "toml_path");
    need(out, "out");
    *out = nullptr;
    auto h = std::make_unique<gsk_setup>();
    h->s = gsk::setup_from_case(toml_path, overrides ? overrides : "");
    *out = h.release();
  });
}

int gsk_setup_destroy(gsk_setup* s) {
  delete s;
  return GSK_OK;
}

int gsk_setup_view_get(const gsk_setup* s, gsk_setup_view* out) {
  return guarded([&] {
    need(s, "setup");
    need(out, "out");
    *out = s->s.view();
  });
}

int gsk_setup_sizes(const gsk_setup* s, int32_t* n_cells, int32_t* nv, int32_t* n_bfaces, int32_t* n_blocks,
                    int32_t* n_patches) {
  return guarded([&] {
    need(s, "setup");
    if (n_cells) *n_cells = s->s.n_cells();
    if (nv) *nv = s->s.nv();
    if (n_bfaces) *n_bfaces = s->s.n_bfaces();
    if (n_blocks) *n_blocks = s->s.n_blocks();
    if (n_patches) *n_patches = static_cast<int32_t>(s->s.patches.size());
  });
}

int gsk_setup_cell_new_to_old(const gsk_setup* s, int32_t* out) {
  return guarded([&] {
    need(s, "setup");
    need(out, "out");
    std::memcpy(out, s->s.cell_new_to_old.data(), s->s.cell_new_to_old.size() * sizeof(int32_t));
  });
}

int gsk_setup_storage_to_grid(const gsk_setup* s, int32_t* out) {
  return guarded([&] {
    need(s, "setup");
    need(out, "out");
    for (std::size_t p = 0; p < s->s.velocity.storage_to_grid.size(); ++p) out[p] = s->s.velocity.storage_to_grid[p];
  });
}

int gsk_setup_cell_centers(const gsk_setup* s, double* out) {
  return guarded([&] {
    need(s, "setup");
    need(out, "out");
    for (std::size_t c = 0; c < s->s.geometry.cell_center.size(); ++c) {
      out[3 * c + 0] = s->s.geometry.cell_center[c].x;
      out[3 * c + 1] = s->s.geometry.cell_center[c].y;
      out[3 * c + 2] = s->s.geometry.cell_center[c].z;
    }
  });
}

int gsk_setup_mass_flow_scale(const gsk_setup* s, double* out) {
  return guarded([&] {
    need(s, "setup");
    need(out, "out");
    // nondimensional mass flow -> Mdot: (mass flux scale L^2) / mass_flow_scale (cmd_cis_run.cpp:451-455)
    const auto& d = s->s.derived;
    *out = d.ref_mass_flux * d.ref_length * d.ref_length / d.mass_flow_scale;
  });
}

int gsk_setup_write_pmesh(const gsk_setup* s, const char* path) {
  return guarded([&] {
    need(s, "setup");
    need(path, "path");
    gsk::write_pmesh(path, s->s.mesh);
  });
}

// ------------------------------------------------------------------ backend
int gsk_backend_create(const gsk_setup_view* view, gsk_backend** out) {
  return guarded([&] {
    need(view, "view");
    need(out, "out");
    *out = nullptr;
    auto h = std::make_unique<gsk_backend>();
    h->owner = gsk::make_hip_backend(*view);
    h->b = h->owner.get();
    *out = h.release();
  });
}

int gsk_backend_destroy(gsk_backend* b) {
  return guarded([&] {
    if (b && b->owner) delete b;
  });
}

const char* gsk_backend_name(const gsk_backend* b) { return b && b->b ? b->b->name() : ""; }

int gsk_backend_chunk_width(const gsk_backend* b, int32_t* out) {
  return guarded([&] {
    need(b, "backend");
    need(out, "out");
    *out = b->b->chunk_width();
  });
}

int gsk_backend_memory_bytes(const gsk_backend* b, uint64_t* out) {
  return guarded([&] {
    need(b, "backend");
    need(out, "out");
    *out = b->b->memory_bytes();
  });
}

#define GSK_BACKEND_CALL(b, expr) \
  return guarded([&] {            \
    need(b, "backend");           \
    b->b->expr;                   \
  })

int gsk_backend_set_equilibrium(gsk_backend* b, const double* states, size_t n) {
  GSK_BACKEND_CALL(b, set_equilibrium(states, n));
}
int gsk_backend_set_distribution(gsk_backend* b, const double* f, size_t n) { GSK_BACKEND_CALL(b, set_distribution(f, n)); }
int gsk_backend_get_distribution(gsk_backend* b, double* f, size_t n) { GSK_BACKEND_CALL(b, get_distribution(f, n)); }
int gsk_backend_compute_moments(gsk_backend* b, double* out, size_t n) { GSK_BACKEND_CALL(b, compute_moments(out, n)); }
int gsk_backend_scale_distribution(gsk_backend* b, double s) { GSK_BACKEND_CALL(b, scale_distribution(s)); }
int gsk_backend_update_cell_params(gsk_backend* b) { GSK_BACKEND_CALL(b, update_cell_params()); }
int gsk_backend_set_inflow(gsk_backend* b, const gsk_inflow_state* inflow, size_t n) {
  GSK_BACKEND_CALL(b, set_inflow(inflow, n));
}
int gsk_backend_boundary_sums(gsk_backend* b, double* out, size_t n) { GSK_BACKEND_CALL(b, boundary_sums(out, n)); }
int gsk_backend_sweep(gsk_backend* b, double* residual_cell, size_t n) { GSK_BACKEND_CALL(b, sweep(residual_cell, n)); }

int gsk_backend_gsis_correction(gsk_backend* b, const double* w_half, const double* w_new, size_t n_w,
                                int32_t* replaced_flags, size_t n_flags, long long* replaced) {
  return guarded([&] {
    need(b, "backend");
    const long long r = b->b->gsis_correction(w_half, w_new, n_w, replaced_flags, replaced_flags ? n_flags : 0);
    if (replaced) *replaced = r;
  });
}

int gsk_backend_internal_face_mass_flux(gsk_backend* b, const int32_t* cells, const int32_t* entries, size_t n,
                                        double* out) {
  GSK_BACKEND_CALL(b, internal_face_mass_flux(cells, entries, n, out));
}

int gsk_backend_take_timings(gsk_backend* b, gsk_timings* out) {
  return guarded([&] {
    need(b, "backend");
    need(out, "out");
    *out = b->b->take_timings();
  });
}

int gsk_backend_deq_counts(gsk_backend* b, gsk_deq_counts* out) {
  return guarded([&] {
    need(b, "backend");
    need(out, "out");
    *out = b->b->deq_counts();
  });
}

// ------------------------------------------------------------------ solver
int gsk_solver_create(const gsk_setup* s, gsk_solver** out) {
  return guarded([&] {
    need(s, "setup");
    need(out, "out");
    *out = nullptr;
    auto h = std::make_unique<gsk_solver>();
    const gsk_setup_view v = s->s.view();
    h->sv = std::make_unique<gsk::KineticSolver>(s->s, gsk::make_hip_backend(v));
    *out = h.release();
  });
}

namespace {
// Backend whose methods are C function pointers supplied by the caller.
class ExternalBackend final : public gsk::Backend {
 public:
  explicit ExternalBackend(const gsk_backend_ops& ops) : o_(ops) {
    if (!o_.set_equilibrium || !o_.set_distribution || !o_.get_distribution || !o_.compute_moments ||
        !o_.scale_distribution || !o_.update_cell_params || !o_.set_inflow || !o_.boundary_sums || !o_.sweep ||
        !o_.gsis_correction || !o_.internal_face_mass_flux)
      throw gsk::Error("external backend ops table is incomplete");
  }
  const char* name() const override { return o_.name ? o_.name(o_.ctx) : "external"; }
  int chunk_width() const override {
    int32_t k = 0;
    if (o_.chunk_width) chk(o_.chunk_width(o_.ctx, &k));
    return k;
  }
  std::size_t memory_bytes() const override {
    uint64_t m = 0;
    if (o_.memory_bytes) chk(o_.memory_bytes(o_.ctx, &m));
    return m;
  }
  void set_equilibrium(const double* s, std::size_t n) override { chk(o_.set_equilibrium(o_.ctx, s, n)); }
  void set_distribution(const double* f, std::size_t n) override { chk(o_.set_distribution(o_.ctx, f, n)); }
  void get_distribution(double* f, std::size_t n) override { chk(o_.get_distribution(o_.ctx, f, n)); }
  void compute_moments(double* out, std::size_t n) override { chk(o_.compute_moments(o_.ctx, out, n)); }
  void scale_distribution(double s) override { chk(o_.scale_distribution(o_.ctx, s)); }
  void update_cell_params() override { chk(o_.update_cell_params(o_.ctx)); }
  void set_inflow(const gsk_inflow_state* in, std::size_t n) override { chk(o_.set_inflow(o_.ctx, in, n)); }
  void boundary_sums(double* out, std::size_t n) override { chk(o_.boundary_sums(o_.ctx, out, n)); }
  void sweep(double* r, std::size_t n) override { chk(o_.sweep(o_.ctx, r, n)); }
  long long gsis_correction(const double* wh, const double* wn, std::size_t nw, int* flags, std::size_t nf) override {
    long long r = 0;
    chk(o_.gsis_correction(o_.ctx, wh, wn, nw, flags, nf, &r));
    return r;
  }
  void internal_face_mass_flux(const int* c, const int* e, std::size_t n, double* out) override {
    chk(o_.internal_face_mass_flux(o_.ctx, c, e, n, out));
  }
  gsk_timings take_timings() override {
    gsk_timings t{};
    if (o_.take_timings) chk(o_.take_timings(o_.ctx, &t));
    return t;
  }
  gsk_deq_counts deq_counts() override {
    gsk_deq_counts d{0, 0, 0};
    if (o_.deq_counts) chk(o_.deq_counts(o_.ctx, &d));
    return d;
  }

 private:
  void chk(int status) const {
    if (status != GSK_OK)
      throw gsk::Error(std::string("external backend: ") + (o_.last_error ? o_.last_error(o_.ctx) : "failed"));
  }
  gsk_backend_ops o_;
};
}  // namespace

int gsk_solver_create_with_backend(const gsk_setup* s, const gsk_backend_ops* ops, gsk_solver** out) {
  return guarded([&] {
    need(s, "setup");
    need(ops, "ops");
    need(out, "out");
    *out = nullptr;
    auto h = std::make_unique<gsk_solver>();
    h->sv = std::make_unique<gsk::KineticSolver>(s->s, std::make_unique<ExternalBackend>(*ops));
    *out = h.release();
  });
}

int gsk_solver_destroy(gsk_solver* sv) {
  return guarded([&] {
    delete sv;
  });
}

int gsk_solver_initialize_uniform(gsk_solver* sv, double rho, double ux, double uy, double uz, double T) {
  return guarded([&] {
    need(sv, "solver");
    sv->sv->initialize_uniform(rho, gsk::Vec3{ux, uy, uz}, T);
  });
}

int gsk_solver_initialize_equilibrium(gsk_solver* sv, const double* states, size_t n) {
  return guarded([&] {
    need(sv, "solver");
    sv->sv->initialize_equilibrium(states, n);
  });
}

int gsk_solver_prepare(gsk_solver* sv, gsk_iteration_info* info) {
  return guarded([&] {
    need(sv, "solver");
    const gsk_iteration_info& i = sv->sv->prepare();
    if (info) *info = i;
  });
}

int gsk_solver_update_mass_flow_pressure(gsk_solver* sv) {
  return guarded([&] {
    need(sv, "solver");
    sv->sv->update_mass_flow_pressure();
  });
}

int gsk_solver_sweep(gsk_solver* sv, int32_t update_pressure, gsk_iteration_info* info) {
  return guarded([&] {
    need(sv, "solver");
    const gsk_iteration_info& i = sv->sv->sweep(update_pressure != 0);
    if (info) *info = i;
  });
}

int gsk_solver_patch_mass_flow(gsk_solver* sv, double* out, size_t n_patches) {
  return guarded([&] {
    need(sv, "solver");
    need(out, "out");
    const auto& m = sv->sv->patch_mass_flow();
    if (n_patches != sv->sv->setup().patches.size()) throw gsk::Error("patch count differs");
    for (std::size_t i = 0; i < n_patches; ++i) out[i] = i < m.size() ? m[i] : 0.0;
  });
}

int gsk_solver_moments(gsk_solver* sv, double* out, size_t n) {
  return guarded([&] {
    need(sv, "solver");
    need(out, "out");
    sv->sv->moments(out, n);
  });
}

int gsk_solver_boundary_fluxes(gsk_solver* sv, double* f_plus, double* f_minus, double* face_state, size_t n_bfaces) {
  return guarded([&] {
    need(sv, "solver");
    need(f_plus, "f_plus");
    need(f_minus, "f_minus");
    need(face_state, "face_state");
    sv->sv->boundary_fluxes(f_plus, f_minus, face_state, n_bfaces);
  });
}

int gsk_solver_apply_gsis_correction(gsk_solver* sv, const double* w_half, const double* w_new, size_t n,
                                     long long* replaced) {
  return guarded([&] {
    need(sv, "solver");
    const long long r = sv->sv->apply_gsis_correction(w_half, w_new, n);
    if (replaced) *replaced = r;
  });
}

int gsk_solver_distribution(gsk_solver* sv, double* f, size_t n) {
  return guarded([&] {
    need(sv, "solver");
    need(f, "f");
    sv->sv->distribution(f, n);
  });
}

int gsk_solver_set_distribution(gsk_solver* sv, const double* f, size_t n) {
  return guarded([&] {
    need(sv, "solver");
    need(f, "f");
    sv->sv->set_distribution(f, n);
  });
}

int gsk_solver_residual_per_cell(gsk_solver* sv, double* out, size_t n) {
  return guarded([&] {
    need(sv, "solver");
    need(out, "out");
    sv->sv->residual_per_cell(out, n);
  });
}

int gsk_solver_element_size(gsk_solver* sv, int32_t* out) {
  return guarded([&] {
    need(sv, "solver");
    need(out, "out");
    *out = sv->sv->element_size();
  });
}

int gsk_solver_distribution_device(gsk_solver* sv, void* out, size_t n) {
  return guarded([&] {
    need(sv, "solver");
    need(out, "out");
    sv->sv->distribution_device(out, n);
  });
}

int gsk_solver_set_distribution_device(gsk_solver* sv, const void* in, size_t n) {
  return guarded([&] {
    need(sv, "solver");
    need(in, "in");
    sv->sv->set_distribution_device(in, n);
  });
}

int gsk_solver_residual_device(gsk_solver* sv, void* out, size_t n) {
  return guarded([&] {
    need(sv, "solver");
    need(out, "out");
    sv->sv->residual_device(out, n);
  });
}

int gsk_solver_precondition_device(gsk_solver* sv, const void* r, void* out, size_t n) {
  return guarded([&] {
    need(sv, "solver");
    need(r, "r");
    need(out, "out");
    sv->sv->precondition_device(r, out, n);
  });
}

int gsk_solver_mu0(gsk_solver* sv, double* out) {
  return guarded([&] {
    need(sv, "solver");
    need(out, "out");
    *out = sv->sv->mu0();
  });
}

int gsk_solver_set_mu0(gsk_solver* sv, double mu0) {
  return guarded([&] {
    need(sv, "solver");
    sv->sv->set_mu0(mu0);
  });
}

int gsk_solver_set_accommodation(gsk_solver* sv, int32_t patch, double alpha) {
  return guarded([&] {
    need(sv, "solver");
    sv->sv->set_accommodation(patch, alpha);
  });
}

int gsk_solver_set_newton_tolerances(gsk_solver* sv, double cell_tol, double inflow_tol) {
  return guarded([&] {
    need(sv, "solver");
    sv->sv->set_newton_tolerances(cell_tol, inflow_tol);
  });
}

int gsk_solver_set_boundary_iterations(gsk_solver* sv, int64_t n) {
  return guarded([&] {
    need(sv, "solver");
    sv->sv->set_boundary_iterations(n);
  });
}

gsk_backend* gsk_solver_backend(gsk_solver* sv) {
  if (!sv) return nullptr;
  sv->handle.b = &sv->sv->backend();
  return &sv->handle;
}

}  // extern "C"
