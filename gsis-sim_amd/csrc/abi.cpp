// extern "C" entry points declared in include/gsis_kinetic.h. Every call catches exceptions and
// turns them into a status code plus a thread-local message (the reference throws gsis::Error,
// src/core/error.hpp:16-47; a C++ host rethrows the message).

#include <algorithm>
#include <cstdio>
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

// Minimal JSON object writer for the case description: keys in insertion order, doubles with 17
// significant digits (they read back to the same value), strings escaped.
class JsonOut {
 public:
  JsonOut& key(const char* k) {
    sep();
    str(k);
    out_ += ':';
    pending_value_ = true;
    return *this;
  }
  JsonOut& val(double d) {
    char b[40];
    std::snprintf(b, sizeof(b), "%.17g", d);
    std::string t = b;
    if (t.find_first_of(".eEn") == std::string::npos) t += ".0";  // stays a real number when read back
    return raw(t);
  }
  JsonOut& val(long long i) { return raw(std::to_string(i)); }
  JsonOut& val(int i) { return raw(std::to_string(i)); }
  JsonOut& val(bool v) { return raw(v ? "true" : "false"); }
  JsonOut& val(const std::string& v) {
    if (!pending_value_) sep();
    pending_value_ = false;
    str(v.c_str());
    return *this;
  }
  JsonOut& open(char c) {
    if (!pending_value_) sep();
    pending_value_ = false;
    out_ += c;
    first_ = true;
    return *this;
  }
  JsonOut& close(char c) {
    out_ += c;
    first_ = false;
    return *this;
  }
  const std::string& text() const { return out_; }

 private:
  JsonOut& raw(const std::string& t) {
    if (!pending_value_) sep();
    pending_value_ = false;
    out_ += t;
    return *this;
  }
  void sep() {
    if (!first_) out_ += ',';
    first_ = false;
  }
  void str(const char* t) {
    out_ += '"';
    for (const char* c = t; *c; ++c) {
      const unsigned char u = static_cast<unsigned char>(*c);
      if (*c == '"' || *c == '\\') {
        out_ += '\\';
        out_ += *c;
      } else if (u < 0x20) {
        char b[8];
        std::snprintf(b, sizeof(b), "\\u%04x", u);
        out_ += b;
      } else {
        out_ += *c;
      }
    }
    out_ += '"';
  }
  std::string out_;
  bool first_ = true, pending_value_ = false;
};

const char* boundary_type_name(gsk::BoundaryType t) {
  switch (t) {
    case gsk::BoundaryType::MaxwellWall: return "maxwell_wall";
    case gsk::BoundaryType::MassFlowInlet: return "mass_flow_inlet";
    case gsk::BoundaryType::DiffuseWall: return "diffuse_wall";
    case gsk::BoundaryType::PressureInlet: return "pressure_inlet";
    case gsk::BoundaryType::PressureOutlet: return "pressure_outlet";
    case gsk::BoundaryType::Vacuum: return "vacuum";
    case gsk::BoundaryType::Symmetry: return "symmetry";
    case gsk::BoundaryType::PeriodicSlab: return "periodic_slab";
  }
  return "unknown";
}

// What a cis-run driver writes into its result files besides the iteration records (the case,
// sizes, scales, [kinetic] parameters and patches: src/app/cmd_cis_run.cpp:360-443), as JSON.
std::string describe_json(const gsk::KineticSetup& k) {
  const gsk::DerivedQuantities& d = k.derived;
  const gsk::KineticParams& p = k.params;
  long long cut = 0;  // internal faces between sweep blocks (block_partition_stats, src/mesh/ordering.cpp:265-286)
  gsk::Index min_cells = k.n_cells(), max_cells = 0;
  {
    std::vector<int> block(static_cast<std::size_t>(k.n_cells()));
    for (int b = 0; b < k.n_blocks(); ++b) {
      const gsk::Index lo = k.block_offsets[static_cast<std::size_t>(b)], hi = k.block_offsets[static_cast<std::size_t>(b) + 1];
      min_cells = std::min(min_cells, hi - lo);
      max_cells = std::max(max_cells, hi - lo);
      for (gsk::Index c = lo; c < hi; ++c) block[static_cast<std::size_t>(c)] = b;
    }
    for (gsk::Index f = 0; f < k.mesh.n_internal_faces(); ++f)
      if (block[static_cast<std::size_t>(k.mesh.owner[static_cast<std::size_t>(f)])] !=
          block[static_cast<std::size_t>(k.mesh.neighbour[static_cast<std::size_t>(f)])])
        ++cut;
  }
  JsonOut j;
  j.open('{');
  j.key("case").val(k.case_name).key("mesh").val(k.mesh_file);
  j.key("n_cells").val(static_cast<long long>(k.n_cells())).key("n_boundary_faces").val(static_cast<long long>(k.n_bfaces()));
  j.key("n_velocity").val(static_cast<long long>(k.nv())).key("orbit_size").val(k.velocity.orbit_size);
  j.key("Kn").val(d.Kn).key("delta").val(d.delta_from_Kn).key("has_p_in").val(d.has_p_in);
  j.key("mass_flow_scale_kg_s").val(d.mass_flow_scale).key("length_scale").val(k.length_scale);
  j.key("closed_domain").val(k.closed_domain);
  j.key("tolerance").val(k.gsis_tolerance).key("max_iterations").val(p.max_iterations);
  j.key("mass_balance_tolerance").val(k.mass_balance_tolerance).key("mass_balance_required").val(k.mass_balance_required);
  j.key("sweep_blocks").val(static_cast<long long>(k.n_blocks())).key("block_cut_faces").val(cut);
  j.key("block_min_cells").val(static_cast<long long>(min_cells)).key("block_max_cells").val(static_cast<long long>(max_cells));
  j.key("sections").open('{').key("axis").val(std::string(k.section_axis == 0 ? "x" : k.section_axis == 1 ? "y" : "z"));
  j.key("positions").open('[');
  for (const double x : k.section_positions) j.val(x);
  j.close(']').close('}');
  j.key("reference").open('{');
  j.key("length_m").val(d.ref_length).key("temperature_K").val(d.ref_temperature).key("pressure_Pa").val(d.ref_pressure);
  j.key("density_kg_m3").val(d.ref_density).key("velocity_m_s").val(d.ref_velocity);
  j.key("heat_flux_W_m2").val(d.ref_heat_flux).key("mass_flux_kg_m2_s").val(d.ref_mass_flux);
  j.close('}');
  j.key("params").open('{');
  j.key("pseudo_cfl").val(p.pseudo_cfl).key("pseudo_time").val(p.pseudo_time).key("limiter").val(p.limiter);
  j.key("limiter_k").val(p.limiter_k).key("limiter_eps").val(p.limiter_eps).key("lsq_weight_power").val(p.lsq_weight_power);
  j.key("cell_ordering").val(p.cell_ordering).key("initial_pressure").val(p.initial_pressure);
  j.key("initial_temperature").val(p.initial_temperature).key("convergence_u_floor").val(p.convergence_u_floor);
  j.key("wall_relaxation").val(p.wall_relaxation).key("boundary_iterations").val(p.boundary_iterations);
  j.key("interface_damping").val(p.interface_damping).key("positivity_limiter").val(p.positivity_limiter);
  j.key("sweep_order").val(p.sweep_order).key("inner_sweeps").val(p.inner_sweeps);
  j.key("inner_wall_update").val(p.inner_wall_update).key("block_passes").val(p.block_passes);
  j.key("equilibrium").val(p.equilibrium).key("deq_fallback").val(p.deq_fallback);
  j.key("pressure_boundary_reconstruction").val(p.pressure_boundary_reconstruction);
  j.key("cold_bgk_temperature").val(p.cold_bgk_temperature).key("precision").val(p.precision);
  j.key("backend").val(p.backend).key("moment_sums").val(p.moment_sums);
  j.close('}');
  j.key("patches").open('[');
  for (const gsk::PatchCondition& pc : k.patches) {
    j.open('{');
    j.key("name").val(pc.name).key("type").val(std::string(boundary_type_name(pc.type)));
    j.key("n_faces").val(static_cast<long long>(pc.n_faces)).key("accommodation").val(pc.accommodation);
    j.key("p").val(pc.p).key("T").val(pc.T).key("target_mass_flow").val(pc.target_mass_flow);
    j.close('}');
  }
  j.close(']');
  j.close('}');
  return j.text();
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

int gsk_setup_describe(const gsk_setup* s, char* buf, size_t cap, size_t* needed) {
  return guarded([&] {
    need(s, "setup");
    const std::string text = describe_json(s->s);
    if (needed) *needed = text.size() + 1;
    if (buf == nullptr || cap == 0) return;
    GSK_CHECK(cap > text.size(), "description buffer too small");
    std::memcpy(buf, text.c_str(), text.size() + 1);
  });
}

int gsk_setup_bface_geometry(const gsk_setup* s, int32_t* patch, double* area, double* normal) {
  return guarded([&] {
    need(s, "setup");
    const gsk::KineticSetup& k = s->s;
    const gsk::Index n_int = k.mesh.n_internal_faces();
    for (gsk::Index b = 0; b < k.n_bfaces(); ++b) {
      const auto o = static_cast<std::size_t>(k.bface_new_to_old[static_cast<std::size_t>(b)]);
      const auto f = static_cast<std::size_t>(n_int + b);
      if (patch) patch[o] = k.bface_patch[static_cast<std::size_t>(b)];
      if (area) area[o] = k.geometry.face_area[f];
      if (normal) {
        normal[3 * o + 0] = k.geometry.face_normal[f].x;
        normal[3 * o + 1] = k.geometry.face_normal[f].y;
        normal[3 * o + 2] = k.geometry.face_normal[f].z;
      }
    }
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
    need(w_half, "w_half");
    need(w_new, "w_new");
    const long long r = b->b->gsis_correction(w_half, w_new, n_w, replaced_flags, replaced_flags ? n_flags : 0);
    if (replaced) *replaced = r;
  });
}

int gsk_backend_gsis_correction_noalloc(gsk_backend* b, const double* w_half, const double* w_new, size_t n_w,
                                        int32_t* replaced_flags, size_t n_flags, long long* replaced) {
  return guarded([&] {
    need(b, "backend");
    need(w_half, "w_half");
    need(w_new, "w_new");
    const long long r =
        b->b->gsis_correction_noalloc(w_half, w_new, n_w, replaced_flags, replaced_flags ? n_flags : 0);
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
    need(w_half, "w_half");
    need(w_new, "w_new");
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

int gsk_solver_section_mass_flow(gsk_solver* sv, int32_t axis, const double* positions, size_t n, double* out) {
  return guarded([&] {
    need(sv, "solver");
    if (n == 0) return;
    need(positions, "positions");
    need(out, "out");
    sv->sv->section_mass_flow(axis, std::vector<double>(positions, positions + n), out);
  });
}

int gsk_solver_section_cuts(gsk_solver* sv, int32_t axis, const double* positions, size_t n, int64_t* n_faces,
                            double* area, double* projected_area, double* face_position) {
  return guarded([&] {
    need(sv, "solver");
    if (n == 0) return;
    need(positions, "positions");
    const auto& cuts = sv->sv->section_cuts(axis, std::vector<double>(positions, positions + n));
    for (std::size_t k = 0; k < cuts.size(); ++k) {
      if (n_faces) n_faces[k] = static_cast<int64_t>(cuts[k].faces.size());
      if (area) area[k] = cuts[k].area;
      if (projected_area) projected_area[k] = cuts[k].projected_area;
      if (face_position) face_position[k] = cuts[k].face_position;
    }
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

int gsk_solver_freeze_limiters(gsk_solver* sv, int32_t on) {
  return guarded([&] {
    need(sv, "solver");
    sv->sv->freeze_limiters(on != 0);
  });
}

int gsk_solver_adjoint_prepare(gsk_solver* sv) {
  return guarded([&] {
    need(sv, "solver");
    sv->sv->adjoint_prepare();
  });
}

int gsk_solver_adjoint_apply_device(gsk_solver* sv, const void* lam, const double* patch_weights, size_t n_patches,
                                    void* out, size_t n) {
  return guarded([&] {
    need(sv, "solver");
    need(patch_weights, "patch_weights");
    need(out, "out");
    sv->sv->adjoint_apply_device(lam, patch_weights, n_patches, out, n);
  });
}

int gsk_solver_set_boundary_iterations(gsk_solver* sv, int64_t n) {
  return guarded([&] {
    need(sv, "solver");
    sv->sv->set_boundary_iterations(n);
  });
}

int gsk_solver_set_device_prepare(gsk_solver* sv, int32_t on) {
  return guarded([&] {
    need(sv, "solver");
    sv->sv->set_device_prepare(on != 0);
  });
}

int gsk_solver_device_prepare(gsk_solver* sv, int32_t* on_device) {
  return guarded([&] {
    need(sv, "solver");
    need(on_device, "on_device");
    *on_device = sv->sv->prepared_on_device() ? 1 : 0;
  });
}

gsk_backend* gsk_solver_backend(gsk_solver* sv) {
  if (!sv) return nullptr;
  sv->handle.b = &sv->sv->backend();
  return &sv->handle;
}

}  // extern "C"
