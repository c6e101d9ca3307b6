/*
 * gsis_kinetic.h - C ABI of the MI355X-native kinetic step (Shakhov DVM: LU-SGS sweeps,
 * velocity-space moments, discrete Shakhov equilibrium).
 *
 * Three layers, each replacing one reference interface (paths relative to the reference tree):
 *
 *   setup   gsk_setup_*    build_kinetic_setup + load_case + load_pmesh
 *                          (src/kinetic/kinetic_setup.hpp:56-115, :118-120;
 *                           src/config/case_config.hpp:174; src/mesh/pmesh_io.hpp:22)
 *   backend gsk_backend_*  class KineticBackend (src/kinetic/kinetic_backend.hpp:59-100) and
 *                          make_cuda_backend (kinetic_backend.hpp:105)
 *   solver  gsk_solver_*   class KineticSolver (src/kinetic/kinetic_solver.hpp:61-115)
 *
 * Conventions (kinetic_backend.hpp:1-7): host data is double, nondimensional (p = rho T),
 * in the solver's internal numbering (RCB/RCM-renumbered cells and boundary faces); f is
 * [cell][storage position]. The solver layer converts to the original numbering and grid
 * order like KineticSolver does (kinetic_solver.cpp:584-722).
 *
 * Errors: every call returns GSK_OK (0) or a negative status; gsk_last_error() returns the
 * message of the last failure on the calling thread (the reference raises gsis::Error,
 * src/core/error.hpp:16-47; a C++ host rethrows it). Sizes passed in are checked exactly.
 * Threading: one host thread per handle, not re-entrant; calls that return host data
 * synchronize the device (kinetic_backend_cuda.cpp:221).
 */
#ifndef GSIS_KINETIC_H
#define GSIS_KINETIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSK_OK 0
#define GSK_ERR_INVALID (-1) /* bad argument / size mismatch / invalid state (gsis::Error) */
#define GSK_ERR_DEVICE (-2)  /* HIP runtime failure (throw_cuda_error, src/gpu/cuda_check.hpp:16-24) */
#define GSK_ERR_MEMORY (-3)  /* does not fit in device memory (kinetic_backend_cuda.cpp:103-106) */

/* Boundary kinds (src/kinetic/kinetic_types.hpp:21-28) */
#define GSK_BC_DIFFUSE_WALL 0
#define GSK_BC_PRESSURE 1
#define GSK_BC_VACUUM 2
#define GSK_BC_SYMMETRY 3
#define GSK_BC_PERIODIC 4
#define GSK_BC_MAXWELL_WALL 5

/* One cell->face CSR entry seen from cell i (CsrEntry, kinetic_types.hpp:34-42), FP64. */
typedef struct gsk_csr_entry {
  int32_t other;    /* neighbour cell, or -1 on a boundary face */
  int32_t bface;    /* boundary face number, or -1 on an internal face */
  double s[3];      /* outward area vector S_out (nondimensional) */
  double ds[3];     /* x_f - x_i */
  double dother[3]; /* x_f - x_j (internal faces only) */
  double c[3];      /* least-squares coefficient c_k */
} gsk_csr_entry;

/* Fixed boundary-face data (BoundaryFaceInfo, kinetic_types.hpp:45-53), FP64. */
typedef struct gsk_bface {
  int32_t owner, kind, mirror_xor, patch;
  double n[3]; /* outward unit normal */
  double area;
  double d[3]; /* x_f - x_owner */
} gsk_bface;

/* Inflow state scale * f_eq(rho, u, T) per boundary face (InflowState, kinetic_backend.hpp:46-51). */
typedef struct gsk_inflow_state {
  double scale, rho, ux, uy, uz, T;
} gsk_inflow_state;

/* Per-phase seconds since the last take_timings (StepTimings, kinetic_backend.hpp:19-41). */
typedef struct gsk_timings {
  double moments, boundary, gradient, residual, sweep_forward, sweep_backward, other;
} gsk_timings;

/* Discrete-equilibrium Newton record of the last update_cell_params (DeqCounts, kinetic_backend.hpp:54-58). */
typedef struct gsk_deq_counts {
  long long corrected, failed;
  int max_iterations;
} gsk_deq_counts;

/*
 * Everything a backend needs from the setup, as plain arrays (the tables the reference builds in
 * kinetic_setup.cpp:516-599 and reads through KineticView, kinetic_types.hpp:142-192).
 * All pointers are borrowed for the duration of gsk_backend_create only (the backend copies).
 */
typedef struct gsk_setup_view {
  int32_t n_cells, nv, n_bfaces, n_blocks;
  int32_t orbit_size;
  int64_t velocity_chunk; /* [kinetic].velocity_chunk (0 = automatic) */
  const double *xi_x, *xi_y, *xi_z; /* velocity points in storage order */
  double weight;                    /* uniform quadrature weight */
  const int32_t* csr_offsets;       /* n_cells + 1 */
  const gsk_csr_entry* csr;         /* csr_offsets[n_cells] */
  const double* eps2;               /* K^3 V per cell */
  const double* volume;             /* V per cell */
  const double* vdt;                /* V / dt per cell */
  const gsk_bface* bfaces;          /* n_bfaces */
  const int32_t* block_offsets;     /* n_blocks + 1 */
  int32_t n_sweep_dirs;             /* 0: sweep in internal numbering */
  const int32_t* sweep_order;       /* n_sweep_dirs * n_cells */
  const int32_t* sweep_pos;         /* n_sweep_dirs * n_cells */
  const uint8_t* sweep_dir_of;      /* nv */
  const int32_t* grid_to_storage;   /* nv (used by Maxwell-wall interpolation) */
  int32_t grid_n[3];
  int32_t n_maxwell_faces;
  const int64_t* maxwell_offset;    /* n_bfaces (or NULL when n_maxwell_faces == 0) */
  const double* maxwell_specular;   /* n_bfaces specular fraction 1 - alpha */
  double deq_lo[3], deq_hi[3], deq_d[3];
  double deq_criterion;
  int32_t deq_max_iterations;
  double deq_tolerance;
  int32_t discrete_eq;          /* [kinetic].equilibrium == "discrete" */
  int32_t deq_fallback_collisionless;
  int32_t limiter_mode;         /* 0 none, 1 Venkatakrishnan absolute, 2 relative */
  int32_t positivity;           /* [kinetic].positivity_limiter */
  double diag_scale;            /* 1, or 1 + 1/CFL for pseudo_time = "velocity" */
  double interface_damping;     /* beta */
  double cold_bgk_temperature;
  double mu0, omega, Pr;        /* mu = mu0 T^omega */
  int32_t block_passes;
  int32_t negative_check;       /* 0 all, 1 created, 2 limit, 3 blend */
  int32_t precision_fp64;       /* 1: f stored and swept in FP64; 0: FP32 */
  int32_t moments_mfma;         /* [kinetic].moment_sums == "mfma" (FP32 form only): moment sums as an
                                   FP32 matrix-core contraction, added in FP64 (no reference counterpart) */
  /* [velocity_grid].stretch > 0 (no reference counterpart): per-velocity quadrature weights as
   * weight * weight_rel[p] (storage order), and the node coordinates of axes x, y, z
   * (grid_n[0] + grid_n[1] + grid_n[2] values) for the Maxwell-wall interpolation. Both NULL on the
   * reference's uniform grid. */
  const double* weight_rel;
  const double* grid_nodes;
} gsk_setup_view;

const char* gsk_last_error(void);
const char* gsk_version(void);

/* ---------------------------------------------------------------- setup ----------------- */
typedef struct gsk_setup gsk_setup;

/* Reads a case TOML and its .pmesh and builds the setup (load_case + load_pmesh +
 * build_kinetic_setup, kinetic_setup.cpp:293-514). overrides: "key=value" lines applied to
 * [kinetic] like `cis-run --set` (src/app/cmd_cis_run.cpp:182-203); may be NULL. */
int gsk_setup_from_case(const char* toml_path, const char* overrides, gsk_setup** out);
int gsk_setup_destroy(gsk_setup* s);
/* Borrowed view of the backend tables; valid until gsk_setup_destroy. */
int gsk_setup_view_get(const gsk_setup* s, gsk_setup_view* out);
/* Sizes and numbering maps (for converting to original numbering). */
int gsk_setup_sizes(const gsk_setup* s, int32_t* n_cells, int32_t* nv, int32_t* n_bfaces, int32_t* n_blocks,
                    int32_t* n_patches);
int gsk_setup_cell_new_to_old(const gsk_setup* s, int32_t* out);          /* n_cells */
int gsk_setup_storage_to_grid(const gsk_setup* s, int32_t* out);          /* nv */
int gsk_setup_cell_centers(const gsk_setup* s, double* out);              /* 3 n_cells, internal numbering, m */
int gsk_setup_mass_flow_scale(const gsk_setup* s, double* out);           /* nondimensional -> Mdot denominator */
int gsk_setup_write_pmesh(const gsk_setup* s, const char* path);          /* renumbered mesh (diagnostics) */
/* What a cis-run driver records besides the iteration data (src/app/cmd_cis_run.cpp:360-443,
 * src/kinetic/kinetic_output.cpp:198-389): a JSON object with the case name and mesh file, sizes,
 * Kn and delta, the reference scales, the [gsis] tolerance and mass-balance settings, the sweep-block
 * statistics, the [kinetic] parameters and the patches (name, type, faces, accommodation, p, T,
 * target mass flow) in patch order. Written NUL-terminated into buf (cap bytes); *needed = its
 * length + 1 (call with buf = NULL to size it). */
int gsk_setup_describe(const gsk_setup* s, char* buf, size_t cap, size_t* needed);
/* Per boundary face in the ORIGINAL boundary-face numbering: patch index, area (mesh units^2) and
 * outward unit normal (3 per face); any output may be NULL. */
int gsk_setup_bface_geometry(const gsk_setup* s, int32_t* patch, double* area, double* normal);

/* ---------------------------------------------------------------- backend --------------- */
typedef struct gsk_backend gsk_backend;

/* make_cuda_backend: HIP backend on the current device; uploads all tables. */
int gsk_backend_create(const gsk_setup_view* view, gsk_backend** out);
int gsk_backend_destroy(gsk_backend* b);
const char* gsk_backend_name(const gsk_backend* b);
int gsk_backend_chunk_width(const gsk_backend* b, int32_t* out);
int gsk_backend_memory_bytes(const gsk_backend* b, uint64_t* out);
int gsk_backend_set_equilibrium(gsk_backend* b, const double* states, size_t n);   /* n = 5 n_cells */
int gsk_backend_set_distribution(gsk_backend* b, const double* f, size_t n);       /* n = n_cells nv */
int gsk_backend_get_distribution(gsk_backend* b, double* f, size_t n);
int gsk_backend_compute_moments(gsk_backend* b, double* out, size_t n);            /* n = 15 n_cells */
int gsk_backend_scale_distribution(gsk_backend* b, double s);
int gsk_backend_update_cell_params(gsk_backend* b);
int gsk_backend_set_inflow(gsk_backend* b, const gsk_inflow_state* inflow, size_t n); /* n = n_bfaces */
int gsk_backend_boundary_sums(gsk_backend* b, double* out, size_t n);              /* n = 20 n_bfaces */
int gsk_backend_sweep(gsk_backend* b, double* residual_cell, size_t n);            /* n = n_cells */
int gsk_backend_gsis_correction(gsk_backend* b, const double* w_half, const double* w_new, size_t n_w,
                                int32_t* replaced_flags, size_t n_flags, long long* replaced);
int gsk_backend_internal_face_mass_flux(gsk_backend* b, const int32_t* cells, const int32_t* entries, size_t n,
                                        double* out);
/* gsk_backend_gsis_correction with no device allocation. No counterpart in the reference: for
 * callers that cannot call hipMalloc (code under HIP graph capture, serving loops, harnesses that
 * forbid allocation). Same arguments, results and errors as gsk_backend_gsis_correction (the same
 * kernel), but w_half and w_new are staged in a fixed device array of GSK_NOALLOC_SCRATCH_BYTES that
 * each GPU the library runs on holds until the process ends, so at most GSK_NOALLOC_MAX_CELLS cells
 * (larger cases return GSK_ERR_INVALID). Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (several backends or threads) race. Records no phase timing. */
#define GSK_NOALLOC_SCRATCH_BYTES (16 * 1024 * 1024)
#define GSK_NOALLOC_MAX_CELLS ((GSK_NOALLOC_SCRATCH_BYTES - 256) / 80)
int gsk_backend_gsis_correction_noalloc(gsk_backend* b, const double* w_half, const double* w_new, size_t n_w,
                                        int32_t* replaced_flags, size_t n_flags, long long* replaced);
int gsk_backend_take_timings(gsk_backend* b, gsk_timings* out);
int gsk_backend_deq_counts(gsk_backend* b, gsk_deq_counts* out);

/* ---------------------------------------------------------------- solver ---------------- */
typedef struct gsk_solver gsk_solver;

/* One CIS iteration record (KineticIterationInfo, kinetic_solver.hpp:37-59). */
typedef struct gsk_iteration_info {
  long long iteration;
  double epsilon, eps_rho, eps_u, eps_T;
  int32_t has_epsilon;
  double residual;
  int32_t has_residual;
  double total_mass, mass_scale;
  double q_mean[3];
  double sigma_mean[6];
  double seconds;
  gsk_timings timings;
  long long deq_corrected, deq_failed;
  int32_t deq_max_iterations;
} gsk_iteration_info;

/* KineticSolver(config, mesh, geometry, params) with the HIP backend; the setup must outlive it. */
int gsk_solver_create(const gsk_setup* s, gsk_solver** out);

/* A backend implemented outside this library with the same method set (the reference selects
 * its CPU or CUDA backend behind one interface, kinetic_solver.cpp:83-87). Every function returns
 * GSK_OK or a negative status; last_error explains a failure. */
typedef struct gsk_backend_ops {
  void* ctx;
  const char* (*name)(void* ctx);
  int (*chunk_width)(void* ctx, int32_t* out);
  int (*memory_bytes)(void* ctx, uint64_t* out);
  int (*set_equilibrium)(void* ctx, const double* states, size_t n);
  int (*set_distribution)(void* ctx, const double* f, size_t n);
  int (*get_distribution)(void* ctx, double* f, size_t n);
  int (*compute_moments)(void* ctx, double* out, size_t n);
  int (*scale_distribution)(void* ctx, double s);
  int (*update_cell_params)(void* ctx);
  int (*set_inflow)(void* ctx, const gsk_inflow_state* inflow, size_t n);
  int (*boundary_sums)(void* ctx, double* out, size_t n);
  int (*sweep)(void* ctx, double* residual_cell, size_t n);
  int (*gsis_correction)(void* ctx, const double* w_half, const double* w_new, size_t n_w, int32_t* flags,
                         size_t n_flags, long long* replaced);
  int (*internal_face_mass_flux)(void* ctx, const int32_t* cells, const int32_t* entries, size_t n, double* out);
  int (*take_timings)(void* ctx, gsk_timings* out);
  int (*deq_counts)(void* ctx, gsk_deq_counts* out);
  const char* (*last_error)(void* ctx);
} gsk_backend_ops;

/* KineticSolver over an external backend (ops copied; ops->ctx must outlive the solver). */
int gsk_solver_create_with_backend(const gsk_setup* s, const gsk_backend_ops* ops, gsk_solver** out);
int gsk_solver_destroy(gsk_solver* sv);
int gsk_solver_initialize_uniform(gsk_solver* sv, double rho, double ux, double uy, double uz, double T);
/* states: 5 per cell (rho, ux, uy, uz, T) in the ORIGINAL cell numbering */
int gsk_solver_initialize_equilibrium(gsk_solver* sv, const double* states, size_t n);
int gsk_solver_prepare(gsk_solver* sv, gsk_iteration_info* info);
int gsk_solver_update_mass_flow_pressure(gsk_solver* sv);
int gsk_solver_sweep(gsk_solver* sv, int32_t update_pressure, gsk_iteration_info* info);
/* patch mass flows of the current prepare (nondimensional, outward positive) */
int gsk_solver_patch_mass_flow(gsk_solver* sv, double* out, size_t n_patches);
/* 15 moments per cell in the original numbering (KineticSolver::moments) */
int gsk_solver_moments(gsk_solver* sv, double* out, size_t n);
/* F+ / F- (5 each) and face state (5) per boundary face, original numbering */
int gsk_solver_boundary_fluxes(gsk_solver* sv, double* f_plus, double* f_minus, double* face_state, size_t n_bfaces);
int gsk_solver_apply_gsis_correction(gsk_solver* sv, const double* w_half, const double* w_new, size_t n,
                                     long long* replaced);
/* f in the original cell numbering x velocity-grid order */
int gsk_solver_distribution(gsk_solver* sv, double* f, size_t n);
int gsk_solver_set_distribution(gsk_solver* sv, const double* f, size_t n);
int gsk_solver_residual_per_cell(gsk_solver* sv, double* out, size_t n);
/* Mass flow through sections normal to axis (0 x, 1 y, 2 z) at n positions in L_ref units: the sum
 * of the discrete face fluxes over the internal faces between a cell whose centre lies before the
 * position and one whose centre does not, along +axis, nondimensional (KineticSolver::
 * section_mass_flow, kinetic_solver.cpp:669-686; build_section_cuts, src/mesh/section_cut.cpp:48-85).
 * Prepares first. */
int gsk_solver_section_mass_flow(gsk_solver* sv, int32_t axis, const double* positions, size_t n, double* out);
/* The same sections' faces: per section the face count, area, projected area (L_ref^2 units) and the
 * area-weighted face position (KineticSolver::section_cuts); any output may be NULL. */
int gsk_solver_section_cuts(gsk_solver* sv, int32_t axis, const double* positions, size_t n, int64_t* n_faces,
                            double* area, double* projected_area, double* face_position);

/* Device-resident operators for sensitivity solves (dMdot/d delta_rp, dMdot/d alpha). No counterpart
 * in the reference; for callers that keep vectors in HBM (the Python sensitivity driver). Arrays
 * are device pointers to n = cells x velocities values in the f precision (element_size bytes
 * each), internal cell numbering and storage velocity order, all on the solver's device. */
int gsk_solver_element_size(gsk_solver* sv, int32_t* out);                         /* 8 FP64, 4 FP32 */
int gsk_solver_distribution_device(gsk_solver* sv, void* out, size_t n);           /* f -> out */
int gsk_solver_set_distribution_device(gsk_solver* sv, const void* in, size_t n);  /* keeps boundary state */
int gsk_solver_residual_device(gsk_solver* sv, void* out, size_t n);   /* prepare, then steady residual R(f) */
int gsk_solver_precondition_device(gsk_solver* sv, const void* r, void* out, size_t n); /* LU-SGS on r */
int gsk_solver_mu0(gsk_solver* sv, double* out);           /* nondimensional viscosity scale */
int gsk_solver_set_mu0(gsk_solver* sv, double mu0);        /* = 1 / (sqrt(2) delta_rp) for delta_rp cases */
int gsk_solver_set_accommodation(gsk_solver* sv, int32_t patch, double alpha);  /* Maxwell-wall patch */
/* discrete-equilibrium Newton tolerances of the cells and of the inflow states (defaults: 1e-11 FP64 /
 * 1e-6 FP32, and 1e-10); 0 iterates until the Newton step stops improving */
int gsk_solver_set_newton_tolerances(gsk_solver* sv, double cell_tol, double inflow_tol);
/* inflow / boundary-sum rounds per prepare ([kinetic].boundary_iterations); the pressure-inlet
 * reconstruction reads the face states of the previous round */
int gsk_solver_set_boundary_iterations(gsk_solver* sv, int64_t n);
/* Where each CIS iteration's prepare runs (no counterpart in the reference, whose
 * KineticSolver::prepare runs on the host, src/kinetic/kinetic_solver.cpp:365-452): on != 0 (the
 * default where the backend offers it) keeps the moments, the positivity check, eq. (59) epsilon,
 * closed-domain rescaling, pressure-face states and the boundary rounds on the backend's device and
 * exchanges only the boundary state and the iteration's scalars with the host; 0 runs them on the
 * host. While limiter branches are recorded or followed the host runs them either way.
 * gsk_solver_device_prepare: *on_device = 1 when the last prepare ran on the device. */
int gsk_solver_set_device_prepare(gsk_solver* sv, int32_t on);
int gsk_solver_device_prepare(gsk_solver* sv, int32_t* on_device);
/* on != 0: the next gsk_solver_residual_device records, at the current f, the branch every limiter
 * takes (the limited gradients' extremes and clamps, and the pressure-face reconstruction's), and
 * later evaluations follow that record, so that R is smooth around the recorded state and
 * differences of it give the Jacobian products of that branch; on = 0 restores the normal limiter.
 * Costs 4 bytes per cell and velocity of device memory while on. */
int gsk_solver_freeze_limiters(gsk_solver* sv, int32_t on);
/* Adjoint operators (no counterpart in the reference, which has no derivative capability; for
 * sensitivity solves, FP64 solvers only). gsk_solver_adjoint_prepare: at the current f, whose limiter
 * branches gsk_solver_freeze_limiters has recorded there, tabulates what the transposed Jacobian of
 * the steady residual needs (Jacobians of the boundary model's face maps and of the cell parameters);
 * closed domains and mass-flow inlets are refused. gsk_solver_adjoint_apply_device:
 * out = J^T lam + (d/df sum_k patch_weights[k] patch_mass_flow[k])^T at that state (n_patches
 * weights; lam, out device arrays [cell][storage position], lam may be NULL for 0); fails once the
 * state has changed since the prepare. Needs 8 doubles per cell and velocity of device workspace. */
int gsk_solver_adjoint_prepare(gsk_solver* sv);
int gsk_solver_adjoint_apply_device(gsk_solver* sv, const void* lam, const double* patch_weights, size_t n_patches,
                                    void* out, size_t n);
gsk_backend* gsk_solver_backend(gsk_solver* sv);

#ifdef __cplusplus
}
#endif
#endif /* GSIS_KINETIC_H */
