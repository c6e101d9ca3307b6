// ORACLE -- test infrastructure only. Nothing in the shipped library links, loads or calls this
// file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, to check results.
//
// A CPU restatement (C++17 + OpenMP) of the reference's CPU kinetic backend, the oracle the HIP
// backend is compared with. Paths are relative to the reference tree; each function names the
// lines it follows.
//   backend loops          src/kinetic/kinetic_backend_cpu.cpp:17-522
//   per-point kernels      src/kinetic/kinetic_kernels.hpp:36-609
//   discrete equilibrium   src/kinetic/discrete_equilibrium.hpp:25-188
//   Maxwellian / limiter   src/physics/maxwell.hpp:17-64, src/physics/limiter.hpp:17-50
//
// Parity pins (tests/): oracle/_ref (the reference's own header kernels compiled by g++ from the
// reference tree, oracle/Makefile) on identical inputs; closed forms (Maxwellian moments, limiter
// values, linear-field LSQ exactness); conservation invariants of the discrete equilibrium; the
// reference's converged slab fluxes (cases/verification/kinetic/results/slab_comparison.csv).
//
// Input is the same gsk_setup_view the HIP backend takes; it exports the gsk_backend_ops method
// set so the library's solver orchestration can drive it (gsk_solver_create_with_backend).
//
// Extension beyond the reference: the tanh-stretched velocity grid (gsk_setup_view.weight_rel /
// grid_nodes non-NULL). Every velocity sum then weighs term p by weight_rel[p], and the Maxwell-wall
// interpolation brackets the reflected velocity between stretched nodes. The reference has no such
// grid, so these paths are pinned by closed forms instead (tests/test_stretched_grid.py: Gaussian
// quadrature of the weights, Maxwellian moments to 1e-10, exact conservation of the discrete
// equilibrium, equilibrium preservation). The Maxwell-wall interpolation on a stretched grid is
// pinned through the kernels' reflected_value, compiled on the host and checked against the closed
// form M_eq(r) g(r) for f = M_eq g with g linear (tests/test_kernel_math.py); this restatement
// matches the kernels to 1e-10 over whole CIS iterations on a stretched-grid curved Maxwell tube. On the uniform
// grid both pointers are NULL and every path is the reference's, bit for bit.

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/gsis_kinetic.h"

namespace orc {

using i64 = long long;
constexpr double kPi = 3.14159265358979323846;

struct Err : std::runtime_error {
  using std::runtime_error::runtime_error;
};
inline void check(bool ok, const char* msg) {
  if (!ok) throw Err(msg);
}

// ---- Maxwellian (maxwell.hpp:17-38) ----
template <class F>
struct MaxwellParams {
  F ux, uy, uz, A, B;
};

template <class F>
inline F maxwellian(const MaxwellParams<F>& m, F x, F y, F z) {
  const F cx = x - m.ux, cy = y - m.uy, cz = z - m.uz;
  return m.A * std::exp(-m.B * (cx * cx + cy * cy + cz * cz));
}

// make_maxwell_params<double>(rho, u, T, R = 1), cast to F (kinetic_kernels.hpp:475-485)
template <class F>
inline MaxwellParams<F> params_from(double rho, double ux, double uy, double uz, double T) {
  const double theta = 1.0 * T;
  const double two_pi_theta = double(2.0 * kPi) * theta;
  MaxwellParams<F> m;
  m.ux = F(ux);
  m.uy = F(uy);
  m.uz = F(uz);
  m.A = F(rho / (two_pi_theta * std::sqrt(two_pi_theta)));
  m.B = F(1.0 / (2.0 * theta));
  return m;
}

// ---- Shakhov target (kinetic_types.hpp:68-90) ----
template <class F>
struct Shakhov {
  MaxwellParams<F> eq;
  F ucx, ucy, ucz, qx, qy, qz, theta, coef, g0, g4, gx, gy, gz;
};

template <class F>
inline F shakhov_poly(const Shakhov<F>& s, F cx, F cy, F cz, F c2) {
  return s.coef * (cx * s.qx + cy * s.qy + cz * s.qz) * (c2 / s.theta - F(5));
}

template <class F>
inline F shakhov_target(const Shakhov<F>& s, F x, F y, F z) {
  const F cx = x - s.ucx, cy = y - s.ucy, cz = z - s.ucz;
  const F c2 = cx * cx + cy * cy + cz * cz;
  const F m = maxwellian(s.eq, x, y, z);
  return m * (F(1) + shakhov_poly(s, cx, cy, cz, c2) + (s.g0 + (cx * s.gx + cy * s.gy + cz * s.gz) + F(0.5) * s.g4 * c2));
}

template <class F>
struct CellParams {
  Shakhov<F> s;
  F coll, diag;
};

// make_cell_params (kinetic_kernels.hpp:451-471)
template <class F>
inline CellParams<F> cell_params(const double* m, double V, double vdt, double mu0, double omega, double Pr,
                                 double cold_T) {
  const MaxwellParams<double> eq = params_from<double>(m[0], m[1], m[2], m[3], m[4]);
  const double theta = 1.0 * m[4];
  const double p = m[0] * theta;
  const double coef = (1.0 - Pr) / (5.0 * p * theta);
  CellParams<F> cp{};
  cp.s.eq = {F(eq.ux), F(eq.uy), F(eq.uz), F(eq.A), F(eq.B)};
  cp.s.ucx = cp.s.eq.ux;
  cp.s.ucy = cp.s.eq.uy;
  cp.s.ucz = cp.s.eq.uz;
  cp.s.qx = F(m[12]);
  cp.s.qy = F(m[13]);
  cp.s.qz = F(m[14]);
  cp.s.theta = F(theta);
  cp.s.coef = m[4] < cold_T ? F(0) : F(coef);
  const double mu = mu0 * std::pow(m[4], omega);
  const double coll = V * m[5] / mu;
  cp.coll = F(coll);
  cp.diag = F(coll + vdt);
  return cp;
}

// ---- Venkatakrishnan (limiter.hpp:17-50) ----
template <class F>
inline F venkat(F d2, F dmx, F dmn, F eps2) {
  F d1;
  if (d2 > F(0))
    d1 = dmx;
  else if (d2 < F(0))
    d1 = dmn;
  else
    return F(1);
  const F num = (d1 * d1 + eps2) + F(2) * d1 * d2;
  const F den = d1 * d1 + F(2) * d2 * d2 + d1 * d2 + eps2;
  if (!(den > F(0))) return F(0);
  const F r = num / den;
  return std::min(std::max(r, F(0)), F(1));
}

// ---- discrete equilibrium (discrete_equilibrium.hpp:25-188) ----
struct Grid {
  double lo[3], hi[3], d[3], criterion;
  int max_it;
  double tol;
};

inline bool resolved(const Grid& g, double ux, double uy, double uz, double T) {
  if (!(T > 0.0)) return false;
  const double u[3] = {ux, uy, uz};
  for (int a = 0; a < 3; ++a) {
    if (2.0 * kPi * kPi * T / (g.d[a] * g.d[a]) < g.criterion) return false;
    const double m1 = g.hi[a] - u[a], m2 = u[a] - g.lo[a];
    if (!(m1 > 0.0 && m2 > 0.0)) return false;
    if (m1 * m1 < 2.0 * T * g.criterion || m2 * m2 < 2.0 * T * g.criterion) return false;
  }
  return true;
}

template <class F>
inline void newton_terms(F e, F pm, F cx, F cy, F cz, F* t) {
  const F h = F(0.5) * (cx * cx + cy * cy + cz * cz);
  const F ex = e * cx, ey = e * cy, ez = e * cz, eh = e * h;
  t[0] = e;
  t[1] = ex;
  t[2] = ey;
  t[3] = ez;
  t[4] = eh;
  t[5] = ex * cx;
  t[6] = ex * cy;
  t[7] = ex * cz;
  t[8] = ex * h;
  t[9] = ey * cy;
  t[10] = ey * cz;
  t[11] = ey * h;
  t[12] = ez * cz;
  t[13] = ez * h;
  t[14] = eh * h;
  const F ep = e * pm;
  t[15] = ep;
  t[16] = ep * cx;
  t[17] = ep * cy;
  t[18] = ep * cz;
  t[19] = ep * h;
}

inline void jacobian(const double* s, double w, double J[5][5]) {
  int k = 0;
  for (int a = 0; a < 5; ++a)
    for (int b = a; b < 5; ++b, ++k) J[a][b] = J[b][a] = w * s[k];
}

inline bool gauss5(double A[5][5], double b[5], double x[5]) {
  for (int k = 0; k < 5; ++k) {
    int piv = k;
    double best = std::fabs(A[k][k]);
    for (int r = k + 1; r < 5; ++r)
      if (std::fabs(A[r][k]) > best) {
        best = std::fabs(A[r][k]);
        piv = r;
      }
    if (!(best > 0.0)) return false;
    if (piv != k) {
      for (int c = 0; c < 5; ++c) std::swap(A[k][c], A[piv][c]);
      std::swap(b[k], b[piv]);
    }
    for (int r = k + 1; r < 5; ++r) {
      const double f = A[r][k] / A[k][k];
      for (int c = k; c < 5; ++c) A[r][c] -= f * A[k][c];
      b[r] -= f * b[k];
    }
  }
  for (int k = 4; k >= 0; --k) {
    double sum = b[k];
    for (int c = k + 1; c < 5; ++c) sum -= A[k][c] * x[c];
    x[k] = sum / A[k][k];
    if (std::isnan(x[k])) return false;
  }
  return true;
}

template <class F>
inline MaxwellParams<F> beta_to_params(const double* beta, double ux, double uy, double uz) {
  const double B = -0.5 * beta[4];
  const double dx = -beta[1] / beta[4], dy = -beta[2] / beta[4], dz = -beta[3] / beta[4];
  return {F(ux + dx), F(uy + dy), F(uz + dz), F(std::exp(beta[0] + B * (dx * dx + dy * dy + dz * dz))), F(B)};
}

inline int newton_step(const double* s, double w, double rho, double T, double tol, double* beta, double& rel) {
  const double G[5] = {w * s[0] - rho, w * s[1], w * s[2], w * s[3], w * s[4] - 1.5 * rho * T};
  const double su = rho * std::sqrt(T);
  rel = std::max(std::fabs(G[0]) / rho,
                 std::max(std::max(std::fabs(G[1]), std::max(std::fabs(G[2]), std::fabs(G[3]))) / su,
                          std::fabs(G[4]) / (rho * T)));
  if (std::isnan(rel)) return -1;
  if (rel < tol) return 1;
  double J[5][5];
  jacobian(s, w, J);
  double rhs[5] = {-G[0], -G[1], -G[2], -G[3], -G[4]}, d[5] = {0, 0, 0, 0, 0};
  if (!gauss5(J, rhs, d)) return -1;
  double scale = 1.0;
  if (std::fabs(d[4]) > 0.5 * std::fabs(beta[4])) scale = 0.5 * std::fabs(beta[4]) / std::fabs(d[4]);
  if (std::fabs(d[0]) * scale > 1.0) scale = 1.0 / std::fabs(d[0]);
  for (int k = 0; k < 5; ++k) beta[k] += scale * d[k];
  return beta[4] < 0.0 ? 0 : -1;
}

inline bool accept(int r, double rel, double rel_prev, bool last) {
  return r == 1 || (r == 0 && rel < 1e-4 && (rel > 0.5 * rel_prev || last));
}

// discrete_maxwell_seq (kinetic_kernels.hpp:499-540)
template <class F>
int discrete_maxwell(const F* xx, const F* xy, const F* xz, const F* wq, int nv, double w, const Grid& g, double rho,
                     double ux, double uy, double uz, double T, const Shakhov<F>* poly, MaxwellParams<F>& out,
                     double* gamma) {
  for (int k = 0; k < 5; ++k) gamma[k] = 0.0;
  out = params_from<F>(rho, ux, uy, uz, T);
  if (resolved(g, ux, uy, uz, T)) return 0;
  const double two_pi_T = 2.0 * kPi * T;
  double beta[5] = {std::log(rho / (two_pi_T * std::sqrt(two_pi_T))), 0.0, 0.0, 0.0, -1.0 / T};
  const F fux = F(ux), fuy = F(uy), fuz = F(uz);
  double rel_prev = 1e300;
  for (int it = 0; it < g.max_it; ++it) {
    const MaxwellParams<F> m = beta_to_params<F>(beta, ux, uy, uz);
    double sums[20] = {};
    F t[20];
    for (int p = 0; p < nv; ++p) {
      const F cx = xx[p] - fux, cy = xy[p] - fuy, cz = xz[p] - fuz;
      F e = maxwellian(m, xx[p], xy[p], xz[p]);
      if (wq) e *= wq[p];  // stretched grid: quadrature weight w * wq[p]
      const F pm = poly ? shakhov_poly(*poly, cx, cy, cz, cx * cx + cy * cy + cz * cz) : F(0);
      newton_terms<F>(e, pm, cx, cy, cz, t);
      for (int k = 0; k < 20; ++k) sums[k] += double(t[k]);
    }
    double b[5] = {beta[0], beta[1], beta[2], beta[3], beta[4]};
    double rel = 0.0;
    const int r = newton_step(sums, w, rho, T, g.tol, b, rel);
    if (accept(r, rel, rel_prev, it + 1 == g.max_it)) {
      out = m;
      if (poly) {
        double J[5][5];
        jacobian(sums, w, J);
        double rhs[5] = {-w * sums[15], -w * sums[16], -w * sums[17], -w * sums[18], -w * sums[19]};
        if (!gauss5(J, rhs, gamma))
          for (int k = 0; k < 5; ++k) gamma[k] = 0.0;
      }
      return it + 1;
    }
    if (r < 0) break;
    std::copy(b, b + 5, beta);
    rel_prev = rel;
  }
  out = params_from<F>(rho, ux, uy, uz, T);
  for (int k = 0; k < 5; ++k) gamma[k] = 0.0;
  return -1;
}

// ---- tables ----
template <class F>
struct Entry {
  int other, bface;
  F s[3], ds[3], dj[3], c[3];
};
template <class F>
struct Face {
  int owner, kind, mirror, patch;
  F n[3], area, d[3];
};
template <class F>
struct Inflow {
  MaxwellParams<F> m;
  F scale;
};

template <class F>
class Backend {
 public:
  explicit Backend(const gsk_setup_view& v) {
    nc_ = v.n_cells;
    nv_ = v.nv;
    nb_ = v.n_bfaces;
    check(nc_ > 0 && nv_ > 0, "empty setup");
    xd_.assign(v.xi_x, v.xi_x + nv_);
    yd_.assign(v.xi_y, v.xi_y + nv_);
    zd_.assign(v.xi_z, v.xi_z + nv_);
    for (int p = 0; p < nv_; ++p) {
      xx_.push_back(F(xd_[p]));
      xy_.push_back(F(yd_[p]));
      xz_.push_back(F(zd_[p]));
    }
    w_ = v.weight;
    if (v.weight_rel != nullptr) {  // stretched velocity grid (an extension of the reference's uniform grid)
      wqd_.assign(v.weight_rel, v.weight_rel + nv_);
      for (int p = 0; p < nv_; ++p) wq_.push_back(F(wqd_[p]));
      check(v.grid_nodes != nullptr, "stretched grid without node coordinates");
      nodes_.assign(v.grid_nodes, v.grid_nodes + v.grid_n[0] + v.grid_n[1] + v.grid_n[2]);
    }
    off_.assign(v.csr_offsets, v.csr_offsets + nc_ + 1);
    for (int k = 0; k < off_[nc_]; ++k) {
      const gsk_csr_entry& e = v.csr[k];
      Entry<F> o;
      o.other = e.other;
      o.bface = e.bface;
      for (int a = 0; a < 3; ++a) {
        o.s[a] = F(e.s[a]);
        o.ds[a] = F(e.ds[a]);
        o.dj[a] = F(e.dother[a]);
        o.c[a] = F(e.c[a]);
      }
      csr_.push_back(o);
    }
    for (int c = 0; c < nc_; ++c) eps2_.push_back(F(v.eps2[c]));
    vol_.assign(v.volume, v.volume + nc_);
    vdt_.assign(v.vdt, v.vdt + nc_);
    for (int b = 0; b < nb_; ++b) {
      const gsk_bface& i = v.bfaces[b];
      Face<F> o;
      o.owner = i.owner;
      o.kind = i.kind;
      o.mirror = i.mirror_xor;
      o.patch = i.patch;
      for (int a = 0; a < 3; ++a) {
        o.n[a] = F(i.n[a]);
        o.d[a] = F(i.d[a]);
      }
      o.area = F(i.area);
      bf_.push_back(o);
    }
    inflow_.assign(nb_, Inflow<F>{});
    blocks_.assign(v.block_offsets, v.block_offsets + v.n_blocks + 1);
    ndir_ = v.n_sweep_dirs;
    if (ndir_ > 0) {
      order_.assign(v.sweep_order, v.sweep_order + static_cast<i64>(ndir_) * nc_);
      pos_.assign(v.sweep_pos, v.sweep_pos + static_cast<i64>(ndir_) * nc_);
      dir_of_.assign(v.sweep_dir_of, v.sweep_dir_of + nv_);
    }
    nmw_ = v.n_maxwell_faces;
    if (nmw_ > 0) {
      mw_off_.assign(v.maxwell_offset, v.maxwell_offset + nb_);
      mw_frac_.assign(v.maxwell_specular, v.maxwell_specular + nb_);
      spec_.assign(static_cast<std::size_t>(nmw_) * nv_, F(0));
      spec1_.assign(spec_.size(), F(0));
      g2s_.assign(v.grid_to_storage, v.grid_to_storage + nv_);
    }
    for (int a = 0; a < 3; ++a) {
      grid_n_[a] = v.grid_n[a];
      g_.lo[a] = v.deq_lo[a];
      g_.hi[a] = v.deq_hi[a];
      g_.d[a] = v.deq_d[a];
    }
    g_.criterion = v.deq_criterion;
    g_.max_it = v.deq_max_iterations;
    g_.tol = v.deq_tolerance;
    deq_ = v.discrete_eq != 0;
    coll_fallback_ = v.deq_fallback_collisionless != 0;
    limiter_ = v.limiter_mode;
    positivity_ = v.positivity;
    diag_scale_ = F(v.diag_scale);
    beta_ = F(v.interface_damping);
    cold_ = v.cold_bgk_temperature;
    mu0_ = v.mu0;
    omega_ = v.omega;
    Pr_ = v.Pr;
    passes_ = v.block_passes;
    mode_ = v.negative_check;
    long long k = v.velocity_chunk > 0 ? v.velocity_chunk : nv_;
    k = (k + v.orbit_size - 1) / v.orbit_size * v.orbit_size;
    chunk_ = static_cast<int>(std::min<long long>(k, nv_));
    cells_.assign(nc_, CellParams<F>{});
    f_.assign(static_cast<std::size_t>(nc_) * nv_, F(0));
    mom_.assign(static_cast<std::size_t>(nc_) * 15, 0.0);
    shift_.assign(static_cast<std::size_t>(nc_) * 3, 0.0);
    const std::size_t work = static_cast<std::size_t>(nc_) * chunk_;
    gx_.assign(work, F(0));
    gy_.assign(work, F(0));
    gz_.assign(work, F(0));
    r_.assign(work, F(0));
  }

  int chunk() const { return chunk_; }

  // set_equilibrium (kinetic_backend_cpu.cpp:64-80)
  void set_equilibrium(const double* st, std::size_t n) {
    check(n == static_cast<std::size_t>(nc_) * 5, "states length");
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc_; ++c) {
      const double* s = st + static_cast<i64>(c) * 5;
      const MaxwellParams<F> m = equilibrium(s[0], s[1], s[2], s[3], s[4], nullptr);
      for (int p = 0; p < nv_; ++p) f_[static_cast<i64>(c) * nv_ + p] = maxwellian(m, xx_[p], xy_[p], xz_[p]);
      shift_[c * 3 + 0] = s[1];
      shift_[c * 3 + 1] = s[2];
      shift_[c * 3 + 2] = s[3];
    }
  }
  void set_distribution(const double* f, std::size_t n) {
    check(n == f_.size(), "distribution length");
    for (std::size_t i = 0; i < n; ++i) f_[i] = F(f[i]);
  }
  void get_distribution(double* f, std::size_t n) {
    check(n == f_.size(), "distribution length");
    for (std::size_t i = 0; i < n; ++i) f[i] = double(f_[i]);
  }

  // compute_moments (kinetic_backend_cpu.cpp:92-120; kinetic_kernels.hpp:382-442)
  void compute_moments(double* out, std::size_t n) {
    check(n == mom_.size(), "moments length");
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc_; ++c) {
      const double u0x = shift_[c * 3], u0y = shift_[c * 3 + 1], u0z = shift_[c * 3 + 2];
      double s[13] = {};
      for (int p = 0; p < nv_; ++p) {
        double fv = double(f_[static_cast<i64>(c) * nv_ + p]);
        if (!wqd_.empty()) fv *= wqd_[p];
        const double cx = xd_[p] - u0x, cy = yd_[p] - u0y, cz = zd_[p] - u0z;
        const double f1 = fv * cx, f2 = fv * cy, f3 = fv * cz;
        const double fc2 = fv * (cx * cx + cy * cy + cz * cz);
        const double t[13] = {fv, f1, f2, f3, f1 * cx, f2 * cy, f3 * cz, f1 * cy, f2 * cz, f1 * cz, fc2 * cx, fc2 * cy, fc2 * cz};
        for (int k = 0; k < 13; ++k) s[k] += t[k];
      }
      double* m = &mom_[static_cast<std::size_t>(c) * 15];
      finalize(s, w_, u0x, u0y, u0z, m);
      if (m[0] > 0.0) {
        shift_[c * 3 + 0] = m[1];
        shift_[c * 3 + 1] = m[2];
        shift_[c * 3 + 2] = m[3];
      }
    }
    std::copy(mom_.begin(), mom_.end(), out);
  }

  void scale_distribution(double s) {
    const F sf = F(s);
    for (F& x : f_) x *= sf;
    for (int c = 0; c < nc_; ++c) {
      double* m = &mom_[static_cast<std::size_t>(c) * 15];
      m[0] *= s;
      for (int k = 5; k < 15; ++k) m[k] *= s;
    }
  }

  // update_cell_params (kinetic_backend_cpu.cpp:134-160)
  void update_cell_params() {
    flags_.assign(nc_, 0);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc_; ++c) {
      const double* m = &mom_[static_cast<std::size_t>(c) * 15];
      CellParams<F> cp = cell_params<F>(m, vol_[c], vdt_[c], mu0_, omega_, Pr_, cold_);
      if (deq_) {
        MaxwellParams<F> eq;
        double gam[5];
        const int r = discrete_maxwell<F>(xx_.data(), xy_.data(), xz_.data(), wq(), nv_, w_, g_, m[0], m[1], m[2], m[3],
                                          m[4], &cp.s, eq, gam);
        if (r > 0) {
          cp.s.eq = eq;
          cp.s.g0 = F(gam[0]);
          cp.s.gx = F(gam[1]);
          cp.s.gy = F(gam[2]);
          cp.s.gz = F(gam[3]);
          cp.s.g4 = F(gam[4]);
        } else if (r < 0 && coll_fallback_) {
          cp.coll = F(0);
          cp.diag = F(vdt_[c]);
        }
        flags_[c] = r;
      }
      cells_[c] = cp;
    }
  }

  gsk_deq_counts deq_counts() const {
    gsk_deq_counts o{0, 0, 0};
    for (int x : flags_) {
      if (x > 0) ++o.corrected;
      if (x < 0) ++o.failed;
      o.max_iterations = std::max(o.max_iterations, x);
    }
    return o;
  }

  void set_inflow(const gsk_inflow_state* in, std::size_t n) {
    check(n == static_cast<std::size_t>(nb_), "inflow length");
    for (std::size_t b = 0; b < n; ++b) {
      inflow_[b].m = params_from<F>(in[b].rho, in[b].ux, in[b].uy, in[b].uz, in[b].T);
      inflow_[b].scale = F(in[b].scale);
    }
  }

  // boundary_sums (kinetic_backend_cpu.cpp:184-240)
  void boundary_sums(double* out, std::size_t n) {
    check(n == static_cast<std::size_t>(nb_) * 20, "boundary sums length");
    if (nmw_ > 0) {
#pragma omp parallel for schedule(dynamic, 1)
      for (int b = 0; b < nb_; ++b) {
        if (mw_off_[b] < 0) continue;
        for (int p = 0; p < nv_; ++p)
          spec1_[mw_off_[b] + p] = dotn(bf_[b], p) < F(0) ? reflected<false>(bf_[b], p) : F(0);
      }
#pragma omp parallel for schedule(dynamic, 1)
      for (int b = 0; b < nb_; ++b) {
        if (mw_off_[b] < 0) continue;
        for (int p = 0; p < nv_; ++p)
          spec_[mw_off_[b] + p] = dotn(bf_[b], p) < F(0) ? reflected<true>(bf_[b], p) : F(0);
      }
    }
    const F w = F(w_);
#pragma omp parallel for schedule(dynamic, 4)
    for (int b = 0; b < nb_; ++b) {
      const Face<F>& f = bf_[b];
      double acc[20] = {};
      for (int p = 0; p < nv_; ++p) {
        const F x = xx_[p], y = xy_[p], z = xz_[p];
        const F an = x * f.n[0] + y * f.n[1] + z * f.n[2];
        F fb;
        int o;
        if (an >= F(0) || f.kind == GSK_BC_PERIODIC) {
          fb = reconstruct(f, p);
          o = 0;
        } else {
          if (f.kind == GSK_BC_SYMMETRY)
            fb = reconstruct(f, p ^ f.mirror);
          else if (f.kind == GSK_BC_MAXWELL_WALL)
            fb = maxwell_inflow(b, p);
          else
            fb = maxwellian(inflow_[b].m, x, y, z);
          o = 5;
        }
        if (!wq_.empty()) fb *= wq_[p];
        const F wf = w * fb;
        const F e = F(0.5) * (x * x + y * y + z * z);
        const F af = an * wf;
        acc[o + 0] += double(af);
        acc[o + 1] += double(af * x);
        acc[o + 2] += double(af * y);
        acc[o + 3] += double(af * z);
        acc[o + 4] += double(af * e);
        acc[o + 10] += double(wf);
        acc[o + 11] += double(wf * x);
        acc[o + 12] += double(wf * y);
        acc[o + 13] += double(wf * z);
        acc[o + 14] += double(wf * e);
      }
      std::copy(acc, acc + 20, out + static_cast<i64>(b) * 20);
    }
  }

  // internal_face_mass_flux (kinetic_backend_cpu.cpp:242-260)
  void face_flux(const int* cells, const int* entries, std::size_t n, double* out) {
#pragma omp parallel for schedule(dynamic, 4)
    for (i64 i = 0; i < static_cast<i64>(n); ++i) {
      const int c = cells[i];
      const Entry<F>& e = csr_[entries[i]];
      double acc = 0.0;
      for (int p = 0; p < nv_; ++p) {
        const F x = xx_[p], y = xy_[p], z = xz_[p];
        const F a = x * e.s[0] + y * e.s[1] + z * e.s[2];
        F ff;
        if (a >= F(0)) {
          const F fi = fat(c, p);
          F g[3];
          gradient(c, p, fi, g);
          ff = fi + g[0] * e.ds[0] + g[1] * e.ds[1] + g[2] * e.ds[2];
        } else {
          const F fj = fat(e.other, p);
          F g[3];
          gradient(e.other, p, fj, g);
          ff = fj + g[0] * e.dj[0] + g[1] * e.dj[1] + g[2] * e.dj[2];
        }
        acc += wqd_.empty() ? double(a * ff) : double(a * ff) * wqd_[p];
      }
      out[i] = double(F(w_)) * acc;
    }
  }

  // sweep (kinetic_backend_cpu.cpp:262-344)
  void sweep(double* rc, std::size_t n) {
    check(n == static_cast<std::size_t>(nc_), "residual length");
    std::fill(rc, rc + n, 0.0);
    const int nblk = static_cast<int>(blocks_.size()) - 1;
    for (int k0 = 0; k0 < nv_; k0 += chunk_) {
      const int kc = std::min(chunk_, nv_ - k0);
#pragma omp parallel for schedule(static)
      for (int c = 0; c < nc_; ++c)
        for (int t = 0; t < kc; ++t) {
          F g[3];
          gradient(c, k0 + t, fat(c, k0 + t), g);
          const i64 wi = static_cast<i64>(c) * chunk_ + t;
          gx_[wi] = g[0];
          gy_[wi] = g[1];
          gz_[wi] = g[2];
        }
#pragma omp parallel for schedule(static)
      for (int c = 0; c < nc_; ++c) {
        double r2 = 0.0;
        for (int t = 0; t < kc; ++t) {
          const F R = residual(k0, c, t);
          r_[static_cast<i64>(c) * chunk_ + t] = R;
          r2 += wqd_.empty() ? double(R) * double(R) : double(R) * double(R) * wqd_[k0 + t];
        }
        rc[c] += r2;
      }
      // block passes: gx <- R, gy <- previous dF (the reference reuses the gradient arrays)
      const F* r0 = nullptr;
      const F* dprev = nullptr;
      if (passes_ > 1) gx_ = r_;
      for (int pass = 0; pass < passes_; ++pass) {
        if (pass > 0) {
          gy_ = r_;
          r0 = gx_.data();
          dprev = gy_.data();
        }
        const bool update = pass + 1 == passes_;
#pragma omp parallel for schedule(static)
        for (int lane = 0; lane < nblk * kc; ++lane) {
          const int b = lane / kc, t = lane % kc;
          const int c0 = blocks_[b], c1 = blocks_[b + 1];
          const int* ord = nullptr;
          const int* pos = nullptr;
          direction(k0 + t, ord, pos);
          for (int k = c0; k < c1; ++k) forward_cell(ord ? ord[k] : k, k0, t, c0, c1, pos, r0, dprev);
        }
#pragma omp parallel for schedule(static)
        for (int lane = 0; lane < nblk * kc; ++lane) {
          const int b = lane / kc, t = lane % kc;
          const int c0 = blocks_[b], c1 = blocks_[b + 1];
          const int* ord = nullptr;
          const int* pos = nullptr;
          direction(k0 + t, ord, pos);
          for (int k = c1 - 1; k >= c0; --k) backward_cell(ord ? ord[k] : k, k0, t, c0, c1, pos, dprev, update);
        }
      }
    }
  }

  // gsis_correction (kinetic_backend_cpu.cpp:346-432)
  long long gsis_correction(const double* wh_all, const double* wn_all, std::size_t nw, int* flags, std::size_t nf) {
    check(nw == static_cast<std::size_t>(nc_) * 5, "correction states length");
    long long replaced = 0;
#pragma omp parallel for schedule(static) reduction(+ : replaced)
    for (int c = 0; c < nc_; ++c) {
      const double* wh = wh_all + static_cast<i64>(c) * 5;
      const double* wn = wn_all + static_cast<i64>(c) * 5;
      int rh = 0, rn = 0;
      const MaxwellParams<F> mh = equilibrium(wh[0], wh[1], wh[2], wh[3], wh[4], &rh);
      const MaxwellParams<F> mn = equilibrium(wn[0], wn[1], wn[2], wn[3], wn[4], &rn);
      if (coll_fallback_ && (rh < 0 || rn < 0)) {
        ++replaced;
        if (nf) flags[c] = 2;
        for (int a = 0; a < 3; ++a) shift_[c * 3 + a] = wh[1 + a];
        continue;
      }
      F* fc = &f_[static_cast<i64>(c) * nv_];
      const double thr = -1e-8 * wn[0] / (wn[4] * std::sqrt(wn[4]));
      bool rep = false;
      auto dcorr = [&](int p) { return maxwellian(mn, xx_[p], xy_[p], xz_[p]) - maxwellian(mh, xx_[p], xy_[p], xz_[p]); };
      if (mode_ == 2) {  // limit
        F alpha = F(1);
        for (int p = 0; p < nv_; ++p) {
          const F d = dcorr(p);
          F lim = F(1);
          if (d < F(0)) lim = std::min(F(1), (fc[p] - (std::min(fc[p], F(0)) + F(thr))) / (-d));
          alpha = std::min(alpha, lim);
        }
        for (int p = 0; p < nv_; ++p) fc[p] += alpha * dcorr(p);
        rep = alpha < F(1);
      } else if (mode_ == 3) {  // blend
        F theta = F(0);
        for (int p = 0; p < nv_; ++p) {
          const F e = maxwellian(mn, xx_[p], xy_[p], xz_[p]);
          const F g = fc[p] + (e - maxwellian(mh, xx_[p], xy_[p], xz_[p]));
          const F bound = std::min(fc[p], F(0)) + F(thr);
          F th = F(0);
          if (g < bound) {
            const F den = e - g;
            th = den > F(0) ? std::min(F(1), (bound - g) / den) : F(1);
          }
          theta = std::max(theta, th);
        }
        for (int p = 0; p < nv_; ++p) {
          const F e = maxwellian(mn, xx_[p], xy_[p], xz_[p]);
          const F g = fc[p] + (e - maxwellian(mh, xx_[p], xy_[p], xz_[p]));
          fc[p] = (F(1) - theta) * g + theta * e;
        }
        rep = theta > F(0);
      } else {  // all / created
        F fmin = F(0);
        for (int p = 0; p < nv_; ++p) {
          const F old = fc[p];
          const F x = old + dcorr(p);
          fc[p] = x;
          const F chk = mode_ == 1 ? x - std::min(old, F(0)) : x;
          fmin = p == 0 ? chk : std::min(fmin, chk);
        }
        rep = double(fmin) < thr;
        if (rep)
          for (int p = 0; p < nv_; ++p) fc[p] = maxwellian(mn, xx_[p], xy_[p], xz_[p]);
      }
      if (rep) ++replaced;
      if (nf) flags[c] = rep ? 1 : 0;
      for (int a = 0; a < 3; ++a) shift_[c * 3 + a] = wn[1 + a];
    }
    return replaced;
  }

 private:
  static void finalize(const double* s, double w, double u0x, double u0y, double u0z, double* out) {
    const double s0 = s[0];
    if (!(s0 > 0.0)) {
      out[0] = w * s0;
      for (int k = 1; k < 15; ++k) out[k] = 0.0;
      return;
    }
    const double dx = s[1] / s0, dy = s[2] / s0, dz = s[3] / s0;
    const double rho = w * s0;
    const double pxx = w * (s[4] - s0 * dx * dx), pyy = w * (s[5] - s0 * dy * dy), pzz = w * (s[6] - s0 * dz * dz);
    const double pxy = w * (s[7] - s0 * dx * dy), pyz = w * (s[8] - s0 * dy * dz), pxz = w * (s[9] - s0 * dx * dz);
    const double p = (pxx + pyy + pzz) / 3.0;
    const double tr2 = s[4] + s[5] + s[6];
    const double du2 = dx * dx + dy * dy + dz * dz;
    const double sx = s[4] * dx + s[7] * dy + s[9] * dz;
    const double sy = s[7] * dx + s[5] * dy + s[8] * dz;
    const double sz = s[9] * dx + s[8] * dy + s[6] * dz;
    const double o[15] = {rho, u0x + dx, u0y + dy, u0z + dz, p / rho, p, pxx - p, pyy - p, pzz - p, pxy, pyz, pxz,
                          0.5 * w * (s[10] - dx * tr2 - 2.0 * sx + 2.0 * s0 * du2 * dx),
                          0.5 * w * (s[11] - dy * tr2 - 2.0 * sy + 2.0 * s0 * du2 * dy),
                          0.5 * w * (s[12] - dz * tr2 - 2.0 * sz + 2.0 * s0 * du2 * dz)};
    std::copy(o, o + 15, out);
  }

  MaxwellParams<F> equilibrium(double rho, double ux, double uy, double uz, double T, int* result) const {
    if (result) *result = 0;
    if (!deq_) return params_from<F>(rho, ux, uy, uz, T);
    MaxwellParams<F> m;
    double gam[5];
    const int r = discrete_maxwell<F>(xx_.data(), xy_.data(), xz_.data(), wq(), nv_, w_, g_, rho, ux, uy, uz, T, nullptr, m,
                                      gam);
    if (result) *result = r;
    return m;
  }

  F fat(int c, int p) const { return f_[static_cast<i64>(c) * nv_ + p]; }
  const F* wq() const { return wq_.empty() ? nullptr : wq_.data(); }
  F dotn(const Face<F>& f, int p) const { return xx_[p] * f.n[0] + xy_[p] * f.n[1] + xz_[p] * f.n[2]; }
  F inflow_at(int b, int p) const { return inflow_[b].scale * maxwellian(inflow_[b].m, xx_[p], xy_[p], xz_[p]); }
  F maxwell_inflow(int b, int p) const { return inflow_at(b, p) + F(mw_frac_[b]) * spec_[mw_off_[b] + p]; }

  // boundary_value_first_order (kinetic_kernels.hpp:145-162)
  F boundary_first_order(const Entry<F>& e, int cell, int p, F fi) const {
    const Face<F>& b = bf_[e.bface];
    if (b.kind == GSK_BC_PERIODIC) return fi;
    if (b.kind == GSK_BC_SYMMETRY) return F(0.5) * (fi + fat(cell, p ^ b.mirror));
    const F a = xx_[p] * e.s[0] + xy_[p] * e.s[1] + xz_[p] * e.s[2];
    if (a >= F(0)) return fi;
    if (b.kind == GSK_BC_MAXWELL_WALL) return inflow_at(e.bface, p) + F(mw_frac_[e.bface]) * spec1_[mw_off_[e.bface] + p];
    return inflow_at(e.bface, p);
  }

  // limited_gradient (kinetic_kernels.hpp:165-201)
  void gradient(int cell, int p, F fi, F* out) const {
    F g[3] = {F(0), F(0), F(0)};
    F fmx = fi, fmn = fi;
    for (int k = off_[cell]; k < off_[cell + 1]; ++k) {
      const Entry<F>& e = csr_[k];
      const F fk = e.other >= 0 ? fat(e.other, p) : boundary_first_order(e, cell, p, fi);
      const F d = fk - fi;
      for (int a = 0; a < 3; ++a) g[a] += e.c[a] * d;
      fmx = std::max(fmx, fk);
      fmn = std::min(fmn, fk);
    }
    if (limiter_ == 0 && positivity_ == 0) {
      std::copy(g, g + 3, out);
      return;
    }
    if (positivity_ != 0 && !(fi > F(0))) {
      out[0] = out[1] = out[2] = F(0);
      return;
    }
    F eps2 = eps2_[cell];
    if (limiter_ == 2) {
      const F s = std::max(std::fabs(fmx), std::fabs(fmn));
      eps2 *= s * s;
    }
    F phi = F(1), proj_min = F(0);
    for (int k = off_[cell]; k < off_[cell + 1]; ++k) {
      const Entry<F>& e = csr_[k];
      if (e.other < 0 && bf_[e.bface].kind == GSK_BC_PERIODIC) continue;
      const F proj = g[0] * e.ds[0] + g[1] * e.ds[1] + g[2] * e.ds[2];
      if (limiter_ != 0) phi = std::min(phi, venkat(proj, fmx - fi, fmn - fi, eps2));
      proj_min = std::min(proj_min, proj);
    }
    if (positivity_ != 0 && proj_min < F(0)) phi = fi > F(0) ? std::min(phi, fi / (-proj_min)) : F(0);
    for (int a = 0; a < 3; ++a) out[a] = phi * g[a];
  }

  // reconstruct_at_boundary (kinetic_kernels.hpp:204-211)
  F reconstruct(const Face<F>& b, int p) const {
    const F fi = fat(b.owner, p);
    if (b.kind == GSK_BC_PERIODIC) return fi;
    F g[3];
    gradient(b.owner, p, fi, g);
    return fi + g[0] * b.d[0] + g[1] * b.d[1] + g[2] * b.d[2];
  }

  // reflected_value (kinetic_kernels.hpp:80-127)
  template <bool Rec>
  F reflected(const Face<F>& b, int p) const {
    if (b.mirror != 0) {
      const int q = p ^ b.mirror;
      return Rec ? reconstruct(b, q) : fat(b.owner, q);
    }
    const F x = xx_[p], y = xy_[p], z = xz_[p];
    const F un = x * b.n[0] + y * b.n[1] + z * b.n[2];
    const F r[3] = {x - F(2) * un * b.n[0], y - F(2) * un * b.n[1], z - F(2) * un * b.n[2]};
    int base[3];
    F t[3];
    int node_off = 0;
    for (int a = 0; a < 3; ++a) {
      if (!nodes_.empty()) {  // stretched grid: bracketing nodes, linear in xi between them
        const double* nd = nodes_.data() + node_off;
        const int n = grid_n_[a];
        node_off += n;
        const double ra = double(r[a]);
        if (ra < nd[0] - 1e-5 * (nd[1] - nd[0]) || ra > nd[n - 1] + 1e-5 * (nd[n - 1] - nd[n - 2])) return F(0);
        const int lo = std::max(0, std::min(n - 2, static_cast<int>(std::upper_bound(nd, nd + n, ra) - nd) - 1));
        base[a] = lo;
        t[a] = F(std::max(0.0, std::min(1.0, (ra - nd[lo]) / (nd[lo + 1] - nd[lo]))));
        continue;
      }
      const double zz = (double(r[a]) - g_.lo[a]) / g_.d[a] - 0.5;
      if (zz < -1e-5 || zz > grid_n_[a] - 1.0 + 1e-5) return F(0);
      const double cl = std::max(0.0, std::min(zz, grid_n_[a] - 1.0));
      base[a] = static_cast<int>(cl);
      if (base[a] >= grid_n_[a] - 1) base[a] = grid_n_[a] - 2;
      t[a] = F(cl - base[a]);
    }
    const MaxwellParams<F>& ref = cells_[b.owner].s.eq;
    F result = F(0);
    for (int k = 0; k < 8; ++k) {
      const int ix = base[0] + (k & 1), iy = base[1] + ((k >> 1) & 1), iz = base[2] + ((k >> 2) & 1);
      const F w = ((k & 1) ? t[0] : F(1) - t[0]) * ((k & 2) ? t[1] : F(1) - t[1]) * ((k & 4) ? t[2] : F(1) - t[2]);
      if (w == F(0)) continue;
      const int q = g2s_[ix + grid_n_[0] * (iy + grid_n_[1] * iz)];
      const F value = Rec ? reconstruct(b, q) : fat(b.owner, q);
      const double drx = double(r[0]) - ref.ux, dry = double(r[1]) - ref.uy, drz = double(r[2]) - ref.uz;
      const double dqx = double(xx_[q]) - ref.ux, dqy = double(xy_[q]) - ref.uy, dqz = double(xz_[q]) - ref.uz;
      const double ratio = std::exp(-double(ref.B) * (drx * drx + dry * dry + drz * drz - dqx * dqx - dqy * dqy - dqz * dqz));
      result += w * F(double(value) * ratio);
    }
    return result;
  }

  // residual_value (kinetic_kernels.hpp:238-278)
  F residual(int k0, int cell, int t) const {
    const int p = k0 + t;
    const F x = xx_[p], y = xy_[p], z = xz_[p];
    const F fi = fat(cell, p);
    const i64 wi = static_cast<i64>(cell) * chunk_ + t;
    F R = F(0);
    for (int k = off_[cell]; k < off_[cell + 1]; ++k) {
      const Entry<F>& e = csr_[k];
      int kind = -1;
      if (e.other < 0) {
        kind = bf_[e.bface].kind;
        if (kind == GSK_BC_PERIODIC) continue;
      }
      const F a = x * e.s[0] + y * e.s[1] + z * e.s[2];
      F ff;
      if (a >= F(0)) {
        ff = fi + gx_[wi] * e.ds[0] + gy_[wi] * e.ds[1] + gz_[wi] * e.ds[2];
      } else if (e.other >= 0) {
        const i64 wj = static_cast<i64>(e.other) * chunk_ + t;
        ff = fat(e.other, p) + gx_[wj] * e.dj[0] + gy_[wj] * e.dj[1] + gz_[wj] * e.dj[2];
      } else if (kind == GSK_BC_SYMMETRY) {
        const int tm = t ^ bf_[e.bface].mirror;
        const i64 wm = static_cast<i64>(cell) * chunk_ + tm;
        ff = fat(cell, k0 + tm) + gx_[wm] * e.ds[0] + gy_[wm] * e.ds[1] + gz_[wm] * e.ds[2];
      } else if (kind == GSK_BC_MAXWELL_WALL) {
        ff = maxwell_inflow(e.bface, p);
      } else {
        ff = inflow_at(e.bface, p);
      }
      R += a * ff;
    }
    const CellParams<F>& cp = cells_[cell];
    R -= cp.coll * (shakhov_target(cp.s, x, y, z) - fi);
    return R;
  }

  void direction(int p, const int*& ord, const int*& pos) const {
    if (ndir_ > 0) {
      const i64 o = static_cast<i64>(dir_of_[p]) * nc_;
      ord = order_.data() + o;
      pos = pos_.data() + o;
    }
  }

  // sweep_forward_cell (kinetic_kernels.hpp:290-324)
  void forward_cell(int cell, int k0, int t, int c0, int c1, const int* pos, const F* r0, const F* dprev) {
    const int p = k0 + t;
    const F x = xx_[p], y = xy_[p], z = xz_[p];
    F ap = F(0), amo = F(0), L = F(0), Lo = F(0);
    const int pi = pos ? pos[cell] : cell;
    for (int k = off_[cell]; k < off_[cell + 1]; ++k) {
      const Entry<F>& e = csr_[k];
      if (e.other < 0 && bf_[e.bface].kind == GSK_BC_PERIODIC) continue;
      const F a = x * e.s[0] + y * e.s[1] + z * e.s[2];
      if (a > F(0)) {
        ap += a;
      } else if (e.other >= 0) {
        if (e.other < c0 || e.other >= c1) {
          if (dprev)
            Lo += a * dprev[static_cast<i64>(e.other) * chunk_ + t];
          else
            amo -= a;
        } else if ((pos ? pos[e.other] : e.other) < pi) {
          L += a * r_[static_cast<i64>(e.other) * chunk_ + t];
        }
      }
    }
    const F D = cells_[cell].diag + diag_scale_ * ap + beta_ * amo;
    const i64 wi = static_cast<i64>(cell) * chunk_ + t;
    const F Ri = r0 ? r0[wi] : r_[wi];
    r_[wi] = (-Ri - Lo - L) / D;
  }

  // sweep_backward_cell (kinetic_kernels.hpp:328-358)
  void backward_cell(int cell, int k0, int t, int c0, int c1, const int* pos, const F* dprev, bool update) {
    const int p = k0 + t;
    const F x = xx_[p], y = xy_[p], z = xz_[p];
    F ap = F(0), amo = F(0), U = F(0);
    const int pi = pos ? pos[cell] : cell;
    for (int k = off_[cell]; k < off_[cell + 1]; ++k) {
      const Entry<F>& e = csr_[k];
      if (e.other < 0 && bf_[e.bface].kind == GSK_BC_PERIODIC) continue;
      const F a = x * e.s[0] + y * e.s[1] + z * e.s[2];
      if (a > F(0)) {
        ap += a;
      } else if (e.other >= 0) {
        if (e.other < c0 || e.other >= c1) {
          if (!dprev) amo -= a;
        } else if ((pos ? pos[e.other] : e.other) > pi) {
          U += a * r_[static_cast<i64>(e.other) * chunk_ + t];
        }
      }
    }
    const F D = cells_[cell].diag + diag_scale_ * ap + beta_ * amo;
    const i64 wi = static_cast<i64>(cell) * chunk_ + t;
    const F df = r_[wi] - U / D;
    r_[wi] = df;
    if (update) f_[static_cast<i64>(cell) * nv_ + p] += df;
  }

  int nc_ = 0, nv_ = 0, nb_ = 0, chunk_ = 0, ndir_ = 0, nmw_ = 0, passes_ = 1, mode_ = 0;
  int limiter_ = 1, positivity_ = 0;
  int grid_n_[3] = {0, 0, 0};
  bool deq_ = false, coll_fallback_ = false;
  double w_ = 0, cold_ = 0, mu0_ = 0, omega_ = 0, Pr_ = 0;
  F diag_scale_ = F(1), beta_ = F(0);
  Grid g_{};
  std::vector<double> xd_, yd_, zd_, vol_, vdt_, mom_, shift_, mw_frac_;
  std::vector<F> xx_, xy_, xz_, eps2_, f_, gx_, gy_, gz_, r_, spec_, spec1_;
  std::vector<int> off_, blocks_, order_, pos_, g2s_, flags_;
  std::vector<unsigned char> dir_of_;
  std::vector<long long> mw_off_;
  std::vector<Entry<F>> csr_;
  std::vector<Face<F>> bf_;
  std::vector<Inflow<F>> inflow_;
  std::vector<CellParams<F>> cells_;
  std::vector<F> wq_;                 // stretched grid: relative weights (empty: uniform)
  std::vector<double> wqd_, nodes_;  // the same in FP64; node coordinates of axes x, y, z
};

// ---- C ABI: one handle = either precision ----
struct Handle {
  std::unique_ptr<Backend<double>> d;
  std::unique_ptr<Backend<float>> s;
  std::string err;
  double t_sweep = 0.0;
};

template <class Fn>
int call(void* ctx, Fn&& fn) {
  Handle* h = static_cast<Handle*>(ctx);
  try {
    if (h->d)
      fn(*h->d);
    else
      fn(*h->s);
    return GSK_OK;
  } catch (const std::exception& e) {
    h->err = e.what();
    return GSK_ERR_INVALID;
  }
}

}  // namespace orc

extern "C" {

static thread_local std::string g_orc_error;

const char* orc_last_error(void) { return g_orc_error.c_str(); }

// precision: 1 = FP64, 0 = FP32 storage (as view->precision_fp64 when -1)
void* orc_create(const gsk_setup_view* view, int precision) {
  try {
    auto h = std::make_unique<orc::Handle>();
    const int fp64 = precision < 0 ? view->precision_fp64 : precision;
    if (fp64)
      h->d = std::make_unique<orc::Backend<double>>(*view);
    else
      h->s = std::make_unique<orc::Backend<float>>(*view);
    return h.release();
  } catch (const std::exception& e) {
    g_orc_error = e.what();
    return nullptr;
  }
}

void orc_destroy(void* h) { delete static_cast<orc::Handle*>(h); }

int orc_set_num_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

static const char* o_name(void* ctx) { return static_cast<orc::Handle*>(ctx)->d ? "oracle-fp64" : "oracle-fp32"; }
static int o_chunk(void* ctx, int32_t* out) {
  return orc::call(ctx, [&](auto& b) { *out = b.chunk(); });
}
static int o_mem(void* ctx, uint64_t* out) {
  *out = 0;
  return GSK_OK;
}
static int o_set_eq(void* ctx, const double* s, size_t n) {
  return orc::call(ctx, [&](auto& b) { b.set_equilibrium(s, n); });
}
static int o_set_f(void* ctx, const double* f, size_t n) {
  return orc::call(ctx, [&](auto& b) { b.set_distribution(f, n); });
}
static int o_get_f(void* ctx, double* f, size_t n) {
  return orc::call(ctx, [&](auto& b) { b.get_distribution(f, n); });
}
static int o_moments(void* ctx, double* out, size_t n) {
  return orc::call(ctx, [&](auto& b) { b.compute_moments(out, n); });
}
static int o_scale(void* ctx, double s) {
  return orc::call(ctx, [&](auto& b) { b.scale_distribution(s); });
}
static int o_params(void* ctx) {
  return orc::call(ctx, [&](auto& b) { b.update_cell_params(); });
}
static int o_inflow(void* ctx, const gsk_inflow_state* in, size_t n) {
  return orc::call(ctx, [&](auto& b) { b.set_inflow(in, n); });
}
static int o_bsums(void* ctx, double* out, size_t n) {
  return orc::call(ctx, [&](auto& b) { b.boundary_sums(out, n); });
}
static int o_sweep(void* ctx, double* r, size_t n) {
  const auto t0 = std::chrono::steady_clock::now();
  const int st = orc::call(ctx, [&](auto& b) { b.sweep(r, n); });
  static_cast<orc::Handle*>(ctx)->t_sweep += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}
static int o_gsis(void* ctx, const double* wh, const double* wn, size_t nw, int32_t* flags, size_t nf, long long* rep) {
  return orc::call(ctx, [&](auto& b) { *rep = b.gsis_correction(wh, wn, nw, flags, flags ? nf : 0); });
}
static int o_flux(void* ctx, const int32_t* c, const int32_t* e, size_t n, double* out) {
  return orc::call(ctx, [&](auto& b) { b.face_flux(c, e, n, out); });
}
static int o_timings(void* ctx, gsk_timings* out) {
  orc::Handle* h = static_cast<orc::Handle*>(ctx);
  *out = gsk_timings{0, 0, 0, 0, h->t_sweep, 0, 0};
  h->t_sweep = 0.0;
  return GSK_OK;
}
static int o_deq(void* ctx, gsk_deq_counts* out) {
  return orc::call(ctx, [&](auto& b) { *out = b.deq_counts(); });
}
static const char* o_err(void* ctx) { return static_cast<orc::Handle*>(ctx)->err.c_str(); }

int orc_fill_ops(void* h, gsk_backend_ops* ops) {
  if (!h || !ops) return GSK_ERR_INVALID;
  *ops = gsk_backend_ops{h,       o_name,  o_chunk,  o_mem,  o_set_eq, o_set_f,    o_get_f,  o_moments, o_scale, o_params,
                         o_inflow, o_bsums, o_sweep, o_gsis, o_flux,   o_timings, o_deq,    o_err};
  return GSK_OK;
}

}  // extern "C"
