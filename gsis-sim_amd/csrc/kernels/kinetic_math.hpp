#pragma once
// Per (cell, velocity) arithmetic of the kinetic step for the HIP kernels, templated on the
// storage precision F of f (double for the drop-in FP64 form, float for the mixed form).
//
// Each function restates one reference formula; the operation order is kept so that the FP64
// form agrees with the reference CPU path to rounding (all files relative to the reference tree):
//   Maxwellian / Shakhov params         src/physics/maxwell.hpp:17-64
//   Shakhov target with conservation    src/kinetic/kinetic_types.hpp:68-90
//   Venkatakrishnan limiter             src/physics/limiter.hpp:17-50
//   LSQ gradient + limiter              src/kinetic/kinetic_kernels.hpp:145-201
//   residual (2nd-order upwind)         src/kinetic/kinetic_kernels.hpp:238-278
//   LU-SGS forward/backward cell        src/kinetic/kinetic_kernels.hpp:290-358
//   moment terms / finalize             src/kinetic/kinetic_kernels.hpp:378-442
//   cell params                         src/kinetic/kinetic_kernels.hpp:451-471
//   discrete equilibrium Newton         src/kinetic/discrete_equilibrium.hpp:25-188,
//                                       src/kinetic/kinetic_kernels.hpp:489-493
//   GSIS-correction helpers             src/kinetic/kinetic_kernels.hpp:557-587
//   boundary terms / reflections        src/kinetic/kinetic_kernels.hpp:76-133, :596-609

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GSK_HD __host__ __device__ __forceinline__
#else
#define GSK_HD inline
#endif

namespace gsk::dev {

inline constexpr double kPi = 3.14159265358979323846;
inline constexpr int kMomentSums = 13;
inline constexpr int kMoments = 15;
inline constexpr int kBoundarySums = 20;
inline constexpr int kDeqSums = 20;

enum : int { kBcDiffuse = 0, kBcPressure = 1, kBcVacuum = 2, kBcSymmetry = 3, kBcPeriodic = 4, kBcMaxwell = 5 };
enum : int { kLimNone = 0, kLimAbs = 1, kLimRel = 2 };

GSK_HD double dexp(double x) { return ::exp(x); }
GSK_HD float dexp(float x) { return ::expf(x); }
GSK_HD double dsqrt(double x) { return ::sqrt(x); }
GSK_HD float dsqrt(float x) { return ::sqrtf(x); }
GSK_HD double dabs(double x) { return ::fabs(x); }
GSK_HD float dabs(float x) { return ::fabsf(x); }
template <class T>
GSK_HD T dmin(T a, T b) { return b < a ? b : a; }
template <class T>
GSK_HD T dmax(T a, T b) { return a < b ? b : a; }
// the same for operands that are neither NaN nor -0 (one v_max / v_min instead of compare + select)
GSK_HD double nmax(double a, double b) { return ::fmax(a, b); }
GSK_HD float nmax(float a, float b) { return ::fmaxf(a, b); }
GSK_HD double nmin(double a, double b) { return ::fmin(a, b); }
GSK_HD float nmin(float a, float b) { return ::fminf(a, b); }

// Dot products and multiply-adds with separately rounded products and left-to-right sums, the
// reference's form, in both precisions (the files build with -ffp-contract=off). FP64 needs it for
// ulp-level parity. In FP32, fused forms (one v_fma_f32 per term) saved ~1% of a step but raised the
// iteration-to-iteration jitter of a converged FP32 solution about six-fold (the convergence measure
// of eq. (59) then stalls near 2e-6 on the reference's slab cases instead of 3e-7 to 5e-7).
template <class T>
GSK_HD T dot3(T x, T y, T z, T a, T b, T c) {
  return x * a + y * b + z * c;
}
template <class T>
GSK_HD T madd(T acc, T a, T b) {
  return acc + a * b;
}

// f_eq = A exp(-B |xi - u|^2)
template <class F>
struct Maxw {
  F ux, uy, uz, A, B;
};

template <class F>
GSK_HD F maxw_eval(const Maxw<F>& m, F x, F y, F z) {
  const F cx = x - m.ux, cy = y - m.uy, cz = z - m.uz;
  return m.A * dexp(-m.B * dot3(cx, cy, cz, cx, cy, cz));
}

// (rho, u, T) -> (A, B) with R = 1, computed in double and stored in F
template <class F>
GSK_HD Maxw<F> maxw_from_state(double rho, double ux, double uy, double uz, double T) {
  const double theta = T;
  const double tpt = (2.0 * kPi) * theta;
  Maxw<F> m;
  m.ux = static_cast<F>(ux);
  m.uy = static_cast<F>(uy);
  m.uz = static_cast<F>(uz);
  m.A = static_cast<F>(rho / (tpt * ::sqrt(tpt)));
  m.B = static_cast<F>(1.0 / (2.0 * theta));
  return m;
}

// Shakhov target f_S = eq [1 + coef (c.q)(|c|^2/theta - 5) + g0 + g.c + g4 |c|^2/2], c = xi - uc
template <class F>
struct Target {
  Maxw<F> eq;
  F ucx, ucy, ucz;
  F qx, qy, qz;
  F theta, coef;
  F g0, g4, gx, gy, gz;
};

template <class F>
struct CellP {
  Target<F> s;
  F coll;  // V / tau
  F diag;  // V / dt + V / tau
};

template <class F>
GSK_HD F shakhov_poly(const Target<F>& s, F cx, F cy, F cz, F c2) {
  return s.coef * dot3(cx, cy, cz, s.qx, s.qy, s.qz) * (c2 / s.theta - F(5));
}

template <class F>
GSK_HD F target_eval(const Target<F>& s, F x, F y, F z) {
  const F cx = x - s.ucx, cy = y - s.ucy, cz = z - s.ucz;
  const F c2 = dot3(cx, cy, cz, cx, cy, cz);
  const F m = maxw_eval(s.eq, x, y, z);
  return m * (F(1) + shakhov_poly(s, cx, cy, cz, c2) + madd(s.g0 + dot3(cx, cy, cz, s.gx, s.gy, s.gz), F(0.5) * s.g4, c2));
}

// cell params from 15 moments (kinetic_kernels.hpp:451-471)
template <class F>
GSK_HD CellP<F> make_cell_params(const double* m, double volume, double vdt, double mu0, double omega, double Pr,
                                 double cold_bgk_T) {
  const Maxw<double> eqd = maxw_from_state<double>(m[0], m[1], m[2], m[3], m[4]);
  const double theta = m[4];
  const double p = m[0] * theta;
  const double coef = (1.0 - Pr) / (5.0 * p * theta);
  CellP<F> cp;
  cp.s.eq.ux = static_cast<F>(eqd.ux);
  cp.s.eq.uy = static_cast<F>(eqd.uy);
  cp.s.eq.uz = static_cast<F>(eqd.uz);
  cp.s.eq.A = static_cast<F>(eqd.A);
  cp.s.eq.B = static_cast<F>(eqd.B);
  cp.s.ucx = cp.s.eq.ux;
  cp.s.ucy = cp.s.eq.uy;
  cp.s.ucz = cp.s.eq.uz;
  cp.s.qx = static_cast<F>(m[12]);
  cp.s.qy = static_cast<F>(m[13]);
  cp.s.qz = static_cast<F>(m[14]);
  cp.s.theta = static_cast<F>(theta);
  cp.s.coef = m[4] < cold_bgk_T ? F(0) : static_cast<F>(coef);
  cp.s.g0 = F(0);
  cp.s.g4 = F(0);
  cp.s.gx = cp.s.gy = cp.s.gz = F(0);
  const double mu = mu0 * ::pow(m[4], omega);
  const double coll = volume * m[5] / mu;
  cp.coll = static_cast<F>(coll);
  cp.diag = static_cast<F>(coll + vdt);
  return cp;
}

// ---- Venkatakrishnan limiter ----
template <class F>
GSK_HD F venkat_face(F d2, F dmaxv, F dminv, F eps2) {
  F d1;
  if (d2 > F(0)) {
    d1 = dmaxv;
  } else if (d2 < F(0)) {
    d1 = dminv;
  } else {
    return F(1);
  }
  const F num = (d1 * d1 + eps2) + F(2) * d1 * d2;
  const F den = d1 * d1 + F(2) * d2 * d2 + d1 * d2 + eps2;
  if (!(den > F(0))) return F(0);
  const F r = num / den;
  return r < F(0) ? F(0) : (F(1) < r ? F(1) : r);
}

// The same with the branch it took (clamp: 0 none, 1 to 0 (also a non-positive denominator), 2 to 1
// (also d2 = 0)), and its value on a given branch with d1 already chosen by the sign of d2 (the
// limiter-branch evaluations of sensitivity solves).
template <class F>
GSK_HD F venkat_tracked(F d2, F dmaxv, F dminv, F eps2, unsigned& clamp) {
  F d1;
  if (d2 > F(0)) {
    d1 = dmaxv;
  } else if (d2 < F(0)) {
    d1 = dminv;
  } else {
    clamp = 2;
    return F(1);
  }
  const F num = (d1 * d1 + eps2) + F(2) * d1 * d2;
  const F den = d1 * d1 + F(2) * d2 * d2 + d1 * d2 + eps2;
  clamp = 1;
  if (!(den > F(0))) return F(0);
  const F r = num / den;
  if (r < F(0)) return F(0);
  clamp = 2;
  if (F(1) < r) return F(1);
  clamp = 0;
  return r;
}

template <class F>
GSK_HD F venkat_frozen(F d2, F d1, F eps2, unsigned clamp) {
  if (clamp == 1) return F(0);
  if (clamp == 2) return F(1);
  const F num = (d1 * d1 + eps2) + F(2) * d1 * d2;
  const F den = d1 * d1 + F(2) * d2 * d2 + d1 * d2 + eps2;
  return num / den;
}

// ---- moments about the per-cell shift u0 ----
template <class F>
GSK_HD void moment_terms(F f, F cx, F cy, F cz, F (&t)[kMomentSums]) {
  t[0] = f;
  t[1] = f * cx;
  t[2] = f * cy;
  t[3] = f * cz;
  t[4] = t[1] * cx;
  t[5] = t[2] * cy;
  t[6] = t[3] * cz;
  t[7] = t[1] * cy;
  t[8] = t[2] * cz;
  t[9] = t[1] * cz;
  const F fc2 = f * (cx * cx + cy * cy + cz * cz);
  t[10] = fc2 * cx;
  t[11] = fc2 * cy;
  t[12] = fc2 * cz;
}

// 13 raw sums (before the weight) and the shift -> 15 moments; rho <= 0 leaves out[1..] = 0
GSK_HD void finalize_moments(const double* s, double w, double u0x, double u0y, double u0z, double* out) {
  const double s0 = s[0];
  if (!(s0 > 0.0)) {
    out[0] = w * s0;
    for (int k = 1; k < kMoments; ++k) out[k] = 0.0;
    return;
  }
  const double dx = s[1] / s0, dy = s[2] / s0, dz = s[3] / s0;
  const double rho = w * s0;
  const double pxx = w * (s[4] - s0 * dx * dx);
  const double pyy = w * (s[5] - s0 * dy * dy);
  const double pzz = w * (s[6] - s0 * dz * dz);
  const double pxy = w * (s[7] - s0 * dx * dy);
  const double pyz = w * (s[8] - s0 * dy * dz);
  const double pxz = w * (s[9] - s0 * dx * dz);
  const double p = (pxx + pyy + pzz) / 3.0;
  const double tr2 = s[4] + s[5] + s[6];
  const double du2 = dx * dx + dy * dy + dz * dz;
  const double s2dx = s[4] * dx + s[7] * dy + s[9] * dz;
  const double s2dy = s[7] * dx + s[5] * dy + s[8] * dz;
  const double s2dz = s[9] * dx + s[8] * dy + s[6] * dz;
  out[0] = rho;
  out[1] = u0x + dx;
  out[2] = u0y + dy;
  out[3] = u0z + dz;
  out[4] = p / rho;
  out[5] = p;
  out[6] = pxx - p;
  out[7] = pyy - p;
  out[8] = pzz - p;
  out[9] = pxy;
  out[10] = pyz;
  out[11] = pxz;
  out[12] = 0.5 * w * (s[10] - dx * tr2 - 2.0 * s2dx + 2.0 * s0 * du2 * dx);
  out[13] = 0.5 * w * (s[11] - dy * tr2 - 2.0 * s2dy + 2.0 * s0 * du2 * dy);
  out[14] = 0.5 * w * (s[12] - dz * tr2 - 2.0 * s2dz + 2.0 * s0 * du2 * dz);
}

// ---- discrete (conservative) equilibrium ----
struct DeqGrid {
  double lo[3], hi[3], d[3];
  double criterion;
  int max_iterations;
  double tolerance;
};

GSK_HD bool deq_resolved(const DeqGrid& g, double ux, double uy, double uz, double T) {
  if (!(T > 0.0)) return false;
  const double u[3] = {ux, uy, uz};
  for (int a = 0; a < 3; ++a) {
    if (2.0 * kPi * kPi * T / (g.d[a] * g.d[a]) < g.criterion) return false;
    const double m1 = g.hi[a] - u[a];
    const double m2 = u[a] - g.lo[a];
    if (!(m1 > 0.0 && m2 > 0.0)) return false;
    if (m1 * m1 < 2.0 * T * g.criterion || m2 * m2 < 2.0 * T * g.criterion) return false;
  }
  return true;
}

template <class F>
GSK_HD void deq_terms(F e, F pm, F cx, F cy, F cz, F (&t)[kDeqSums]) {
  const F h = F(0.5) * (cx * cx + cy * cy + cz * cz);
  t[0] = e;
  t[1] = e * cx;
  t[2] = e * cy;
  t[3] = e * cz;
  t[4] = e * h;
  t[5] = t[1] * cx;
  t[6] = t[1] * cy;
  t[7] = t[1] * cz;
  t[8] = t[1] * h;
  t[9] = t[2] * cy;
  t[10] = t[2] * cz;
  t[11] = t[2] * h;
  t[12] = t[3] * cz;
  t[13] = t[3] * h;
  t[14] = t[4] * h;
  const F ep = e * pm;
  t[15] = ep;
  t[16] = ep * cx;
  t[17] = ep * cy;
  t[18] = ep * cz;
  t[19] = ep * h;
}

// The same sums on a tensor-product velocity grid in product form. There the trial Maxwellian
// A exp(-B |xi - u'|^2), times the point's relative weight, is A times one factor E_a per axis, and
// every term of deq_terms is e times a monomial cx^i cy^j cz^k in c = xi - u (degree <= 4 for the
// 15 Newton sums; the Shakhov polynomial, cubic in c, times 1, c or |c|^2 / 2 for the last five:
// degree <= 5). A sum over all n0 n1 n2 points of such a term is therefore
//   A * mu[0][i] * mu[1][j] * mu[2][k],   mu[a][k] = sum over the nodes of axis a of E_a c_a^k,
// so one Newton step needs 3 (n0 + n1 + n2) one-axis terms instead of n0 n1 n2 point terms, and
// gives the same sums up to rounding. a = A times the weight normalisation.
inline constexpr int kAxisPowers = 6;  // c^0 .. c^5

// one node's terms of the one-axis sums: E c^k for k = 0..5, E = exp(-B (x - up)^2) w, c = x - u
// (up: the trial Maxwellian's velocity on this axis, u: the cell's, w: the node's weight factor)
GSK_HD void deq_axis_terms(double x, double up, double u, double B, double w, double (&t)[kAxisPowers]) {
  const double d = x - up;
  double e = ::exp(-B * (d * d)) * w;
  const double c = x - u;
  for (int k = 0; k < kAxisPowers; ++k) {
    t[k] = e;
    e *= c;
  }
}

GSK_HD double deq_monomial(const double (&mu)[3][kAxisPowers], double a, int i, int j, int k) {
  return a * (mu[0][i] * mu[1][j] * mu[2][k]);
}

// s[0..15) in deq_terms' order: e {1, cx, cy, cz, h, cx cx, cx cy, cx cz, cx h, cy cy, cy cz, cy h,
// cz cz, cz h, h h}, h = |c|^2 / 2
GSK_HD void deq_product_newton_sums(const double (&mu)[3][kAxisPowers], double a, double* s) {
  s[0] = deq_monomial(mu, a, 0, 0, 0);
  s[1] = deq_monomial(mu, a, 1, 0, 0);
  s[2] = deq_monomial(mu, a, 0, 1, 0);
  s[3] = deq_monomial(mu, a, 0, 0, 1);
  s[4] = 0.5 * (deq_monomial(mu, a, 2, 0, 0) + deq_monomial(mu, a, 0, 2, 0) + deq_monomial(mu, a, 0, 0, 2));
  s[5] = deq_monomial(mu, a, 2, 0, 0);
  s[6] = deq_monomial(mu, a, 1, 1, 0);
  s[7] = deq_monomial(mu, a, 1, 0, 1);
  s[8] = 0.5 * (deq_monomial(mu, a, 3, 0, 0) + deq_monomial(mu, a, 1, 2, 0) + deq_monomial(mu, a, 1, 0, 2));
  s[9] = deq_monomial(mu, a, 0, 2, 0);
  s[10] = deq_monomial(mu, a, 0, 1, 1);
  s[11] = 0.5 * (deq_monomial(mu, a, 2, 1, 0) + deq_monomial(mu, a, 0, 3, 0) + deq_monomial(mu, a, 0, 1, 2));
  s[12] = deq_monomial(mu, a, 0, 0, 2);
  s[13] = 0.5 * (deq_monomial(mu, a, 2, 0, 1) + deq_monomial(mu, a, 0, 2, 1) + deq_monomial(mu, a, 0, 0, 3));
  s[14] = 0.25 * (deq_monomial(mu, a, 4, 0, 0) + deq_monomial(mu, a, 0, 4, 0) + deq_monomial(mu, a, 0, 0, 4) +
                  2.0 * (deq_monomial(mu, a, 2, 2, 0) + deq_monomial(mu, a, 2, 0, 2) + deq_monomial(mu, a, 0, 2, 2)));
}

// sum of e pm c^b, pm = coef (c . q)(|c|^2 / theta - 5) = coef sum_d q_d c_d (sum_g c_g^2 / theta - 5)
GSK_HD double deq_product_shakhov_term(const double (&mu)[3][kAxisPowers], double a, double coef, const double (&q)[3],
                                       double theta, int bx, int by, int bz) {
  double total = 0.0;
  for (int d = 0; d < 3; ++d) {
    const int e[3] = {bx + (d == 0), by + (d == 1), bz + (d == 2)};
    double cube = 0.0;
    for (int g = 0; g < 3; ++g)
      cube += deq_monomial(mu, a, e[0] + 2 * (g == 0), e[1] + 2 * (g == 1), e[2] + 2 * (g == 2));
    total += q[d] * (cube / theta - 5.0 * deq_monomial(mu, a, e[0], e[1], e[2]));
  }
  return coef * total;
}

// s[15..20): e pm {1, cx, cy, cz, h}
GSK_HD void deq_product_shakhov_sums(const double (&mu)[3][kAxisPowers], double a, double coef, const double (&q)[3],
                                     double theta, double* s) {
  s[15] = deq_product_shakhov_term(mu, a, coef, q, theta, 0, 0, 0);
  s[16] = deq_product_shakhov_term(mu, a, coef, q, theta, 1, 0, 0);
  s[17] = deq_product_shakhov_term(mu, a, coef, q, theta, 0, 1, 0);
  s[18] = deq_product_shakhov_term(mu, a, coef, q, theta, 0, 0, 1);
  s[19] = 0.5 * (deq_product_shakhov_term(mu, a, coef, q, theta, 2, 0, 0) +
                 deq_product_shakhov_term(mu, a, coef, q, theta, 0, 2, 0) +
                 deq_product_shakhov_term(mu, a, coef, q, theta, 0, 0, 2));
}

GSK_HD void deq_jacobian(const double* s, double w, double (&J)[5][5]) {
  int k = 0;
  for (int a = 0; a < 5; ++a)
    for (int b = a; b < 5; ++b) {
      J[a][b] = w * s[k];
      J[b][a] = J[a][b];
      ++k;
    }
}

// partial-pivot Gauss elimination; false when singular or NaN
GSK_HD bool deq_solve5(double (&A)[5][5], double (&b)[5], double (&x)[5]) {
  for (int k = 0; k < 5; ++k) {
    int piv = k;
    double best = ::fabs(A[k][k]);
    for (int r = k + 1; r < 5; ++r)
      if (::fabs(A[r][k]) > best) {
        best = ::fabs(A[r][k]);
        piv = r;
      }
    if (!(best > 0.0)) return false;
    if (piv != k) {
      for (int c = 0; c < 5; ++c) {
        const double tmp = A[k][c];
        A[k][c] = A[piv][c];
        A[piv][c] = tmp;
      }
      const double tb = b[k];
      b[k] = b[piv];
      b[piv] = tb;
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
    if (!(x[k] == x[k])) return false;
  }
  return true;
}

GSK_HD void deq_initial_beta(double rho, double T, double (&beta)[5]) {
  const double tpt = 2.0 * kPi * T;
  beta[0] = ::log(rho / (tpt * ::sqrt(tpt)));
  beta[1] = 0.0;
  beta[2] = 0.0;
  beta[3] = 0.0;
  beta[4] = -1.0 / T;
}

template <class F>
GSK_HD Maxw<F> deq_params(const double (&beta)[5], double ux, double uy, double uz) {
  const double B = -0.5 * beta[4];
  const double dx = -beta[1] / beta[4];
  const double dy = -beta[2] / beta[4];
  const double dz = -beta[3] / beta[4];
  Maxw<F> m;
  m.ux = static_cast<F>(ux + dx);
  m.uy = static_cast<F>(uy + dy);
  m.uz = static_cast<F>(uz + dz);
  m.A = static_cast<F>(::exp(beta[0] + B * (dx * dx + dy * dy + dz * dz)));
  m.B = static_cast<F>(B);
  return m;
}

// 1 = converged (beta unchanged), 0 = beta updated, -1 = failure; rel = relative moment error
GSK_HD int deq_newton_step(const double* s, double w, double rho, double T, double tol, double (&beta)[5],
                           double& rel) {
  double G[5] = {w * s[0] - rho, w * s[1], w * s[2], w * s[3], w * s[4] - 1.5 * rho * T};
  const double su = rho * ::sqrt(T);
  rel = dmax(::fabs(G[0]) / rho,
             dmax(dmax(::fabs(G[1]), dmax(::fabs(G[2]), ::fabs(G[3]))) / su, ::fabs(G[4]) / (rho * T)));
  if (!(rel == rel)) return -1;
  if (rel < tol) return 1;
  double J[5][5];
  deq_jacobian(s, w, J);
  double rhs[5] = {-G[0], -G[1], -G[2], -G[3], -G[4]};
  double d[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (!deq_solve5(J, rhs, d)) return -1;
  double scale = 1.0;
  if (::fabs(d[4]) > 0.5 * ::fabs(beta[4])) scale = 0.5 * ::fabs(beta[4]) / ::fabs(d[4]);
  if (::fabs(d[0]) * scale > 1.0) scale = 1.0 / ::fabs(d[0]);
  for (int k = 0; k < 5; ++k) beta[k] += scale * d[k];
  if (!(beta[4] < 0.0)) return -1;
  return 0;
}

GSK_HD bool deq_shakhov_correction(const double* s, double w, double (&gamma)[5]) {
  double J[5][5];
  deq_jacobian(s, w, J);
  double rhs[5] = {-w * s[15], -w * s[16], -w * s[17], -w * s[18], -w * s[19]};
  return deq_solve5(J, rhs, gamma);
}

GSK_HD bool deq_accept(int r, double rel, double rel_prev, bool last) {
  if (r == 1) return true;
  if (r == 0 && rel < 1e-4 && (rel > 0.5 * rel_prev || last)) return true;
  return false;
}

// ---- GSIS correction helpers ----
template <class F>
GSK_HD F negative_check_value(F old, F x, bool created) {
  return created ? x - dmin(old, F(0)) : x;
}
template <class F>
GSK_HD F correction_alpha_limit(F old, F d, F thr) {
  if (!(d < F(0))) return F(1);
  const F room = old - (dmin(old, F(0)) + thr);
  return dmin(F(1), room / (-d));
}
template <class F>
GSK_HD F blend_theta(F old, F d, F e, F thr) {
  const F g = old + d;
  const F bound = dmin(old, F(0)) + thr;
  if (!(g < bound)) return F(0);
  const F den = e - g;
  if (!(den > F(0))) return F(1);
  return dmin(F(1), (bound - g) / den);
}

}  // namespace gsk::dev
