// HIP backend of the kinetic step for gfx950 (MI355X): kernels and the host-side backend object.
//
// Layout in HBM: f[cell][storage position] (velocities contiguous, so the 64 lanes of a wave read
// 64 consecutive velocities of one cell: coalesced), chunk workspaces [cell][chunk width], per-cell
// tables as arrays of structs read wave-uniformly (scalar loads). Per-cell reductions over
// velocities (moments, discrete-equilibrium Newton sums, boundary sums) run one workgroup per cell
// or face with FP64 accumulation and a fixed-order wave + LDS tree, so results are deterministic
// and independent of launch splitting (reference rule: docs/design_cis_gpu.md:121-123).
//
// Reference kernels replaced (src/kinetic/kinetic_cuda_kernels.cu): k_set_equilibrium (:147),
// k_moments (:171), k_scale (:207), k_cell_params / k_cell_params_deq (:221-260),
// k_maxwell_reflection (:262), k_boundary_sums (:274), k_internal_face_mass_flux (:326),
// k_gradient (:343), k_residual + reduce (:357-381), k_sweep_forward / backward (:383-425),
// k_gsis_correction (:427). Host flow: src/kinetic/kinetic_backend_cuda.cpp:89-509.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include <mutex>

#include "../host/backend.hpp"
#include "../host/core.hpp"
#include "kinetic_view.hpp"

// Fixed device scratch of the entry points that allocate nothing (Backend::gsis_correction_noalloc):
// each GPU the library runs on holds its own copy until the process ends. Host code never passes
// this name to a kernel or a copy (on the host it is a stand-in, not a device address): it uses the
// address hipGetSymbolAddress returns for the current device (noalloc_scratch() below).
static __device__ __align__(256) unsigned char scratch[GSK_NOALLOC_SCRATCH_BYTES];

namespace gsk {
namespace dev {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// ---- fixed-order block reductions (wave butterfly-free: lane 0 of each wave, then waves in order) ----
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += __shfl_down(v[k], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) lds[w * N + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double s = 0.0;
    for (int i = 0; i < kWaves; ++i) s += lds[i * N + k];
    v[k] = s;
  }
  __syncthreads();
}

// fixed-order wave reduction: an xor butterfly, after which every lane holds the same sum (each step
// adds a pair of values, and a + b = b + a, so the two lanes of a pair agree)
template <int N>
__device__ __forceinline__ void wave_sum(double (&v)[N]) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += __shfl_xor(v[k], off, 64);
}

template <class F, bool Max>
__device__ __forceinline__ F block_extreme(F v, F* lds) {
  for (int off = 32; off > 0; off >>= 1) {
    const F o = __shfl_down(v, off, 64);
    v = Max ? dmax(v, o) : dmin(v, o);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[w] = v;
  __syncthreads();
  F r = lds[0];
  for (int i = 1; i < kWaves; ++i) r = Max ? dmax(r, lds[i]) : dmin(r, lds[i]);
  __syncthreads();
  return r;
}

// Discrete conservative Maxwellian of one cell, all threads of the block cooperating on the sums
// of each Newton step; every thread ends with the same result (discrete_maxwell_seq,
// kinetic_kernels.hpp:499-540). A Newton step needs the 15 conservation / Jacobian sums; the 5
// Shakhov-polynomial sums enter only the correction gamma of the accepted iterate, so they are
// summed in one extra pass for that iterate (each sum is reduced on its own, in the same order, so
// the values equal those of a single 20-sum pass). Returns iterations, 0 when resolved, -1 on
// failure.
// F is the type of the velocity points and of the terms: the f precision for cells, double for
// boundary inflow states (which the reference corrects in FP64 on the host).
constexpr int kDeqNewtonSums = 15;
template <class F>
__device__ int block_discrete_maxwell(int nv, const F* __restrict__ xs, const F* __restrict__ ys,
                                      const F* __restrict__ zs, double weight, const F* __restrict__ wq,
                                      const DeqGrid& grid, double tolerance,
                                      double rho, double ux, double uy, double uz, double T, const Target<F>* poly,
                                      Maxw<F>& out, double (&gamma)[5], double* lds) {
  for (int k = 0; k < 5; ++k) gamma[k] = 0.0;
  out = maxw_from_state<F>(rho, ux, uy, uz, T);
  if (deq_resolved(grid, ux, uy, uz, T)) return 0;
  double beta[5];
  deq_initial_beta(rho, T, beta);
  const F fux = static_cast<F>(ux), fuy = static_cast<F>(uy), fuz = static_cast<F>(uz);
  double rel_prev = 1e300;
  for (int it = 0; it < grid.max_iterations; ++it) {
    const Maxw<F> m = deq_params<F>(beta, ux, uy, uz);
    double sums[kDeqSums];
#pragma unroll
    for (int k = 0; k < kDeqSums; ++k) sums[k] = 0.0;
    for (int p = threadIdx.x; p < nv; p += blockDim.x) {
      const F x = xs[p], y = ys[p], z = zs[p];
      const F cx = x - fux, cy = y - fuy, cz = z - fuz;
      F t[kDeqSums];
      F e = maxw_eval(m, x, y, z);
      if (wq != nullptr) e *= wq[p];  // stretched grid: weight * wq[p] per velocity
      deq_terms<F>(e, F(0), cx, cy, cz, t);
#pragma unroll
      for (int k = 0; k < kDeqNewtonSums; ++k) sums[k] += static_cast<double>(t[k]);
    }
    {
      double s15[kDeqNewtonSums];
#pragma unroll
      for (int k = 0; k < kDeqNewtonSums; ++k) s15[k] = sums[k];
      block_sum<kDeqNewtonSums>(s15, lds);
#pragma unroll
      for (int k = 0; k < kDeqNewtonSums; ++k) sums[k] = s15[k];
    }
    // The 5x5 Newton step runs in one thread, whose pivoting arrays live in private memory; the
    // block reads its result from LDS.
    if (threadIdx.x == 0) {
      double b[5] = {beta[0], beta[1], beta[2], beta[3], beta[4]};
      double rel = 0.0;
      const int r = deq_newton_step(sums, weight, rho, T, tolerance, b, rel);
      const bool accept = deq_accept(r, rel, rel_prev, it + 1 == grid.max_iterations);
      for (int k = 0; k < 5; ++k) lds[k] = b[k];
      lds[5] = rel;
      lds[6] = static_cast<double>(r);
      lds[7] = accept ? 1.0 : 0.0;
    }
    __syncthreads();
    const double rel = lds[5];
    const int r = static_cast<int>(lds[6]);
    const bool accept = lds[7] != 0.0;
    double b[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) b[k] = lds[k];
    __syncthreads();
    if (accept) {
      out = m;
      if (poly != nullptr) {
        double s5[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int p = threadIdx.x; p < nv; p += blockDim.x) {
          const F x = xs[p], y = ys[p], z = zs[p];
          const F cx = x - fux, cy = y - fuy, cz = z - fuz;
          const F pm = shakhov_poly(*poly, cx, cy, cz, cx * cx + cy * cy + cz * cz);
          F t[kDeqSums];
          F e = maxw_eval(m, x, y, z);
          if (wq != nullptr) e *= wq[p];
          deq_terms<F>(e, pm, cx, cy, cz, t);
#pragma unroll
          for (int k = 0; k < 5; ++k) s5[k] += static_cast<double>(t[kDeqNewtonSums + k]);
        }
        block_sum<5>(s5, lds);
#pragma unroll
        for (int k = 0; k < 5; ++k) sums[kDeqNewtonSums + k] = s5[k];
        if (threadIdx.x == 0) {
          double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
          if (!deq_shakhov_correction(sums, weight, g))
            for (int k = 0; k < 5; ++k) g[k] = 0.0;
          for (int k = 0; k < 5; ++k) lds[k] = g[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 5; ++k) gamma[k] = lds[k];
        __syncthreads();
      }
      return it + 1;
    }
    if (r < 0) break;
#pragma unroll
    for (int k = 0; k < 5; ++k) beta[k] = b[k];
    rel_prev = rel;
  }
  out = maxw_from_state<F>(rho, ux, uy, uz, T);
  for (int k = 0; k < 5; ++k) gamma[k] = 0.0;
  return -1;
}

// The discrete Maxwellian of one cell by one wave, with the Newton sums in product form
// (deq_product_newton_sums): the lanes take the nodes of the three axes, form the one-axis sums
// mu[a][k] = sum_i E_a(i) c_a(i)^k, E_a(i) = exp(-B (x_i - u'_a)^2) times the node's weight factor,
// and share them through wave_sum (which leaves the totals in every lane), so every lane takes the
// Newton step itself and no lane waits on another. 3 (n0 + n1 + n2) exponentials per Newton step
// instead of n0 n1 n2. Same iteration, trial parameters (deq_params<F>, as the f precision stores
// them), step control, accept rule and result as block_discrete_maxwell, whose point sums these
// equal up to rounding; the one-axis terms are FP64 in both forms.
__device__ __forceinline__ void deq_axis_sums(const AxisGrid& g, double B, const double (&up)[3], const double (&u)[3],
                                              double (&mu)[3][kAxisPowers]) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  double part[3 * kAxisPowers];
#pragma unroll
  for (int k = 0; k < 3 * kAxisPowers; ++k) part[k] = 0.0;
  int off = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll 1
    for (int i = lane; i < g.n[a]; i += 64) {
      double t[kAxisPowers];
      deq_axis_terms(g.node[off + i], up[a], u[a], B, g.w != nullptr ? g.w[off + i] : 1.0, t);
#pragma unroll
      for (int k = 0; k < kAxisPowers; ++k) part[a * kAxisPowers + k] += t[k];
    }
    off += g.n[a];
  }
  wave_sum<3 * kAxisPowers>(part);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int k = 0; k < kAxisPowers; ++k) mu[a][k] = part[a * kAxisPowers + k];
}

template <class F>
__device__ int wave_discrete_maxwell(const AxisGrid& g, double weight, const DeqGrid& grid, double tolerance,
                                     double rho, double ux, double uy, double uz, double T, const Target<F>* poly,
                                     Maxw<F>& out, double (&gamma)[5]) {
  for (int k = 0; k < 5; ++k) gamma[k] = 0.0;
  out = maxw_from_state<F>(rho, ux, uy, uz, T);
  if (deq_resolved(grid, ux, uy, uz, T)) return 0;
  double beta[5];
  deq_initial_beta(rho, T, beta);
  // c = xi - u with u as the f precision holds it (the point form's x - fux)
  const double u[3] = {static_cast<double>(static_cast<F>(ux)), static_cast<double>(static_cast<F>(uy)),
                       static_cast<double>(static_cast<F>(uz))};
  double rel_prev = 1e300;
  for (int it = 0; it < grid.max_iterations; ++it) {
    const Maxw<F> m = deq_params<F>(beta, ux, uy, uz);
    const double up[3] = {static_cast<double>(m.ux), static_cast<double>(m.uy), static_cast<double>(m.uz)};
    double mu[3][kAxisPowers];
    deq_axis_sums(g, static_cast<double>(m.B), up, u, mu);
    const double a = static_cast<double>(m.A) * g.wscale;
    double sums[kDeqSums];
    deq_product_newton_sums(mu, a, sums);
    double b[5] = {beta[0], beta[1], beta[2], beta[3], beta[4]};
    double rel = 0.0;
    const int r = deq_newton_step(sums, weight, rho, T, tolerance, b, rel);
    if (deq_accept(r, rel, rel_prev, it + 1 == grid.max_iterations)) {
      out = m;
      if (poly != nullptr) {
        const double q[3] = {static_cast<double>(poly->qx), static_cast<double>(poly->qy),
                             static_cast<double>(poly->qz)};
        deq_product_shakhov_sums(mu, a, static_cast<double>(poly->coef), q, static_cast<double>(poly->theta), sums);
        double gm[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (!deq_shakhov_correction(sums, weight, gm))
          for (int k = 0; k < 5; ++k) gm[k] = 0.0;
        for (int k = 0; k < 5; ++k) gamma[k] = gm[k];
      }
      return it + 1;
    }
    if (r < 0) break;
#pragma unroll
    for (int k = 0; k < 5; ++k) beta[k] = b[k];
    rel_prev = rel;
  }
  out = maxw_from_state<F>(rho, ux, uy, uz, T);
  for (int k = 0; k < 5; ++k) gamma[k] = 0.0;
  return -1;
}

template <class F>
__device__ Maxw<F> equilibrium_params(const View<F>& v, double rho, double ux, double uy, double uz, double T, int* result,
                                      double* lds) {
  if (result) *result = 0;
  if (!v.discrete_eq) return maxw_from_state<F>(rho, ux, uy, uz, T);
  Maxw<F> m;
  double gamma[5];
  const int r = block_discrete_maxwell<F>(v.nv, v.xi_x, v.xi_y, v.xi_z, v.weight_d, v.wq, v.deq, v.deq.tolerance, rho,
                                          ux, uy, uz, T, nullptr, m, gamma, lds);
  if (result) *result = r;
  return m;
}

// ---- kernels ------------------------------------------------------------------------------

// Discrete-equilibrium correction of boundary inflow states in FP64, one block per state
// {scale, rho, ux, uy, uz, T}: a state that is not resolved by the grid is replaced by the
// Maxwellian whose discrete moments match it (KineticSolver::correct_inflow, kinetic_solver.cpp:103-114;
// maxwell_params_to_state, kinetic_kernels.hpp:543-552). Newton tolerance tol (1e-10 as the reference).
// index: the states to correct (null: the first n).
__global__ __launch_bounds__(kBlock) void k_inflow_correct(int nv, const double* __restrict__ xd,
                                                           const double* __restrict__ yd, const double* __restrict__ zd,
                                                           double weight, const double* __restrict__ wqd,
                                                           DeqGrid grid, double tol, double* __restrict__ states,
                                                           const int* __restrict__ index, int n) {
  __shared__ double lds[kWaves * kDeqSums];
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    double* st = states + static_cast<i64>(index != nullptr ? index[i] : i) * 6;
    const double rho = st[1], ux = st[2], uy = st[3], uz = st[4], T = st[5];
    if (!(rho > 0.0) || !(T > 0.0)) continue;
    Maxw<double> m;
    double gamma[5];
    const int r = block_discrete_maxwell<double>(nv, xd, yd, zd, weight, wqd, grid, tol, rho, ux, uy, uz, T, nullptr,
                                                 m, gamma, lds);
    if (r > 0 && threadIdx.x == 0) {
      const double Tn = 1.0 / (2.0 * m.B);
      const double tpt = 2.0 * kPi * Tn;
      st[1] = m.A * tpt * ::sqrt(tpt);
      st[2] = m.ux;
      st[3] = m.uy;
      st[4] = m.uz;
      st[5] = Tn;
    }
  }
}

template <class F>
__global__ __launch_bounds__(kBlock) void k_set_equilibrium(View<F> v, const double* __restrict__ states,
                                                            double* __restrict__ shift) {
  __shared__ double lds[kWaves * kDeqSums];
  for (int c = blockIdx.x; c < v.n_cells; c += gridDim.x) {
    const double* st = states + static_cast<i64>(c) * 5;
    const Maxw<F> m = equilibrium_params<F>(v, st[0], st[1], st[2], st[3], st[4], nullptr, lds);
    F* fc = v.f + static_cast<i64>(c) * v.nv;
    for (int p = threadIdx.x; p < v.nv; p += blockDim.x) fc[p] = maxw_eval(m, v.xi_x[p], v.xi_y[p], v.xi_z[p]);
    if (threadIdx.x == 0) {
      shift[c * 3 + 0] = st[1];
      shift[c * 3 + 1] = st[2];
      shift[c * 3 + 2] = st[3];
    }
  }
}

// 13 shifted sums per cell in FP64 -> 15 moments; updates the shift (k_moments)
template <class F>
__global__ __launch_bounds__(kBlock) void k_moments(View<F> v, const double* __restrict__ xd, const double* __restrict__ yd,
                                                    const double* __restrict__ zd, double* __restrict__ shift,
                                                    double* __restrict__ moments) {
  __shared__ double lds[kWaves * kMomentSums];
  for (int c = blockIdx.x; c < v.n_cells; c += gridDim.x) {
    const double u0x = shift[c * 3 + 0], u0y = shift[c * 3 + 1], u0z = shift[c * 3 + 2];
    double s[kMomentSums];
    for (int k = 0; k < kMomentSums; ++k) s[k] = 0.0;
    const F* fc = v.f + static_cast<i64>(c) * v.nv;
    for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
      double t[kMomentSums];
      double fp = static_cast<double>(fc[p]);
      if (v.wqd != nullptr) fp *= v.wqd[p];  // stretched grid: weight * wqd[p] per velocity
      moment_terms<double>(fp, xd[p] - u0x, yd[p] - u0y, zd[p] - u0z, t);
      for (int k = 0; k < kMomentSums; ++k) s[k] += t[k];
    }
    block_sum<kMomentSums>(s, lds);
    if (threadIdx.x == 0) {
      double* m = moments + static_cast<i64>(c) * kMoments;
      finalize_moments(s, v.weight_d, u0x, u0y, u0z, m);
      if (m[0] > 0.0) {
        shift[c * 3 + 0] = m[1];
        shift[c * 3 + 1] = m[2];
        shift[c * 3 + 2] = m[3];
      }
    }
  }
}

// Moment sums of the FP32 form as a dense contraction on the matrix cores (an extension; the FP64
// form and the reference sum in FP64, k_moments). For a tile of 16 cells (one workgroup),
// D[m][n] = sum_k psi_m(xi_k - s) f[n][k], psi = the 13 monomials of moment_terms (1, c, cc, |c|^2 c;
// rows 13..15 zero) about one shift s per tile: the mean of the tile's cell shifts, rounded to FP32.
// FP32 products and accumulation on v_mfma_f32_16x16x4f32 (A: lane l holds A[l%16][l/16], B: lane l
// holds B[l/16][l%16], D: lane l register i holds D[4(l/16)+i][l%16]). The 4 waves take a quarter of
// the velocities each; FP32 accumulators are added into FP64 every kMfmaBlock velocities, the waves'
// FP64 partials are added in wave order, and finalize_moments forms the moments about s in FP64.
typedef float floatx4 __attribute__((ext_vector_type(4)));
constexpr int kMfmaBlock = 512;
__global__ __launch_bounds__(256) void k_moments_mfma(View<float> v, double* __restrict__ shift,
                                                      double* __restrict__ moments) {
  __shared__ double part[4][16][16];  // [wave][cell][moment row]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = blockIdx.x * 16;
  const int n = lane & 15, g = lane >> 4;
  const int nvalid = min(16, v.n_cells - c0);
  const i64 c = c0 + min(n, nvalid - 1);
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int j = 0; j < nvalid; ++j) {
    sx += shift[(c0 + j) * 3 + 0];
    sy += shift[(c0 + j) * 3 + 1];
    sz += shift[(c0 + j) * 3 + 2];
  }
  const float fsx = static_cast<float>(sx / nvalid), fsy = static_cast<float>(sy / nvalid),
              fsz = static_cast<float>(sz / nvalid);
  // this lane's monomial (row m = n): e1 * e2 * e3 with factors chosen by axis (0 = none)
  const int m = n;
  const int ax1 = (m == 1 || m == 4 || m == 7 || m == 9 || m == 10) ? 1
                  : (m == 2 || m == 5 || m == 8 || m == 11)         ? 2
                  : (m == 3 || m == 6 || m == 12)                   ? 3
                                                                    : 0;
  const int ax2 = (m == 4) ? 1 : (m == 5 || m == 7) ? 2 : (m == 6 || m == 8 || m == 9) ? 3 : 0;
  const bool cube = m >= 10 && m <= 12;
  const float base = m < 13 ? 1.0f : 0.0f;
  // velocity segment of this wave (multiple of 16 long)
  const int seg = ((v.nv + 63) / 64) * 16;
  const int k_begin = wave * seg, k_end = min(v.nv, k_begin + seg);
  const float* frow = v.f + c * v.nv;
  const bool vec4 = (v.nv % 4) == 0;  // rows 16-byte aligned: float4 loads
  double acc64[4] = {0.0, 0.0, 0.0, 0.0};
  floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = k_begin; k0 < k_end; k0 += 16) {
    const int pb = k0 + 4 * g;  // this lane's 4 velocities pb .. pb + 3 (one per MFMA)
    float fv[4], xv[4], yv[4], zv[4], wv[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    if (vec4 && pb + 3 < k_end) {
      const float4 f4 = *reinterpret_cast<const float4*>(frow + pb);
      const float4 x4 = *reinterpret_cast<const float4*>(v.xi_x + pb);
      const float4 y4 = *reinterpret_cast<const float4*>(v.xi_y + pb);
      const float4 z4 = *reinterpret_cast<const float4*>(v.xi_z + pb);
      fv[0] = f4.x, fv[1] = f4.y, fv[2] = f4.z, fv[3] = f4.w;
      xv[0] = x4.x, xv[1] = x4.y, xv[2] = x4.z, xv[3] = x4.w;
      yv[0] = y4.x, yv[1] = y4.y, yv[2] = y4.z, yv[3] = y4.w;
      zv[0] = z4.x, zv[1] = z4.y, zv[2] = z4.z, zv[3] = z4.w;
      if (v.wq != nullptr) {
        const float4 w4 = *reinterpret_cast<const float4*>(v.wq + pb);
        wv[0] = w4.x, wv[1] = w4.y, wv[2] = w4.z, wv[3] = w4.w;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int p = pb + j;
        const bool ok = p < k_end;
        const int q = ok ? p : k_begin;
        fv[j] = ok ? frow[q] : 0.0f;
        xv[j] = v.xi_x[q];
        yv[j] = v.xi_y[q];
        zv[j] = v.xi_z[q];
        if (v.wq != nullptr) wv[j] = v.wq[q];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float cx = xv[j] - fsx, cy = yv[j] - fsy, cz = zv[j] - fsz;
      const float e1 = ax1 == 1 ? cx : ax1 == 2 ? cy : ax1 == 3 ? cz : base;
      const float e2 = ax2 == 1 ? cx : ax2 == 2 ? cy : ax2 == 3 ? cz : 1.0f;
      const float e3 = cube ? cx * cx + cy * cy + cz * cz : 1.0f;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(e1 * e2 * e3 * wv[j], fv[j], acc, 0, 0, 0);
    }
    if ((k0 + 16 - k_begin) % kMfmaBlock == 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        acc64[i] += static_cast<double>(acc[i]);
        acc[i] = 0.0f;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) part[wave][n][4 * g + i] = acc64[i] + static_cast<double>(acc[i]);
  __syncthreads();
  if (threadIdx.x < nvalid) {
    double s[kMomentSums];
#pragma unroll
    for (int r = 0; r < kMomentSums; ++r)
      s[r] = ((part[0][threadIdx.x][r] + part[1][threadIdx.x][r]) + part[2][threadIdx.x][r]) + part[3][threadIdx.x][r];
    const i64 cc = c0 + threadIdx.x;
    double* mo = moments + cc * kMoments;
    finalize_moments(s, v.weight_d, static_cast<double>(fsx), static_cast<double>(fsy), static_cast<double>(fsz), mo);
    if (mo[0] > 0.0) {
      shift[cc * 3 + 0] = mo[1];
      shift[cc * 3 + 1] = mo[2];
      shift[cc * 3 + 2] = mo[3];
    }
  }
}

template <class F>
__global__ void k_scale_f(F* __restrict__ f, i64 n, F s) {
  for (i64 i = blockIdx.x * static_cast<i64>(blockDim.x) + threadIdx.x; i < n; i += static_cast<i64>(gridDim.x) * blockDim.x)
    f[i] *= s;
}

__global__ void k_scale_moments(double* __restrict__ moments, int n_cells, double s) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  double* m = moments + static_cast<i64>(c) * kMoments;
  m[0] *= s;
  for (int k = 5; k < kMoments; ++k) m[k] *= s;
}

// Cell parameters of one cell from its 15 moments: Shakhov target, collision rate and diagonal
// (make_cell_params), then, with the discrete equilibrium, the Newton for its Maxwellian and the
// Shakhov-polynomial correction gamma (k_cell_params_deq, kinetic_cuda_kernels.cu:232-260); a cell
// whose Newton fails keeps the continuous target, or with deq_fallback turns collisionless. Computed
// by one wave (wave_discrete_maxwell); every lane ends with the same result. r: the Newton's
// iterations (0 resolved, -1 failed).
template <class F>
__device__ CellP<F> cell_params_wave(const View<F>& v, const double* m, double volume, double vdt, double mu0,
                                     double omega, double Pr, int& r) {
  CellP<F> cp = make_cell_params<F>(m, volume, vdt, mu0, omega, Pr, v.cold_bgk);
  r = 0;
  if (v.discrete_eq) {
    Maxw<F> eq;
    double gamma[5];
    r = wave_discrete_maxwell<F>(v.axes, v.weight_d, v.deq, v.deq.tolerance, m[0], m[1], m[2], m[3], m[4], &cp.s, eq,
                                 gamma);
    if (r > 0) {
      cp.s.eq = eq;
      cp.s.g0 = static_cast<F>(gamma[0]);
      cp.s.gx = static_cast<F>(gamma[1]);
      cp.s.gy = static_cast<F>(gamma[2]);
      cp.s.gz = static_cast<F>(gamma[3]);
      cp.s.g4 = static_cast<F>(gamma[4]);
    } else if (r < 0 && v.deq_fallback) {
      cp.coll = F(0);
      cp.diag = static_cast<F>(vdt);
    }
  }
  return cp;
}

// Shakhov target, collision rate and diagonal per cell: one wave per cell (cell_params_wave),
// kWaves cells per workgroup, no workgroup barrier.
template <class F>
__global__ __launch_bounds__(kBlock) void k_cell_params(View<F> v, const double* __restrict__ moments,
                                                        const double* __restrict__ volume, const double* __restrict__ vdt,
                                                        double mu0, double omega, double Pr, CellP<F>* __restrict__ cells,
                                                        int* __restrict__ flags) {
  const int lane = static_cast<int>(threadIdx.x & 63), wave = static_cast<int>(threadIdx.x >> 6);
  for (int c = blockIdx.x * kWaves + wave; c < v.n_cells; c += gridDim.x * kWaves) {
    int r = 0;
    const CellP<F> cp = cell_params_wave(v, moments + static_cast<i64>(c) * kMoments, volume[c], vdt[c], mu0, omega,
                                         Pr, r);
    if (lane == 0) {
      cells[c] = cp;
      flags[c] = r;
    }
  }
}

// specular caches of partially accommodating walls (k_maxwell_reflection)
template <class F, bool Reconstruct, int Mode = 0>  // Mode: limiter branches (limited_gradient)
__global__ void k_maxwell_reflection(View<F> v, F* __restrict__ out) {
  const int b = blockIdx.y;
  if (b >= v.n_bfaces || v.mw_off[b] < 0) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= v.nv) return;
  const BFace<F>& bf = v.bf[b];
  const F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
  out[v.mw_off[b] + p] =
      (x * bf.nx + y * bf.ny + z * bf.nz) < F(0) ? reflected_value<Reconstruct, Mode>(v, bf, p, x, y, z) : F(0);
}

// 20 half-space sums per boundary face in FP64 (k_boundary_sums): flux and moment sums (5 each)
// over the outgoing velocities (slots 0..4, 10..14) and over the incoming ones (5..9, 15..19).
// faces: optional subset list. Block z = 0 sums the outgoing side, z = 1 the incoming side (each
// block holds 10 accumulators); the velocities of a face are split over gridDim.y blocks (block y
// takes every gridDim.y-th run of blockDim velocities). k_boundary_reduce adds the partials in order.
constexpr int kSideSums = 10;
template <class F, int Mode = 0>  // Mode: limiter branches (limited_gradient)
__global__ __launch_bounds__(kBlock) void k_boundary_sums(View<F> v, const int* __restrict__ faces, int n_faces,
                                                          double* __restrict__ part) {
  __shared__ double lds[kWaves * kSideSums];
  const int nchunk = static_cast<int>(gridDim.y), chunk = static_cast<int>(blockIdx.y);
  const bool out_side = blockIdx.z == 0;
  for (int i = blockIdx.x; i < n_faces; i += gridDim.x) {
    const int b = faces ? faces[i] : i;
    const BFace<F> bf = v.bf[b];
    double acc[kSideSums];
#pragma unroll
    for (int k = 0; k < kSideSums; ++k) acc[k] = 0.0;
    for (int p = chunk * blockDim.x + threadIdx.x; p < v.nv; p += nchunk * blockDim.x) {
      const F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
      const F an = x * bf.nx + y * bf.ny + z * bf.nz;
      const bool outgoing = an >= F(0) || bf.kind == kBcPeriodic;
      if (outgoing != out_side) continue;
      F fb;
      if (outgoing) {
        fb = reconstruct_at_boundary<Mode>(v, bf, p, x, y, z);
      } else if (bf.kind == kBcSymmetry) {
        const int pm = p ^ bf.mirror_xor;
        fb = reconstruct_at_boundary<Mode>(v, bf, pm, v.xi_x[pm], v.xi_y[pm], v.xi_z[pm]);
      } else if (bf.kind == kBcMaxwell) {
        fb = maxwell_inflow(v, b, p, x, y, z);
      } else {
        fb = maxw_eval(v.inflow[b].m, x, y, z);
      }
      if (v.wq != nullptr) fb *= v.wq[p];  // stretched grid
      const F wf = v.weight * fb;
      const F e = F(0.5) * (x * x + y * y + z * z);
      const F af = an * wf;
      acc[0] += static_cast<double>(af);
      acc[1] += static_cast<double>(af * x);
      acc[2] += static_cast<double>(af * y);
      acc[3] += static_cast<double>(af * z);
      acc[4] += static_cast<double>(af * e);
      acc[5] += static_cast<double>(wf);
      acc[6] += static_cast<double>(wf * x);
      acc[7] += static_cast<double>(wf * y);
      acc[8] += static_cast<double>(wf * z);
      acc[9] += static_cast<double>(wf * e);
    }
    block_sum<kSideSums>(acc, lds);
    if (threadIdx.x == 0) {
      double* o = part + ((static_cast<i64>(i) * nchunk + chunk) * 2 + blockIdx.z) * kSideSums;
#pragma unroll
      for (int k = 0; k < kSideSums; ++k) o[k] = acc[k];
    }
  }
}

__global__ void k_boundary_reduce(const int* __restrict__ faces, int n_faces, int nchunk, const double* __restrict__ part,
                                  double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_faces * kBoundarySums) return;
  const int i = j / kBoundarySums, k = j % kBoundarySums;
  const int b = faces ? faces[i] : i;
  const int side = (k % 10) / 5, slot = (k / 10) * 5 + k % 5;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[((static_cast<i64>(i) * nchunk + c) * 2 + side) * kSideSums + slot];
  out[static_cast<i64>(b) * kBoundarySums + k] = s;
}

// Wall part of the inner-sweep boundary refresh (the reference's KineticSolver::refresh_wall_inflow,
// src/kinetic/kinetic_solver.cpp:456-480) from the wall faces' boundary sums s: a Maxwell wall takes
// scale <- max(0, scale - (s[0] + s[5]) / unit0) and flags a value below -1e-8; a diffuse wall takes
// rho = max(0, -s[0] / s[5]) (0 unless s[5] < 0) and scale <- rho, or w rho + (1 - w) scale when
// w = relaxation < 1. The FP64 scales are the solver's own; inflow[b].scale is their cast to F.
template <class F>
__global__ void k_wall_refresh(const int* __restrict__ faces, const int* __restrict__ maxwell,
                               const double* __restrict__ unit0, int n, double relaxation,
                               const double* __restrict__ bsums, double* __restrict__ scale,
                               Inflow<F>* __restrict__ inflow, int* __restrict__ negative) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = faces[i];
  const double* s = bsums + static_cast<i64>(b) * kBoundarySums;
  double sc = scale[b];
  if (maxwell[i]) {
    const double next = sc - (s[0] + s[5]) / unit0[i];
    if (next < -1e-8) *negative = 1;
    sc = 0.0 < next ? next : 0.0;
  } else {
    const double jp = s[0], jm = s[5];
    const double q = jm < 0.0 ? -jp / jm : 0.0;
    const double rho = jm < 0.0 ? (0.0 < q ? q : 0.0) : 0.0;
    sc = relaxation == 1.0 ? rho : relaxation * rho + (1.0 - relaxation) * sc;
  }
  scale[b] = sc;
  inflow[b].scale = static_cast<F>(sc);
}

// section mass flux through listed internal faces (k_internal_face_mass_flux)
template <class F>
__global__ __launch_bounds__(kBlock) void k_face_mass_flux(View<F> v, const int* __restrict__ cells,
                                                           const int* __restrict__ entries, int n, double* __restrict__ out) {
  __shared__ double lds[kWaves];
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int c = cells[i];
    const CsrE<F> e = v.csr[entries[i]];
    double acc[1] = {0.0};
    for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
      double term = static_cast<double>(internal_face_flux_term(v, c, e, p, v.xi_x[p], v.xi_y[p], v.xi_z[p]));
      if (v.wqd != nullptr) term *= v.wqd[p];  // stretched grid
      acc[0] += term;
    }
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) out[i] = static_cast<double>(v.weight) * acc[0];
  }
}

// XCD-aware (cell, velocity tile) mapping. Workgroups are dealt round-robin to the 8 XCDs; XCD x
// takes the velocity tiles t = x (mod 8) of every cell, cells in order. The f / gradient rows that a
// cell re-reads from its neighbours (close in the RCB/RCM numbering) were then touched by the same
// XCD a few hundred workgroups earlier, with only an eighth of the velocities in between. (Walking
// all cells of one tile first, with 256- or 64-velocity tiles, reads rows at the stride of a full
// velocity row and measured slower.) Grid: 8 * ceil(tiles / 8) * n_cells workgroups.
// L is the work-item index, L0 + blockIdx.x for a launch that starts at L0 (a multiple of 8, so
// L mod 8 is the launching XCD); a large chunk is covered by several launches. The host keeps the
// total below 2^32 work items, so the index arithmetic is 32-bit unsigned (a 64-bit division
// costs each wave ~150 scalar instructions, a quarter of the gradient's whole instruction stream).
__device__ __forceinline__ bool xcd_tile(unsigned L, int n_cells, int tiles, int& cell, int& tile) {
  const unsigned x = L & 7u;
  const unsigned k = L >> 3;
  const int ntx = (tiles - static_cast<int>(x) + 7) / 8;  // tiles owned by this XCD
  if (ntx <= 0) return false;
  const unsigned c = k / static_cast<unsigned>(ntx);
  if (c >= static_cast<unsigned>(n_cells)) return false;
  cell = static_cast<int>(c);
  tile = static_cast<int>(x + 8u * (k - c * static_cast<unsigned>(ntx)));
  return true;
}

template <class F, int Mode = 0>  // Mode: limiter branches (0 off, 1 record, 2 follow the record)
__global__ __launch_bounds__(kBlock) void k_gradient(View<F> v, Chunk<F> ch, int tiles, unsigned L0) {
  int c, tile;
  if (!xcd_tile(L0 + blockIdx.x, v.n_cells, tiles, c, tile)) return;
  const int t = tile * kBlock + threadIdx.x;
  if (t >= ch.kc) return;
  const int p = ch.k0 + t;
  F gx, gy, gz;
  const EllRow<F>& r = v.ell[c];
  const F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p], fi = f_at(v, c, p);
  if constexpr (Mode != 0)
    limited_gradient_row_on_branch<Mode>(v, r, c, p, x, y, z, fi, gx, gy, gz);
  else if ((r.skip | r.bnd) == 0)
    limited_gradient_row<true>(v, r, c, p, x, y, z, fi, gx, gy, gz);
  else
    limited_gradient_row<false>(v, r, c, p, x, y, z, fi, gx, gy, gz);
  const i64 wi = static_cast<i64>(c) * ch.stride + t;
  ch.gx[wi] = gx;
  ch.gy[wi] = gy;
  ch.gz[wi] = gz;
}

// cells with more than kGather faces, from the CSR (after k_gradient, which wrote them wrongly)
template <class F, int Mode = 0>  // Mode: limiter branches (limited_gradient)
__global__ __launch_bounds__(kBlock) void k_gradient_csr(View<F> v, Chunk<F> ch, const int* __restrict__ cells, int n,
                                                         int tiles) {
  for (i64 L = blockIdx.x; L < static_cast<i64>(n) * tiles; L += gridDim.x) {
    const int c = cells[L / tiles], tile = static_cast<int>(L % tiles);
    const int t = tile * kBlock + threadIdx.x;
    if (t >= ch.kc) continue;
    const int p = ch.k0 + t;
    F gx, gy, gz;
    limited_gradient<Mode>(v, c, p, v.xi_x[p], v.xi_y[p], v.xi_z[p], f_at(v, c, p), gx, gy, gz);
    const i64 wi = static_cast<i64>(c) * ch.stride + t;
    ch.gx[wi] = gx;
    ch.gy[wi] = gy;
    ch.gz[wi] = gz;
  }
}

// residual R into rres and per (cell, tile) partial sums of R^2 (k_residual)
template <class F>
__global__ __launch_bounds__(kBlock) void k_residual(View<F> v, Chunk<F> ch, int tiles, unsigned L0, F* __restrict__ rres,
                                                     double* __restrict__ partial) {
  __shared__ double lds[kWaves];
  int c, tile;
  if (!xcd_tile(L0 + blockIdx.x, v.n_cells, tiles, c, tile)) return;
  const int t = tile * kBlock + threadIdx.x;
  double r2[1] = {0.0};
  if (t < ch.kc) {
    const EllRow<F>& r = v.ell[c];
    const F R = (r.skip | r.bnd) == 0 ? residual_value_row<true>(v, ch, r, c, t) : residual_value_row<false>(v, ch, r, c, t);
    rres[static_cast<i64>(c) * ch.stride + t] = R;
    r2[0] = static_cast<double>(R) * static_cast<double>(R);
    if (v.wqd != nullptr) r2[0] *= v.wqd[ch.k0 + t];  // stretched grid: weighted sum of R^2
  }
  block_sum<1>(r2, lds);
  if (threadIdx.x == 0) partial[static_cast<i64>(c) * tiles + tile] = r2[0];
}

// cells with more than kGather faces, from the CSR (after k_residual, which wrote them wrongly)
template <class F>
__global__ __launch_bounds__(kBlock) void k_residual_csr(View<F> v, Chunk<F> ch, const int* __restrict__ cells, int n,
                                                         int tiles, F* __restrict__ rres, double* __restrict__ partial) {
  __shared__ double lds[kWaves];
  for (i64 L = blockIdx.x; L < static_cast<i64>(n) * tiles; L += gridDim.x) {
    const int c = cells[L / tiles], tile = static_cast<int>(L % tiles);
    const int t = tile * kBlock + threadIdx.x;
    double r2[1] = {0.0};
    if (t < ch.kc) {
      const F R = residual_value(v, ch, c, t);
      rres[static_cast<i64>(c) * ch.stride + t] = R;
      r2[0] = static_cast<double>(R) * static_cast<double>(R);
      if (v.wqd != nullptr) r2[0] *= v.wqd[ch.k0 + t];
    }
    block_sum<1>(r2, lds);
    if (threadIdx.x == 0) partial[static_cast<i64>(c) * tiles + tile] = r2[0];
  }
}

__global__ void k_residual_reduce(const double* __restrict__ partial, int tiles, int n_cells, double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  double s = 0.0;
  for (int i = 0; i < tiles; ++i) s += partial[static_cast<i64>(c) * tiles + i];
  out[c] += s;
}

// ---- LU-SGS sweeps, level scheduled -------------------------------------------------------
// The reference walks each spatial block's cells one after another per velocity (one thread per
// (block, velocity), kinetic_cuda_kernels.cu:383-425). Here the cells of a block are grouped by
// dependency level for each sweep direction: a cell's level is one more than the highest level of
// the in-block neighbours that precede it in that direction's order (sweep_pos). Every cell then
// reads only neighbours of lower levels, which are final, so processing the levels in turn gives
// exactly the sequential result. One workgroup owns (block, direction, 64-velocity tile); its waves
// share the cells of a level (lanes = velocities, coalesced) and meet at a barrier between levels.
constexpr int kMaxDirs = 16;
struct VelRange {
  int beg[kMaxDirs];
  int end[kMaxDirs];
};

// One cell of one sweep (forward or backward) for one direction: its faces in CSR order with the
// class each has in that sweep, so the kernel needs no order / position / boundary lookups.
// code = neighbour << 2 | class; class 0 skip (periodic), 1 contributes a+ only (boundary, or an
// in-block neighbour on the downstream side), 2 in-block upstream neighbour, 3 out-of-block
// neighbour. The record keeps the faces of classes 2, 3 and 1 in that order (each group in CSR
// order) with their row offsets precomputed for the backend's chunk width, so the kernels test no
// class and multiply no index. nf = -1: more than kGather faces, use the CSR path.
// diag is the cell's V/tau + V/dt, copied in by k_record_diag whenever the cell parameters change.
//
// The records of a (direction, block) are stored level by level; a level wider than one LDS stage
// is cut into several stages (its cells are independent, so the cut only adds a barrier).
template <class F>
struct SweepRec {
  int cell, nf;        // nf: faces that take part (slots [0, nf)); -1: more than kGather faces, CSR path
  int n_in, n_out;     // the first n_in slots are in-block upstream faces, the next n_out out-of-block ones
  i64 row_w;           // byte offset of the cell's row in the chunk buffers (cell * chunk width * sizeof(F))
  i64 row_f;           // byte offset of the cell's row in f (cell * nv * sizeof(F))
  i64 nb_w[kGather];   // byte offsets of the neighbours' rows in the chunk buffers (first two groups)
  F diag;
  F sx[kGather], sy[kGather], sz[kGather];  // zero in the slots past nf
};
template <class F>
constexpr int kRecWords = static_cast<int>(sizeof(SweepRec<F>) / 4);
static_assert(sizeof(SweepRec<double>) % 8 == 0 && kRecWords<double> <= 64, "record must fit one wave load");
static_assert(sizeof(SweepRec<float>) % 4 == 0 && kRecWords<float> <= 64, "record must fit one wave load");

// LDS staging of the records: a stage holds up to kStageRegs * threads words; two stage buffers
// alternate (one is read while the next stage loads into the other).
constexpr int kMaxSweepThreads = 512;
constexpr int kStageRegs = 6;
constexpr int kStageWords = kStageRegs * kMaxSweepThreads;
template <class F>
constexpr int stage_capacity(int threads) {
  return kStageRegs * threads / kRecWords<F>;
}

template <class F>
__global__ void k_record_diag(const CellP<F>* __restrict__ cells, SweepRec<F>* __restrict__ recs, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) recs[i].diag = cells[recs[i].cell].diag;
}

// Row access with a wave-uniform row pointer (scalar registers) plus the lane's fixed byte offset,
// so no per-lane 64-bit address arithmetic is spent per neighbour.
template <class F>
__device__ __forceinline__ F row_load(const F* row, unsigned byte_off) {
  return *reinterpret_cast<const F*>(reinterpret_cast<const char*>(row) + byte_off);
}
template <class F>
__device__ __forceinline__ F& row_ref(F* row, unsigned byte_off) {
  return *reinterpret_cast<F*>(reinterpret_cast<char*>(row) + byte_off);
}
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Values loaded once before a loop: an empty register-constraint statement marks them as defined
// there, so the compiler does not re-wait on their loads (and on every memory access issued after
// them, stores included) inside the loop.
template <class F>
__device__ __forceinline__ void settle(F& x, F& y, F& z) {
  asm volatile("" : "+v"(x), "+v"(y), "+v"(z));
}

// one cell of the forward sweep (sweep_forward_cell, kinetic_kernels.hpp:290-324); the upwind
// neighbour values are gathered before the in-order accumulation
template <class F>
__device__ __forceinline__ void forward_cell(const View<F>& v, int cell, int t, int c0, int c1, const int* pos, F x, F y,
                                             F z, int stride, const F* __restrict__ rres, F* cur,
                                             const F* __restrict__ prev) {
  F ap = F(0), am_out = F(0), L = F(0), Lout = F(0);
  const int e0 = v.csr_off[cell], e1 = v.csr_off[cell + 1];
  const int pi = pos ? pos[cell] : cell;
  const i64 wi = static_cast<i64>(cell) * stride + t;
  const F Ri = rres[wi];
  if (e1 - e0 <= kGather) {
    F ak[kGather], nb[kGather];
#pragma unroll
    for (int k = 0; k < kGather; ++k)
      if (k < e1 - e0) {
        const CsrE<F>& e = v.csr[e0 + k];
        ak[k] = x * e.sx + y * e.sy + z * e.sz;
        nb[k] = F(0);
        if (e.other >= 0 && !(ak[k] > F(0))) {
          const i64 wj = static_cast<i64>(e.other) * stride + t;
          if (e.other < c0 || e.other >= c1) {
            if (prev != nullptr) nb[k] = prev[wj];
          } else if ((pos ? pos[e.other] : e.other) < pi) {
            nb[k] = cur[wj];
          }
        }
      }
#pragma unroll
    for (int k = 0; k < kGather; ++k)
      if (k < e1 - e0) {
        const CsrE<F>& e = v.csr[e0 + k];
        if (e.other < 0 && v.bf[e.bface].kind == kBcPeriodic) continue;
        const F a = ak[k];
        if (a > F(0)) {
          ap += a;
        } else if (e.other >= 0) {
          if (e.other < c0 || e.other >= c1) {
            if (prev != nullptr)
              Lout += a * nb[k];
            else
              am_out -= a;
          } else if ((pos ? pos[e.other] : e.other) < pi) {
            L += a * nb[k];
          }
        }
      }
  } else {
    for (int q = e0; q < e1; ++q) {
      const CsrE<F>& e = v.csr[q];
      if (e.other < 0 && v.bf[e.bface].kind == kBcPeriodic) continue;
      const F a = x * e.sx + y * e.sy + z * e.sz;
      if (a > F(0)) {
        ap += a;
      } else if (e.other >= 0) {
        if (e.other < c0 || e.other >= c1) {
          if (prev != nullptr)
            Lout += a * prev[static_cast<i64>(e.other) * stride + t];
          else
            am_out -= a;
        } else if ((pos ? pos[e.other] : e.other) < pi) {
          L += a * cur[static_cast<i64>(e.other) * stride + t];
        }
      }
    }
  }
  const F D = v.cells[cell].diag + v.diag_scale * ap + v.interface_damping * am_out;
  cur[wi] = (-Ri - Lout - L) / D;
}

// one cell of the backward sweep (sweep_backward_cell, kinetic_kernels.hpp:328-358)
template <class F>
__device__ __forceinline__ void backward_cell(const View<F>& v, int cell, int t, int p, int c0, int c1, const int* pos,
                                              F x, F y, F z, int stride, F* cur, int has_prev, int update_f) {
  F ap = F(0), am_out = F(0), U = F(0);
  const int e0 = v.csr_off[cell], e1 = v.csr_off[cell + 1];
  const int pi = pos ? pos[cell] : cell;
  const i64 wi = static_cast<i64>(cell) * stride + t;
  const F own = cur[wi];
  if (e1 - e0 <= kGather) {
    F ak[kGather], nb[kGather];
#pragma unroll
    for (int k = 0; k < kGather; ++k)
      if (k < e1 - e0) {
        const CsrE<F>& e = v.csr[e0 + k];
        ak[k] = x * e.sx + y * e.sy + z * e.sz;
        nb[k] = F(0);
        if (e.other >= c0 && e.other < c1 && !(ak[k] > F(0)) && (pos ? pos[e.other] : e.other) > pi)
          nb[k] = cur[static_cast<i64>(e.other) * stride + t];
      }
#pragma unroll
    for (int k = 0; k < kGather; ++k)
      if (k < e1 - e0) {
        const CsrE<F>& e = v.csr[e0 + k];
        if (e.other < 0 && v.bf[e.bface].kind == kBcPeriodic) continue;
        const F a = ak[k];
        if (a > F(0)) {
          ap += a;
        } else if (e.other >= 0) {
          if (e.other < c0 || e.other >= c1) {
            if (!has_prev) am_out -= a;
          } else if ((pos ? pos[e.other] : e.other) > pi) {
            U += a * nb[k];
          }
        }
      }
  } else {
    for (int q = e0; q < e1; ++q) {
      const CsrE<F>& e = v.csr[q];
      if (e.other < 0 && v.bf[e.bface].kind == kBcPeriodic) continue;
      const F a = x * e.sx + y * e.sy + z * e.sz;
      if (a > F(0)) {
        ap += a;
      } else if (e.other >= 0) {
        if (e.other < c0 || e.other >= c1) {
          if (!has_prev) am_out -= a;
        } else if ((pos ? pos[e.other] : e.other) > pi) {
          U += a * cur[static_cast<i64>(e.other) * stride + t];
        }
      }
    }
  }
  const F D = v.cells[cell].diag + v.diag_scale * ap + v.interface_damping * am_out;
  const F df = own - U / D;
  cur[wi] = df;
  if (update_f) v.f[static_cast<i64>(cell) * v.nv + p] += df;
}

// n / d as each precision's sweep computes it: FP64 divides (correctly rounded, the reference's
// value); FP32 multiplies by the hardware reciprocal (v_rcp_f32, 1 ulp), two roundings instead of
// the ~10-instruction correctly rounded division sequence, within the FP32 form's tolerance.
__device__ __forceinline__ double sweep_div(double n, double d) { return n / d; }
__device__ __forceinline__ float sweep_div(float n, float d) { return n * __builtin_amdgcn_rcpf(d); }

__device__ __forceinline__ i64 uniform64(i64 x) {
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(x)));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(x >> 32)));
  return static_cast<i64>((static_cast<unsigned long long>(hi) << 32) | lo);
}
template <class F>
__device__ __forceinline__ F* byte_row(F* base, i64 off) {
  return reinterpret_cast<F*>(reinterpret_cast<char*>(base) + off);
}
template <class F>
__device__ __forceinline__ const F* byte_row(const F* base, i64 off) {
  return reinterpret_cast<const F*>(reinterpret_cast<const char*>(base) + off);
}

// One cell from its sweep record in LDS: the results of forward_cell. The record's integers are
// made wave-uniform (scalar registers), its areas are LDS operands. A neighbour row is loaded for
// the whole wave when any lane is upwind of that face (one scalar branch, no per-lane masks), and
// every lane adds every face, with max(a, 0) / min(a, 0) (times the neighbour value, which is
// finite) in place of the reference's branches: a face a lane skips adds +-0, which leaves every
// partial sum unchanged bit for bit (the sums start at +0 and never become -0). The upwind sums L and Lout run
// over their faces in CSR order as in the reference; the outflow sum ap runs over the record's
// grouped order, which can move it by an ulp. Unused slots have zero areas: a = +-0 adds +0 to ap
// and nothing else. toff = the lane velocity's byte offset in a chunk row. HasPrev: a pass after
// the first (out-of-block upwind values from the previous pass instead of the damping term).
template <bool HasPrev, class F>
__device__ __forceinline__ void forward_rec(const SweepRec<F>& r, F x, F y, F z, F diag_scale, F damping,
                                            unsigned toff, const F* __restrict__ rres, F* cur,
                                            const F* __restrict__ prev) {
  const int n_in = uniform(r.n_in), n_io = n_in + uniform(r.n_out);
  const i64 row = uniform64(r.row_w);
  const F Ri = row_load(byte_row(rres, row), toff);
  // max(a, 0) and min(a, 0) stand for "a > 0 ? a : 0" and "a > 0 ? 0 : a": they differ only in
  // giving -0 for a = -0, and +-0 added to a sum that is never -0 leaves it unchanged. A face whose
  // row is not loaded keeps nb = 0, so adding its am * nb = +-0 to an upwind sum changes nothing.
  F am[kGather], nb[kGather];
  F ap = F(0);
#pragma unroll
  for (int k = 0; k < kGather; ++k) {
    const F a = dot3(x, y, z, r.sx[k], r.sy[k], r.sz[k]);
    ap += nmax(a, F(0));
    am[k] = nmin(a, F(0));
    nb[k] = F(0);
    // one load site per face (in-block rows from this pass, out-of-block rows from the previous one)
    const bool load = HasPrev ? k < n_io : k < n_in;
    const F* base = (!HasPrev || k < n_in) ? static_cast<const F*>(cur) : prev;
    if (load && __builtin_amdgcn_ballot_w64(a < F(0)) != 0) nb[k] = row_load(byte_row(base, uniform64(r.nb_w[k])), toff);
  }
  F am_out = F(0), L = F(0), Lout = F(0);
#pragma unroll
  for (int k = 0; k < kGather; ++k) {
    if (HasPrev) {
      if (k < n_in)
        L = madd(L, am[k], nb[k]);
      else
        Lout = madd(Lout, am[k], nb[k]);
    } else {
      L = madd(L, am[k], nb[k]);
      if (k >= n_in && k < n_io) am_out -= am[k];
    }
  }
  const F D = madd(madd(r.diag, diag_scale, ap), damping, am_out);
  row_ref(byte_row(cur, row), toff) = sweep_div(-Ri - Lout - L, D);
}

// One cell from its sweep record in LDS: the results of backward_cell (same scheme as
// forward_rec). poff = the lane velocity's byte offset in an f row.
template <bool HasPrev, class F>
__device__ __forceinline__ void backward_rec(const SweepRec<F>& r, F x, F y, F z, F diag_scale, F damping,
                                             unsigned toff, unsigned poff, F* cur, F* f, int update_f) {
  const int n_in = uniform(r.n_in), n_io = n_in + uniform(r.n_out);
  const i64 row = uniform64(r.row_w);
  const F own = row_load(byte_row(cur, row), toff);
  // the cell's f is read with its own value, so that load overlaps the neighbour loads instead of
  // following the store of df (no other cell writes this (cell, velocity) during the sweep)
  const F f_old = update_f ? row_load(byte_row(static_cast<const F*>(f), uniform64(r.row_f)), poff) : F(0);
  F am[kGather], nb[kGather];  // as in forward_rec
  F ap = F(0);
#pragma unroll
  for (int k = 0; k < kGather; ++k) {
    const F a = dot3(x, y, z, r.sx[k], r.sy[k], r.sz[k]);
    ap += nmax(a, F(0));
    am[k] = nmin(a, F(0));
    nb[k] = F(0);
    if (k < n_in && __builtin_amdgcn_ballot_w64(a < F(0)) != 0) nb[k] = row_load(byte_row(cur, uniform64(r.nb_w[k])), toff);
  }
  F am_out = F(0), U = F(0);
#pragma unroll
  for (int k = 0; k < kGather; ++k) {
    U = madd(U, am[k], nb[k]);
    if (!HasPrev && k >= n_in && k < n_io) am_out -= am[k];
  }
  const F D = madd(madd(r.diag, diag_scale, ap), damping, am_out);
  const F df = own - sweep_div(U, D);
  row_ref(byte_row(cur, row), toff) = df;
  if (update_f) row_ref(byte_row(f, uniform64(r.row_f)), poff) = f_old + df;
}

// Workgroup -> (direction, block, velocity tile). The (direction, block) pairs are dealt to the 8
// XCDs (pair P on XCD P mod 8) and all velocity tiles of a pair run on that XCD, so the pair's
// records are fetched into one L2 and shared by its tiles. Grid: 8 * ceil(n_pairs / 8) * tiles.
__device__ __forceinline__ bool sweep_tile(int n_pairs, int tiles, int n_blocks, int& g, int& b, int& tile) {
  const int x = static_cast<int>(blockIdx.x % 8);
  const int k = static_cast<int>(blockIdx.x / 8);
  const int pair = x + 8 * (k / tiles);
  if (pair >= n_pairs) return false;
  tile = k % tiles;
  g = pair / n_blocks;
  b = pair % n_blocks;
  return true;
}

struct SweepArgs {
  int k0, stride, n_blocks, n_pairs, tiles;
  const int* block_off;    // block cell ranges
  const int* vel_list;     // storage positions grouped by direction
  const int* stage_off;    // first record of each stage, plus the end
  const int* group_stage;  // first stage of each (direction, block), plus the end
  VelRange vr;
};

// Stage s's records (contiguous in global memory) -> an LDS stage buffer, with direct
// global-to-LDS loads (each wave moves 64 words per instruction; no registers involved). The
// loads complete under the vector memory counter, which the stage-end barrier drains.
#define GSK_STAGE_ISSUE(recs, stage_off, s, buf)                                                     \
  do {                                                                                              \
    const int* src_ = reinterpret_cast<const int*>(recs) + static_cast<i64>((stage_off)[s]) * kRecWords<F>; \
    const int nw_ = ((stage_off)[(s) + 1] - (stage_off)[s]) * kRecWords<F>;                          \
    for (int base_ = wave * 64; base_ < nw_; base_ += nwaves * 64)                                   \
      if (base_ + lane < nw_) __builtin_amdgcn_global_load_lds(src_ + base_ + lane, (buf) + base_, 4, 0, 0); \
  } while (0)

// The workgroup walks the stages of its (direction, block); while it works on one stage from LDS,
// the next stage's records are in flight to registers and go to the other LDS buffer before the
// stage-end barrier. Records come from LDS (not through the vector memory counter), so a cell's
// only memory round trip is its neighbour gather, which overlaps the previous cell's store. The
// waves take every nwaves-th cell of a stage; lanes = velocities of one direction.
template <class F, bool HasPrev>
__global__ __attribute__((amdgpu_waves_per_eu(6))) void k_sweep_forward(View<F> v, SweepArgs a, const SweepRec<F>* __restrict__ recs,
                                const F* __restrict__ rres, F* cur, const F* __restrict__ prev) {
  int g, b, tile;
  if (!sweep_tile(a.n_pairs, a.tiles, a.n_blocks, g, b, tile)) return;
  const int vb = a.vr.beg[g] + tile * 64;
  if (vb >= a.vr.end[g]) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int nwaves = blockDim.x >> 6;
  const bool valid = vb + lane < a.vr.end[g];
  const int p = a.vel_list[valid ? vb + lane : vb];
  const int t = p - a.k0;
  const unsigned toff = static_cast<unsigned>(t) * sizeof(F), poff = static_cast<unsigned>(p) * sizeof(F);
  F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
  settle(x, y, z);
  const int c0 = a.block_off[b], c1 = a.block_off[b + 1];
  const int* pos = v.n_sweep_dirs > 0 ? v.sweep_pos + static_cast<i64>(g) * v.n_cells : nullptr;
  __shared__ __align__(16) int lds[2][kStageWords];
  const int s0 = a.group_stage[g * a.n_blocks + b], s1 = a.group_stage[g * a.n_blocks + b + 1];
  if (s0 < s1) GSK_STAGE_ISSUE(recs, a.stage_off, s0, lds[0]);
  __syncthreads();
  for (int s = s0; s < s1; ++s) {
    const int* buf = lds[(s - s0) & 1];
    if (s + 1 < s1) GSK_STAGE_ISSUE(recs, a.stage_off, s + 1, lds[(s + 1 - s0) & 1]);
    const int n = a.stage_off[s + 1] - a.stage_off[s];
    for (int i = wave; i < n; i += nwaves) {
      const SweepRec<F>& r = reinterpret_cast<const SweepRec<F>*>(buf)[i];
      const int nf = uniform(r.nf);
      if (!valid) continue;
      if (nf >= 0)
        forward_rec<HasPrev>(r, x, y, z, v.diag_scale, v.interface_damping, toff, rres, cur, prev);
      else
        forward_cell(v, uniform(r.cell), t, c0, c1, pos, x, y, z, a.stride, rres, cur, prev);
    }
    __syncthreads();
  }
}

template <class F, bool HasPrev>
__global__ __attribute__((amdgpu_waves_per_eu(6))) void k_sweep_backward(View<F> v, SweepArgs a, const SweepRec<F>* __restrict__ recs, F* cur,
                                 int update_f) {
  constexpr int has_prev = HasPrev ? 1 : 0;
  int g, b, tile;
  if (!sweep_tile(a.n_pairs, a.tiles, a.n_blocks, g, b, tile)) return;
  const int vb = a.vr.beg[g] + tile * 64;
  if (vb >= a.vr.end[g]) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int nwaves = blockDim.x >> 6;
  const bool valid = vb + lane < a.vr.end[g];
  const int p = a.vel_list[valid ? vb + lane : vb];
  const int t = p - a.k0;
  const unsigned toff = static_cast<unsigned>(t) * sizeof(F), poff = static_cast<unsigned>(p) * sizeof(F);
  F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
  settle(x, y, z);
  const int c0 = a.block_off[b], c1 = a.block_off[b + 1];
  const int* pos = v.n_sweep_dirs > 0 ? v.sweep_pos + static_cast<i64>(g) * v.n_cells : nullptr;
  __shared__ __align__(16) int lds[2][kStageWords];
  const int s0 = a.group_stage[g * a.n_blocks + b], s1 = a.group_stage[g * a.n_blocks + b + 1];
  if (s0 < s1) GSK_STAGE_ISSUE(recs, a.stage_off, s0, lds[0]);
  __syncthreads();
  for (int s = s0; s < s1; ++s) {
    const int* buf = lds[(s - s0) & 1];
    if (s + 1 < s1) GSK_STAGE_ISSUE(recs, a.stage_off, s + 1, lds[(s + 1 - s0) & 1]);
    const int n = a.stage_off[s + 1] - a.stage_off[s];
    for (int i = wave; i < n; i += nwaves) {
      const SweepRec<F>& r = reinterpret_cast<const SweepRec<F>*>(buf)[i];
      const int nf = uniform(r.nf);
      if (!valid) continue;
      if (nf >= 0)
        backward_rec<HasPrev>(r, x, y, z, v.diag_scale, v.interface_damping, toff, poff, cur, v.f, update_f);
      else
        backward_cell(v, uniform(r.cell), t, p, c0, c1, pos, x, y, z, a.stride, cur, has_prev, update_f);
    }
    __syncthreads();
  }
}

// GSIS correction f <- f - M(W_half) + M(W_new) with the negative-value rules (k_gsis_correction)
template <class F>
__global__ __launch_bounds__(kBlock) void k_gsis_correction(View<F> v, const double* __restrict__ w_half,
                                                            const double* __restrict__ w_new, int mode,
                                                            double* __restrict__ shift, int* __restrict__ flags) {
  __shared__ double lds[kWaves * kDeqSums];
  __shared__ F ldsf[kWaves];
  for (int c = blockIdx.x; c < v.n_cells; c += gridDim.x) {
    const double* wh = w_half + static_cast<i64>(c) * 5;
    const double* wn = w_new + static_cast<i64>(c) * 5;
    int rh = 0, rn = 0;
    const Maxw<F> mh = equilibrium_params<F>(v, wh[0], wh[1], wh[2], wh[3], wh[4], &rh, lds);
    const Maxw<F> mn = equilibrium_params<F>(v, wn[0], wn[1], wn[2], wn[3], wn[4], &rn, lds);
    if (v.deq_fallback && (rh < 0 || rn < 0)) {
      if (threadIdx.x == 0) {
        flags[c] = 2;
        shift[c * 3 + 0] = wh[1];
        shift[c * 3 + 1] = wh[2];
        shift[c * 3 + 2] = wh[3];
      }
      continue;
    }
    F* fc = v.f + static_cast<i64>(c) * v.nv;
    const double threshold = -1e-8 * wn[0] / (wn[4] * ::sqrt(wn[4]));
    const F thr = static_cast<F>(threshold);
    bool replace = false;
    if (mode == 2) {
      F alpha = F(1);
      for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
        const F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
        alpha = dmin(alpha, correction_alpha_limit(fc[p], maxw_eval(mn, x, y, z) - maxw_eval(mh, x, y, z), thr));
      }
      alpha = block_extreme<F, false>(alpha, ldsf);
      for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
        const F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
        fc[p] += alpha * (maxw_eval(mn, x, y, z) - maxw_eval(mh, x, y, z));
      }
      replace = alpha < F(1);
    } else if (mode == 3) {
      F theta = F(0);
      for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
        const F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
        const F e = maxw_eval(mn, x, y, z);
        theta = dmax(theta, blend_theta(fc[p], e - maxw_eval(mh, x, y, z), e, thr));
      }
      theta = block_extreme<F, true>(theta, ldsf);
      for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
        const F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
        const F e = maxw_eval(mn, x, y, z);
        const F g = fc[p] + (e - maxw_eval(mh, x, y, z));
        fc[p] = (F(1) - theta) * g + theta * e;
      }
      replace = theta > F(0);
    } else {
      F fmin = static_cast<F>(3.0e38);
      bool any = false;
      for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
        const F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
        const F old = fc[p];
        const F xx = old + (maxw_eval(mn, x, y, z) - maxw_eval(mh, x, y, z));
        fc[p] = xx;
        const F chk = negative_check_value(old, xx, mode == 1);
        fmin = any ? dmin(fmin, chk) : chk;
        any = true;
      }
      if (!any) fmin = static_cast<F>(3.0e38);
      fmin = block_extreme<F, false>(fmin, ldsf);
      replace = static_cast<double>(fmin) < threshold;
      if (replace)
        for (int p = threadIdx.x; p < v.nv; p += blockDim.x) fc[p] = maxw_eval(mn, v.xi_x[p], v.xi_y[p], v.xi_z[p]);
    }
    if (threadIdx.x == 0) {
      flags[c] = replace ? 1 : 0;
      shift[c * 3 + 0] = wn[1];
      shift[c * 3 + 1] = wn[2];
      shift[c * 3 + 2] = wn[3];
    }
    __syncthreads();
  }
}

// number of nonzero flags, one workgroup, fixed order (the replaced-cell count of a GSIS correction)
__global__ __launch_bounds__(kBlock) void k_count_nonzero(const int* __restrict__ flags, int n,
                                                          long long* __restrict__ out) {
  __shared__ long long lds[kBlock];
  long long c = 0;
  for (int i = threadIdx.x; i < n; i += kBlock) c += flags[i] != 0 ? 1 : 0;
  lds[threadIdx.x] = c;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if (static_cast<int>(threadIdx.x) < h) lds[threadIdx.x] += lds[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = lds[0];
}

template <class F>
__global__ void k_from_double(const double* __restrict__ src, F* __restrict__ dst, i64 n) {
  for (i64 i = blockIdx.x * static_cast<i64>(blockDim.x) + threadIdx.x; i < n; i += static_cast<i64>(gridDim.x) * blockDim.x)
    dst[i] = static_cast<F>(src[i]);
}
template <class F>
__global__ void k_to_double(const F* __restrict__ src, double* __restrict__ dst, i64 n) {
  for (i64 i = blockIdx.x * static_cast<i64>(blockDim.x) + threadIdx.x; i < n; i += static_cast<i64>(gridDim.x) * blockDim.x)
    dst[i] = static_cast<double>(src[i]);
}

}  // namespace dev
}  // namespace gsk

#include "kinetic_adjoint.hpp"
#include "kinetic_control.hpp"

namespace gsk {

// ==============================================================================================
// Host-side backend
// ==============================================================================================
namespace {

void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) fail_at(file, line, strf("HIP error in %s: %s", what, hipGetErrorString(e)));
}
#define HIPCHK(x) hip_check((x), #x, __FILE__, __LINE__)

// Device address of the fixed scratch array on the current device, looked up once per device
// (each GPU holds its own copy of the array at its own address).
void* noalloc_scratch() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* ptrs[kMaxDevices] = {};
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  GSK_CHECK(dev >= 0 && dev < kMaxDevices, "device ordinal out of range for the fixed scratch table");
  std::lock_guard<std::mutex> lock(mu);
  if (ptrs[dev] == nullptr) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    GSK_CHECK(p != nullptr, "the fixed scratch array has no device address");
    ptrs[dev] = p;
  }
  return ptrs[dev];
}

// owning device buffer
template <class T>
struct DBuf {
  T* p = nullptr;
  std::size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(std::size_t count) {
    release();
    n = count;
    if (count == 0) return;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw Error(strf("device allocation of %zu bytes failed (%s)", count * sizeof(T), hipGetErrorString(e)));
    }
  }
  void upload(const T* src, std::size_t count, hipStream_t s) {
    if (count != n) alloc(count);
    if (count) HIPCHK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  std::size_t bytes() const { return n * sizeof(T); }
};

enum Phase { kMoments, kBoundary, kGradient, kResidual, kForward, kBackward, kOther, kNumPhases };

template <class F>
class HipBackend final : public Backend {
 public:
  explicit HipBackend(const gsk_setup_view& s) {
    HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    scratch_ = noalloc_scratch();
    nc_ = s.n_cells;
    nv_ = s.nv;
    nb_ = s.n_bfaces;
    n_blocks_ = s.n_blocks;
    block_passes_ = s.block_passes;
    negative_mode_ = s.negative_check;
    mu0_ = s.mu0;
    omega_ = s.omega;
    Pr_ = s.Pr;
    GSK_CHECK(nc_ > 0 && nv_ > 0 && n_blocks_ >= 1, "setup view is empty");
    GSK_CHECK(s.orbit_size >= 1 && nv_ % s.orbit_size == 0, "velocity count is not a multiple of the orbit size");
    // velocity tables
    std::vector<F> xf(static_cast<std::size_t>(nv_)), yf(xf.size()), zf(xf.size());
    for (int p = 0; p < nv_; ++p) {
      xf[static_cast<std::size_t>(p)] = static_cast<F>(s.xi_x[p]);
      yf[static_cast<std::size_t>(p)] = static_cast<F>(s.xi_y[p]);
      zf[static_cast<std::size_t>(p)] = static_cast<F>(s.xi_z[p]);
    }
    xi_x_.upload(xf.data(), xf.size(), stream_);
    xi_y_.upload(yf.data(), yf.size(), stream_);
    xi_z_.upload(zf.data(), zf.size(), stream_);
    xd_.upload(s.xi_x, static_cast<std::size_t>(nv_), stream_);
    if (s.weight_rel != nullptr) {  // stretched velocity grid
      std::vector<F> wq(static_cast<std::size_t>(nv_));
      for (int p = 0; p < nv_; ++p) wq[static_cast<std::size_t>(p)] = static_cast<F>(s.weight_rel[p]);
      wq_.upload(wq.data(), wq.size(), stream_);
      wqd_.upload(s.weight_rel, static_cast<std::size_t>(nv_), stream_);
      GSK_CHECK(s.grid_nodes != nullptr, "a stretched velocity grid needs its node coordinates");
      nodes_.upload(s.grid_nodes, static_cast<std::size_t>(s.grid_n[0]) + s.grid_n[1] + s.grid_n[2], stream_);
    }
    yd_.upload(s.xi_y, static_cast<std::size_t>(nv_), stream_);
    zd_.upload(s.xi_z, static_cast<std::size_t>(nv_), stream_);
    // mesh tables
    const int n_entries = s.csr_offsets[nc_];
    for (int c = 0; c < nc_; ++c) max_faces_ = std::max(max_faces_, s.csr_offsets[c + 1] - s.csr_offsets[c]);
    csr_off_.upload(s.csr_offsets, static_cast<std::size_t>(nc_) + 1, stream_);
    std::vector<dev::CsrE<F>> csr(static_cast<std::size_t>(n_entries));
    for (int k = 0; k < n_entries; ++k) {
      const gsk_csr_entry& e = s.csr[k];
      dev::CsrE<F>& o = csr[static_cast<std::size_t>(k)];
      o.other = e.other;
      o.bface = e.bface;
      o.sx = static_cast<F>(e.s[0]);
      o.sy = static_cast<F>(e.s[1]);
      o.sz = static_cast<F>(e.s[2]);
      o.dsx = static_cast<F>(e.ds[0]);
      o.dsy = static_cast<F>(e.ds[1]);
      o.dsz = static_cast<F>(e.ds[2]);
      o.dox = static_cast<F>(e.dother[0]);
      o.doy = static_cast<F>(e.dother[1]);
      o.doz = static_cast<F>(e.dother[2]);
      o.cx = static_cast<F>(e.c[0]);
      o.cy = static_cast<F>(e.c[1]);
      o.cz = static_cast<F>(e.c[2]);
      GSK_CHECK((e.other >= 0 && e.other < nc_) || (e.other < 0 && e.bface >= 0 && e.bface < nb_),
                "CSR entry refers to a cell or face out of range");
    }
    csr_.upload(csr.data(), csr.size(), stream_);
    // adjoint tables: the slot of every internal entry's reverse entry in the neighbour's row, each
    // boundary face's slot in its owner's row, and the owners of boundary faces
    adj_mate_host_.assign(static_cast<std::size_t>(n_entries), -1);
    adj_bslot_host_.assign(static_cast<std::size_t>(nb_), -1);
    for (int c = 0; c < nc_; ++c)
      for (int k = s.csr_offsets[c]; k < s.csr_offsets[c + 1]; ++k) {
        const gsk_csr_entry& e = s.csr[k];
        if (e.other < 0) {
          adj_bslot_host_[static_cast<std::size_t>(e.bface)] = k - s.csr_offsets[c];
          continue;
        }
        int best = -1;
        double best_d = 0.0;
        for (int m = s.csr_offsets[e.other]; m < s.csr_offsets[e.other + 1]; ++m) {
          if (s.csr[m].other != c) continue;
          const double d = std::fabs(s.csr[m].s[0] + e.s[0]) + std::fabs(s.csr[m].s[1] + e.s[1]) +
                           std::fabs(s.csr[m].s[2] + e.s[2]);
          if (best < 0 || d < best_d) {
            best = m - s.csr_offsets[e.other];
            best_d = d;
          }
        }
        GSK_CHECK(best >= 0, "an internal face has no reverse entry in its neighbour's row");
        adj_mate_host_[static_cast<std::size_t>(k)] = best;
      }
    for (int b = 0; b < nb_; ++b)
      GSK_CHECK(adj_bslot_host_[static_cast<std::size_t>(b)] >= 0, "a boundary face is missing from its owner's row");
    std::vector<F> eps2(static_cast<std::size_t>(nc_));
    for (int c = 0; c < nc_; ++c) eps2[static_cast<std::size_t>(c)] = static_cast<F>(s.eps2[c]);
    eps2_.upload(eps2.data(), eps2.size(), stream_);
    volume_.upload(s.volume, static_cast<std::size_t>(nc_), stream_);
    vdt_.upload(s.vdt, static_cast<std::size_t>(nc_), stream_);
    std::vector<dev::BFace<F>> bf(static_cast<std::size_t>(nb_) + 1);  // + the padding face
    for (int b = 0; b < nb_; ++b) {
      const gsk_bface& i = s.bfaces[b];
      dev::BFace<F>& o = bf[static_cast<std::size_t>(b)];
      o.owner = i.owner;
      o.kind = i.kind;
      o.mirror_xor = i.mirror_xor;
      o.patch = i.patch;
      o.nx = static_cast<F>(i.n[0]);
      o.ny = static_cast<F>(i.n[1]);
      o.nz = static_cast<F>(i.n[2]);
      o.area = static_cast<F>(i.area);
      o.dx = static_cast<F>(i.d[0]);
      o.dy = static_cast<F>(i.d[1]);
      o.dz = static_cast<F>(i.d[2]);
      GSK_CHECK(i.owner >= 0 && i.owner < nc_, "boundary face owner out of range");
      GSK_CHECK(i.mirror_xor >= 0 && i.mirror_xor < nv_, "boundary face mirror out of range");
    }
    bf[static_cast<std::size_t>(nb_)].kind = dev::kBcPeriodic;
    bfaces_.upload(bf.data(), bf.size(), stream_);
    host_bface_owner_.resize(static_cast<std::size_t>(nb_));
    for (int b = 0; b < nb_; ++b) host_bface_owner_[static_cast<std::size_t>(b)] = s.bfaces[b].owner;
    host_bface_kind_.resize(static_cast<std::size_t>(nb_));
    for (int b = 0; b < nb_; ++b) host_bface_kind_[static_cast<std::size_t>(b)] = s.bfaces[b].kind;
    inflow_.alloc(static_cast<std::size_t>(nb_));
    wall_scale_.alloc(static_cast<std::size_t>(nb_));
    wall_negative_.alloc(1);
    HIPCHK(hipMemsetAsync(wall_negative_.p, 0, sizeof(int), stream_));
    if (nb_ > 0) HIPCHK(hipMemsetAsync(wall_scale_.p, 0, wall_scale_.bytes(), stream_));
    HIPCHK(hipMemsetAsync(inflow_.p, 0, inflow_.bytes(), stream_));
    block_off_.upload(s.block_offsets, static_cast<std::size_t>(n_blocks_) + 1, stream_);
    n_sweep_dirs_ = s.n_sweep_dirs;
    if (n_sweep_dirs_ > 0) {
      sweep_order_.upload(s.sweep_order, static_cast<std::size_t>(n_sweep_dirs_) * nc_, stream_);
      sweep_pos_.upload(s.sweep_pos, static_cast<std::size_t>(n_sweep_dirs_) * nc_, stream_);
      sweep_dir_of_.upload(s.sweep_dir_of, static_cast<std::size_t>(nv_), stream_);
    }
    // Maxwell walls
    n_mw_ = s.n_maxwell_faces;
    moments_mfma_ = s.moments_mfma != 0;
    GSK_CHECK(!moments_mfma_ || (std::is_same_v<F, float>), "matrix-core moment sums are for the FP32 form");
    GSK_CHECK(n_mw_ == 0 || nb_ <= 65535, "Maxwell-wall reflection launches take at most 65535 boundary faces");
    if (n_mw_ > 0) {
      mw_off_.upload(reinterpret_cast<const long long*>(s.maxwell_offset), static_cast<std::size_t>(nb_), stream_);
      std::vector<F> frac(static_cast<std::size_t>(nb_));
      for (int b = 0; b < nb_; ++b) frac[static_cast<std::size_t>(b)] = static_cast<F>(s.maxwell_specular[b]);
      mw_frac_.upload(frac.data(), frac.size(), stream_);
      specular_.alloc(static_cast<std::size_t>(n_mw_) * nv_);
      specular_first_.alloc(static_cast<std::size_t>(n_mw_) * nv_);
      HIPCHK(hipMemsetAsync(specular_.p, 0, specular_.bytes(), stream_));
      HIPCHK(hipMemsetAsync(specular_first_.p, 0, specular_first_.bytes(), stream_));
      grid_to_storage_.upload(s.grid_to_storage, static_cast<std::size_t>(nv_), stream_);
    }
    upload_axes(s);
    for (int a = 0; a < 3; ++a) {
      grid_n_[a] = s.grid_n[a];
      deq_.lo[a] = s.deq_lo[a];
      deq_.hi[a] = s.deq_hi[a];
      deq_.d[a] = s.deq_d[a];
    }
    deq_.criterion = s.deq_criterion;
    deq_.max_iterations = s.deq_max_iterations;
    deq_.tolerance = s.deq_tolerance;
    discrete_eq_ = s.discrete_eq;
    deq_fallback_ = s.deq_fallback_collisionless;
    cold_bgk_ = s.cold_bgk_temperature;
    limiter_mode_ = s.limiter_mode;
    positivity_ = s.positivity;
    diag_scale_ = s.diag_scale;
    interface_damping_ = s.interface_damping;
    weight_ = s.weight;
    // state
    const std::size_t nf = static_cast<std::size_t>(nc_) * static_cast<std::size_t>(nv_);
    // fit check before the large allocations (kinetic_backend_cuda.cpp:103-106)
    std::size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    GSK_CHECK(nf * sizeof(F) < free_b, strf("f needs %.2f GB but only %.2f GB of device memory is free",
                                            nf * sizeof(F) / 1e9, free_b / 1e9));
    f_.alloc(nf);
    HIPCHK(hipMemsetAsync(f_.p, 0, f_.bytes(), stream_));
    cells_.alloc(static_cast<std::size_t>(nc_));
    HIPCHK(hipMemsetAsync(cells_.p, 0, cells_.bytes(), stream_));
    moments_.alloc(static_cast<std::size_t>(nc_) * dev::kMoments);
    HIPCHK(hipMemsetAsync(moments_.p, 0, moments_.bytes(), stream_));
    shift_.alloc(static_cast<std::size_t>(nc_) * 3);
    HIPCHK(hipMemsetAsync(shift_.p, 0, shift_.bytes(), stream_));
    flags_.alloc(static_cast<std::size_t>(nc_));
    HIPCHK(hipMemsetAsync(flags_.p, 0, flags_.bytes(), stream_));
    residual_.alloc(static_cast<std::size_t>(nc_));
    bsums_.alloc(static_cast<std::size_t>(nb_) * dev::kBoundarySums + 1);
    bsum_chunks_ = std::max(1, std::min(16, (nv_ + 2047) / 2048));
    bsum_part_.alloc(static_cast<std::size_t>(nb_) * bsum_chunks_ * 2 * dev::kSideSums + 1);
    // chunk width: requested (rounded up to the orbit size), or as wide as 6 workspace arrays fit
    const int orbit = s.orbit_size;
    long long k = s.velocity_chunk > 0 ? s.velocity_chunk : nv_;
    k = (k + orbit - 1) / orbit * orbit;
    k = std::min<long long>(k, nv_);
    if (s.velocity_chunk == 0) {
      HIPCHK(hipMemGetInfo(&free_b, &total_b));
      const double budget = 0.85 * static_cast<double>(free_b);
      const double per_k = 6.0 * nc_ * sizeof(F);
      while (k > orbit && (per_k * k > budget || work_items(k) > kMaxWorkItems))
        k = std::max<long long>(orbit, (k / 2 + orbit - 1) / orbit * orbit);
    }
    if (work_items(k) > kMaxWorkItems)
      throw std::runtime_error("gradient / residual work items (" + std::to_string(work_items(k)) +
                               ") exceed 2^32 - 1: use a narrower [kinetic].velocity_chunk");
    chunk_ = static_cast<int>(k);
    tiles_ = (chunk_ + dev::kBlock - 1) / dev::kBlock;
    // padded face rows (kGather faces per cell; padding faces are skipped like periodic ones); cells
    // with more faces are listed for the CSR path
    {
      std::vector<dev::EllRow<F>> ell(static_cast<std::size_t>(nc_));
      std::vector<int> irregular;
      // test hook: GSK_ELL_MAX_FACES lowers the face count above which a cell takes the CSR path
      int ell_max = dev::kGather;
      if (const char* env = std::getenv("GSK_ELL_MAX_FACES")) ell_max = std::max(0, std::min(dev::kGather, std::atoi(env)));
      if (const char* env = std::getenv("GSK_GRID_CAP")) grid_cap_ = std::max(8LL, std::min(grid_cap_, std::atoll(env)) / 8 * 8);
      for (int c = 0; c < nc_; ++c) {
        const int e0 = s.csr_offsets[c], e1 = s.csr_offsets[c + 1];
        const bool csr_path = e1 - e0 > ell_max;  // row left all padding; the CSR-path kernels overwrite it
        if (csr_path) irregular.push_back(c);
        dev::EllRow<F>& r = ell[static_cast<std::size_t>(c)];
        r.skip = r.bnd = 0;
        for (int j = 0; j < dev::kGather; ++j) {
          const int k = e0 + j;
          if (csr_path || k >= e1) {  // padding: the cell's own rows, zero coefficients, never read
            r.nbf[j] = static_cast<long long>(c) * nv_;
            r.nbw[j] = static_cast<long long>(c) * chunk_;
            r.bface[j] = -1;
            r.skip |= 1u << j;
            r.sx[j] = r.sy[j] = r.sz[j] = r.dsx[j] = r.dsy[j] = r.dsz[j] = F(0);
            r.dox[j] = r.doy[j] = r.doz[j] = r.cx[j] = r.cy[j] = r.cz[j] = F(0);
            continue;
          }
          const dev::CsrE<F>& e = csr[static_cast<std::size_t>(k)];
          const int src = e.other >= 0 ? e.other : c;
          r.nbf[j] = static_cast<long long>(src) * nv_;
          r.nbw[j] = static_cast<long long>(src) * chunk_;
          r.bface[j] = e.other >= 0 ? -1 : e.bface;
          if (e.other < 0) {
            r.bnd |= 1u << j;
            if (s.bfaces[e.bface].kind == dev::kBcPeriodic) r.skip |= 1u << j;
          }
          r.sx[j] = e.sx, r.sy[j] = e.sy, r.sz[j] = e.sz;
          r.dsx[j] = e.dsx, r.dsy[j] = e.dsy, r.dsz[j] = e.dsz;
          r.dox[j] = e.dox, r.doy[j] = e.doy, r.doz[j] = e.doz;
          r.cx[j] = e.cx, r.cy[j] = e.cy, r.cz[j] = e.cz;
        }
      }
      ell_.upload(ell.data(), ell.size(), stream_);
      n_irregular_ = static_cast<int>(irregular.size());
      irregular_.upload(irregular.data(), irregular.size(), stream_);
    }
    build_levels(s);  // after chunk_: the sweep records hold row offsets for the chunk width
    const std::size_t work = static_cast<std::size_t>(nc_) * static_cast<std::size_t>(chunk_);
    gx_.alloc(work);
    gy_.alloc(work);
    gz_.alloc(work);
    rres_.alloc(work);
    bufA_.alloc(work);
    bufB_.alloc(work);
    partial_.alloc(static_cast<std::size_t>(nc_) * tiles_);
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&pinned_), pinned_bytes(), hipHostMallocDefault));
    for (int i = 0; i < kNumPhases; ++i) timings_[i] = 0.0;
    HIPCHK(hipStreamSynchronize(stream_));
  }

  ~HipBackend() override {
    for (auto& e : events_) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
    for (auto& e : free_events_) (void)hipEventDestroy(e);
    if (pinned_) (void)hipHostFree(pinned_);
    if (stream_) {
      (void)hipStreamSynchronize(stream_);
      (void)hipStreamDestroy(stream_);
    }
  }

  const char* name() const override { return sizeof(F) == 8 ? "hip-fp64" : "hip-fp32"; }
  int chunk_width() const override { return chunk_; }
  std::size_t memory_bytes() const override {
    return f_.bytes() + gx_.bytes() + gy_.bytes() + gz_.bytes() + rres_.bytes() + bufA_.bytes() + bufB_.bytes() +
           csr_.bytes() + ell_.bytes() + moments_.bytes() + specular_.bytes() + specular_first_.bytes() + partial_.bytes() +
           cells_.bytes() + fw_recs_.bytes() + bw_recs_.bytes() + bsum_part_.bytes() +
           (n_sweep_dirs_ > 0 ? sweep_order_.bytes() + sweep_pos_.bytes() : 0);
  }

  void set_equilibrium(const double* states, std::size_t n) override {
    GSK_CHECK(n == static_cast<std::size_t>(nc_) * 5, "states length differs from 5 x cells");
    begin(kOther);
    DBuf<double> st;
    st.upload(states, n, stream_);
    dev::k_set_equilibrium<F><<<grid_cells(), dev::kBlock, 0, stream_>>>(view(), st.p, shift_.p);
    HIPCHK(hipGetLastError());
    end();
    HIPCHK(hipStreamSynchronize(stream_));
  }

  void set_distribution(const double* f, std::size_t n) override {
    GSK_CHECK(n == f_.n, "distribution length differs from cells x velocities");
    if (sizeof(F) == 8) {
      HIPCHK(hipMemcpyAsync(f_.p, f, n * sizeof(double), hipMemcpyHostToDevice, stream_));
    } else {
      convert_in(f, n);
    }
    HIPCHK(hipStreamSynchronize(stream_));
  }

  void get_distribution(double* f, std::size_t n) override {
    GSK_CHECK(n == f_.n, "distribution length differs from cells x velocities");
    if (sizeof(F) == 8) {
      HIPCHK(hipMemcpyAsync(f, f_.p, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    } else {
      convert_out(f, n);
    }
    HIPCHK(hipStreamSynchronize(stream_));
  }

  void compute_moments(double* out, std::size_t n) override {
    GSK_CHECK(n == moments_.n, "moments output length differs from 15 x cells");
    launch_moments();
    download(out, moments_.p, n);
  }

  // the 13 shifted sums -> moments_ (and the shift), timed as the moment phase
  void launch_moments() {
    begin(kMoments);
    if constexpr (std::is_same_v<F, float>) {
      if (moments_mfma_)
        dev::k_moments_mfma<<<static_cast<unsigned>((nc_ + 15) / 16), 256, 0, stream_>>>(view(), shift_.p, moments_.p);
      else
        dev::k_moments<F><<<grid_cells(), dev::kBlock, 0, stream_>>>(view(), xd_.p, yd_.p, zd_.p, shift_.p, moments_.p);
    } else {
      dev::k_moments<F><<<grid_cells(), dev::kBlock, 0, stream_>>>(view(), xd_.p, yd_.p, zd_.p, shift_.p, moments_.p);
    }
    HIPCHK(hipGetLastError());
    end();
  }

  void scale_distribution(double s) override {
    begin(kOther);
    dev::k_scale_f<F><<<4096, 256, 0, stream_>>>(f_.p, static_cast<long long>(f_.n), static_cast<F>(s));
    dev::k_scale_moments<<<(nc_ + 255) / 256, 256, 0, stream_>>>(moments_.p, nc_, s);
    HIPCHK(hipGetLastError());
    end();
  }

  void update_cell_params() override {
    begin(kOther);
    dev::k_cell_params<F><<<static_cast<unsigned>(std::min((nc_ + dev::kWaves - 1) / dev::kWaves, 1 << 20)), dev::kBlock, 0,
                            stream_>>>(view(), moments_.p, volume_.p, vdt_.p, mu0_, omega_, Pr_,
                                                                     cells_.p, flags_.p);
    for (DBuf<dev::SweepRec<F>>* recs : {&fw_recs_, &bw_recs_}) {
      const int nr = static_cast<int>(recs->n);
      if (nr > 0) dev::k_record_diag<F><<<(nr + 255) / 256, 256, 0, stream_>>>(cells_.p, recs->p, nr);
    }
    HIPCHK(hipGetLastError());
    end();
    deq_flags_.resize(static_cast<std::size_t>(nc_));
    download(deq_flags_.data(), flags_.p, deq_flags_.size());
  }

  gsk_deq_counts deq_counts() override {
    gsk_deq_counts out{0, 0, 0};
    if (!discrete_eq_) return out;
    for (const int x : deq_flags_) {
      if (x > 0) ++out.corrected;
      if (x < 0) ++out.failed;
      out.max_iterations = std::max(out.max_iterations, x);
    }
    return out;
  }

  void set_inflow(const gsk_inflow_state* in, std::size_t n) override {
    GSK_CHECK(n == static_cast<std::size_t>(nb_), "inflow length differs from the boundary-face count");
    std::vector<dev::Inflow<F>> h(n);
    wall_scale_host_.resize(n);
    for (std::size_t b = 0; b < n; ++b) {
      h[b].m = dev::maxw_from_state<F>(in[b].rho, in[b].ux, in[b].uy, in[b].uz, in[b].T);
      h[b].scale = static_cast<F>(in[b].scale);
      wall_scale_host_[b] = in[b].scale;
    }
    if (n) {
      HIPCHK(hipMemcpyAsync(inflow_.p, h.data(), n * sizeof(dev::Inflow<F>), hipMemcpyHostToDevice, stream_));
      HIPCHK(hipMemcpyAsync(wall_scale_.p, wall_scale_host_.data(), n * sizeof(double), hipMemcpyHostToDevice, stream_));
    }
    HIPCHK(hipStreamSynchronize(stream_));
  }

  // The wall refresh between inner sweeps stays on the device: the listed faces' sums, then
  // k_wall_refresh updates the FP64 scales (kept beside inflow_, loaded by set_inflow) and
  // inflow_'s F scales in place. Nothing returns to the host until wall_inflow_scales.
  bool refresh_wall_inflow(const int* faces, const int* maxwell, const double* unit0, std::size_t n,
                           double relaxation) override {
    if (n == 0) return true;
    GSK_CHECK(n <= static_cast<std::size_t>(nb_), "more wall faces than boundary faces");
    if (wall_faces_host_.size() != n || !std::equal(faces, faces + n, wall_faces_host_.begin()) ||
        !std::equal(maxwell, maxwell + n, wall_kind_host_.begin()) ||
        !std::equal(unit0, unit0 + n, wall_unit0_host_.begin())) {
      for (std::size_t i = 0; i < n; ++i) GSK_CHECK(faces[i] >= 0 && faces[i] < nb_, "face index out of range");
      wall_faces_host_.assign(faces, faces + n);
      wall_kind_host_.assign(maxwell, maxwell + n);
      wall_unit0_host_.assign(unit0, unit0 + n);
      wall_faces_.upload(wall_faces_host_.data(), n, stream_);
      wall_kind_.upload(wall_kind_host_.data(), n, stream_);
      wall_unit0_.upload(wall_unit0_host_.data(), n, stream_);
    }
    begin(kBoundary);
    launch_subset_sums(wall_faces_.p, n);
    dev::k_wall_refresh<F><<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream_>>>(
        wall_faces_.p, wall_kind_.p, wall_unit0_.p, static_cast<int>(n), relaxation, bsums_.p, wall_scale_.p, inflow_.p,
        wall_negative_.p);
    HIPCHK(hipGetLastError());
    end();
    return true;
  }

  void wall_inflow_scales(const int* faces, std::size_t n, double* out) override {
    download(wall_scale_host_.data(), wall_scale_.p, wall_scale_host_.size());
    for (std::size_t i = 0; i < n; ++i) {
      GSK_CHECK(faces[i] >= 0 && faces[i] < nb_, "face index out of range");
      out[i] = wall_scale_host_[static_cast<std::size_t>(faces[i])];
    }
    int negative = 0;
    download(&negative, wall_negative_.p, 1);
    if (negative) {
      HIPCHK(hipMemsetAsync(wall_negative_.p, 0, sizeof(int), stream_));
      HIPCHK(hipStreamSynchronize(stream_));
      throw Error("Maxwell wall diffuse density became negative; refine the velocity grid");
    }
  }

  bool correct_inflow(gsk_inflow_state* states, std::size_t n) override {
    if (n == 0) return true;
    static_assert(sizeof(gsk_inflow_state) == 6 * sizeof(double), "inflow state is six doubles");
    begin(kOther);
    inflow_stage_.upload(reinterpret_cast<const double*>(states), n * 6, stream_);
    dev::k_inflow_correct<<<static_cast<unsigned>(std::min<std::size_t>(n, 65535)), dev::kBlock, 0, stream_>>>(
        nv_, xd_.p, yd_.p, zd_.p, weight_, wqd_.p, view().deq, inflow_tol_, inflow_stage_.p, nullptr,
        static_cast<int>(n));
    HIPCHK(hipGetLastError());
    end();
    download(reinterpret_cast<double*>(states), inflow_stage_.p, n * 6);
    return true;
  }

  void boundary_sums(double* out, std::size_t n) override {
    GSK_CHECK(n == static_cast<std::size_t>(nb_) * dev::kBoundarySums, "boundary sums length differs from 20 x faces");
    begin(kBoundary);
    if (n_mw_ > 0) {
      const dim3 grid((nv_ + 255) / 256, nb_);
      dev::k_maxwell_reflection<F, false><<<grid, 256, 0, stream_>>>(view(), specular_first_.p);
      launch_reflection(grid);
    }
    if (nb_ > 0) launch_boundary_sums(nullptr, nb_);
    HIPCHK(hipGetLastError());
    end();
    download(out, bsums_.p, n);
  }

  void boundary_sums_subset(const int* faces, std::size_t n_faces, double* out, std::size_t n) override {
    GSK_CHECK(n == static_cast<std::size_t>(nb_) * dev::kBoundarySums, "boundary sums length differs from 20 x faces");
    if (n_faces == 0) return;
    for (std::size_t i = 0; i < n_faces; ++i) GSK_CHECK(faces[i] >= 0 && faces[i] < nb_, "face index out of range");
    if (subset_host_.size() != n_faces || !std::equal(faces, faces + n_faces, subset_host_.begin())) {
      subset_host_.assign(faces, faces + n_faces);
      subset_.upload(subset_host_.data(), n_faces, stream_);
    }
    begin(kBoundary);
    launch_subset_sums(subset_.p, n_faces);
    HIPCHK(hipGetLastError());
    end();
    download(out, bsums_.p, n);
  }

  // Maxwell-wall reflection caches, then the boundary sums of the listed faces (device list faces_dev)
  void launch_subset_sums(const int* faces_dev, std::size_t n_faces) {
    if (n_mw_ > 0) {
      const dim3 grid((nv_ + 255) / 256, nb_);
      dev::k_maxwell_reflection<F, false><<<grid, 256, 0, stream_>>>(view(), specular_first_.p);
      launch_reflection(grid);
    }
    launch_boundary_sums(faces_dev, static_cast<int>(n_faces));
  }

  // partial sums over velocity chunks, then their fixed-order sum into bsums_
  void launch_boundary_sums(const int* faces, int n_faces) {
    const dim3 grid(static_cast<unsigned>(std::min(n_faces, 65535)), static_cast<unsigned>(bsum_chunks_), 2);
    if (lim_mode_ == 1)
      dev::k_boundary_sums<F, 1><<<grid, dev::kBlock, 0, stream_>>>(view(), faces, n_faces, bsum_part_.p);
    else if (lim_mode_ == 2)
      dev::k_boundary_sums<F, 2><<<grid, dev::kBlock, 0, stream_>>>(view(), faces, n_faces, bsum_part_.p);
    else
      dev::k_boundary_sums<F><<<grid, dev::kBlock, 0, stream_>>>(view(), faces, n_faces, bsum_part_.p);
    const int nj = n_faces * dev::kBoundarySums;
    dev::k_boundary_reduce<<<(nj + 255) / 256, 256, 0, stream_>>>(faces, n_faces, bsum_chunks_, bsum_part_.p, bsums_.p);
  }

  // gradient / residual work items (workgroups) of a chunk of width k: 8 XCD lanes x ceil(tiles / 8)
  // x cells; the kernels index them with 32-bit unsigned arithmetic
  static constexpr long long kMaxWorkItems = 0xFFFFFFFFLL;
  long long work_items(long long k) const {
    const long long t = (k + dev::kBlock - 1) / dev::kBlock;
    return 8LL * ((t + 7) / 8) * nc_;
  }

  // Workgroups of one launch over `work` items: at most grid_cap_ (a multiple of 8, so a kernel's
  // L mod 8 stays its XCD), keeping blocks x threads far below the 2^32 a launch allows; the
  // gradient / residual cover larger work with several launches, the CSR-path kernels loop.
  // the reconstructed specular cache of the Maxwell walls, on the current limiter-branch mode
  void launch_reflection(dim3 grid) {
    if (lim_mode_ == 1)
      dev::k_maxwell_reflection<F, true, 1><<<grid, 256, 0, stream_>>>(view(), specular_.p);
    else if (lim_mode_ == 2)
      dev::k_maxwell_reflection<F, true, 2><<<grid, 256, 0, stream_>>>(view(), specular_.p);
    else
      dev::k_maxwell_reflection<F, true><<<grid, 256, 0, stream_>>>(view(), specular_.p);
  }

  unsigned capped_grid(long long work) const {
    const long long g = std::min<long long>(std::max<long long>(work, 8), grid_cap_);
    return static_cast<unsigned>((g + 7) / 8 * 8);
  }

  // gradient and residual R of one velocity chunk into rres_ (and per-cell sums of R^2 into residual_)
  void chunk_residual(const dev::View<F>& v, const dev::Chunk<F>& ch) {
    const long long work = work_items(chunk_);
    const unsigned grid_irr = capped_grid(static_cast<long long>(n_irregular_) * tiles_);
    begin(kGradient);
    for (long long L0 = 0; L0 < work; L0 += grid_cap_) {
      const unsigned grid = capped_grid(work - L0), l0 = static_cast<unsigned>(L0);
      if (v.lim_mode == 1)
        dev::k_gradient<F, 1><<<grid, dev::kBlock, 0, stream_>>>(v, ch, tiles_, l0);
      else if (v.lim_mode == 2)
        dev::k_gradient<F, 2><<<grid, dev::kBlock, 0, stream_>>>(v, ch, tiles_, l0);
      else
        dev::k_gradient<F><<<grid, dev::kBlock, 0, stream_>>>(v, ch, tiles_, l0);
    }
    if (n_irregular_ > 0)
    {
      if (v.lim_mode == 1)
        dev::k_gradient_csr<F, 1><<<grid_irr, dev::kBlock, 0, stream_>>>(v, ch, irregular_.p, n_irregular_, tiles_);
      else if (v.lim_mode == 2)
        dev::k_gradient_csr<F, 2><<<grid_irr, dev::kBlock, 0, stream_>>>(v, ch, irregular_.p, n_irregular_, tiles_);
      else
        dev::k_gradient_csr<F><<<grid_irr, dev::kBlock, 0, stream_>>>(v, ch, irregular_.p, n_irregular_, tiles_);
    }
    HIPCHK(hipGetLastError());
    end();
    begin(kResidual);
    for (long long L0 = 0; L0 < work; L0 += grid_cap_)
      dev::k_residual<F><<<capped_grid(work - L0), dev::kBlock, 0, stream_>>>(v, ch, tiles_, static_cast<unsigned>(L0),
                                                                               rres_.p, partial_.p);
    if (n_irregular_ > 0)
      dev::k_residual_csr<F><<<grid_irr, dev::kBlock, 0, stream_>>>(v, ch, irregular_.p, n_irregular_, tiles_, rres_.p,
                                                                   partial_.p);
    dev::k_residual_reduce<<<(nc_ + 255) / 256, 256, 0, stream_>>>(partial_.p, tiles_, nc_, residual_.p);
    HIPCHK(hipGetLastError());
    end();
  }

  // block_passes x (forward, backward) LU-SGS passes of one chunk on the R in rres_; returns the
  // buffer holding the final increment, added to f when update_f
  F* chunk_passes(const dev::View<F>& v, const dev::Chunk<F>& ch, bool update_f) {
    // velocity tiles of every sweep direction inside this chunk
    dev::VelRange vr{};
    int max_tiles = 0;
    for (int g = 0; g < n_dirs_eff_; ++g) {
      const int* lo = vel_list_host_.data() + dir_off_[static_cast<std::size_t>(g)];
      const int* hi = vel_list_host_.data() + dir_off_[static_cast<std::size_t>(g) + 1];
      vr.beg[g] = static_cast<int>(std::lower_bound(lo, hi, ch.k0) - vel_list_host_.data());
      vr.end[g] = static_cast<int>(std::lower_bound(lo, hi, ch.k0 + ch.kc) - vel_list_host_.data());
      max_tiles = std::max(max_tiles, (vr.end[g] - vr.beg[g] + 63) / 64);
    }
    const int tiles = std::max(max_tiles, 1);
    const int n_pairs = n_dirs_eff_ * n_blocks_;
    const long long sgrid_ll = 8LL * ((n_pairs + 7) / 8) * tiles;
    GSK_CHECK(sgrid_ll * 64 * waves_ < (1LL << 32), "sweep grid too large");
    const unsigned sgrid = static_cast<unsigned>(sgrid_ll);
    const dev::SweepArgs fw{ch.k0, ch.stride, n_blocks_, n_pairs, tiles, block_off_.p, vel_list_.p, fw_stage_.p,
                            fw_group_.p, vr};
    const dev::SweepArgs bw{ch.k0, ch.stride, n_blocks_, n_pairs, tiles, block_off_.p, vel_list_.p, bw_stage_.p,
                            bw_group_.p, vr};
    F* cur = bufA_.p;
    F* prev = nullptr;
    for (int pass = 0; pass < block_passes_; ++pass) {
      const int has_prev = prev != nullptr ? 1 : 0, last = (update_f && pass + 1 == block_passes_) ? 1 : 0;
      begin(kForward);
      if (has_prev)
        dev::k_sweep_forward<F, true><<<sgrid, 64 * waves_, 0, stream_>>>(v, fw, fw_recs_.p, rres_.p, cur, prev);
      else
        dev::k_sweep_forward<F, false><<<sgrid, 64 * waves_, 0, stream_>>>(v, fw, fw_recs_.p, rres_.p, cur, prev);
      HIPCHK(hipGetLastError());
      end();
      begin(kBackward);
      if (has_prev)
        dev::k_sweep_backward<F, true><<<sgrid, 64 * waves_, 0, stream_>>>(v, bw, bw_recs_.p, cur, last);
      else
        dev::k_sweep_backward<F, false><<<sgrid, 64 * waves_, 0, stream_>>>(v, bw, bw_recs_.p, cur, last);
      HIPCHK(hipGetLastError());
      end();
      prev = cur;
      cur = cur == bufA_.p ? bufB_.p : bufA_.p;
    }
    return prev;
  }

  void sweep(double* residual_cell, std::size_t n) override {
    GSK_CHECK(n == static_cast<std::size_t>(nc_), "residual output length differs from the cell count");
    launch_sweep();
    download(residual_cell, residual_.p, n);
  }

  // no download and no stream synchronisation: the next inner sweep queues behind this one
  void sweep_unreported(double* unused, std::size_t n) override {
    (void)unused;
    GSK_CHECK(n == static_cast<std::size_t>(nc_), "residual output length differs from the cell count");
    launch_sweep();
  }

  void launch_sweep() {
    HIPCHK(hipMemsetAsync(residual_.p, 0, residual_.bytes(), stream_));
    const dev::View<F> v = view();
    for (int k0 = 0; k0 < nv_; k0 += chunk_) {
      const dev::Chunk<F> ch{k0, std::min(chunk_, nv_ - k0), chunk_, gx_.p, gy_.p, gz_.p};
      chunk_residual(v, ch);
      chunk_passes(v, ch, true);
    }
  }

  // ---- device-resident operators for sensitivity solves (arrays [cell][storage position], type F) ----
  int element_size() const override { return static_cast<int>(sizeof(F)); }

  // copy columns [k0, k0 + kc) between a [nc][nv] array and a [nc][chunk_] workspace
  void copy_chunk(F* dst, std::size_t dpitch, const F* src, std::size_t spitch, int kc) {
    HIPCHK(hipMemcpy2DAsync(dst, dpitch * sizeof(F), src, spitch * sizeof(F), static_cast<std::size_t>(kc) * sizeof(F),
                            static_cast<std::size_t>(nc_), hipMemcpyDeviceToDevice, stream_));
  }

  void residual_field_device(void* out) override {
    F* o = static_cast<F*>(out);
    HIPCHK(hipMemsetAsync(residual_.p, 0, residual_.bytes(), stream_));
    const dev::View<F> v = view();
    for (int k0 = 0; k0 < nv_; k0 += chunk_) {
      const dev::Chunk<F> ch{k0, std::min(chunk_, nv_ - k0), chunk_, gx_.p, gy_.p, gz_.p};
      chunk_residual(v, ch);
      copy_chunk(o + k0, static_cast<std::size_t>(nv_), rres_.p, static_cast<std::size_t>(chunk_), ch.kc);
    }
    HIPCHK(hipStreamSynchronize(stream_));
    if (lim_mode_ == 1) lim_mode_ = 2;  // recorded at this state: later evaluations follow the record
  }

  void set_limiter_branches(int mode) override {
    GSK_CHECK(mode >= 0 && mode <= 2, "limiter branch mode must be 0 (off), 1 (record) or 2 (follow)");
    if (mode != 0) {
      GSK_CHECK(max_faces_ <= dev::kLimMaxFaces, "limiter branches hold at most 15 faces per cell");
      if (branch_.n == 0) {
        branch_.alloc(static_cast<std::size_t>(nc_) * static_cast<std::size_t>(nv_));
        HIPCHK(hipMemsetAsync(branch_.p, 0, branch_.bytes(), stream_));
      }
      GSK_CHECK(mode == 1 || lim_mode_ != 0, "no limiter branches have been recorded");
    }
    lim_mode_ = mode;
  }

  // ---- the prepare of a CIS iteration on the device (kinetic_control.hpp) ----
  bool control_setup(const ControlTables& t) override {
    const auto nb = static_cast<std::size_t>(nb_);
    GSK_CHECK(t.face_owner.size() == nb && t.face_patch.size() == nb && t.face_area.size() == nb &&
                  t.face_normal.size() == 3 * nb && t.volume.size() == static_cast<std::size_t>(nc_),
              "control tables do not match the setup");
    const std::size_t npf = t.press_face.size(), nsl = t.slot_cell.size();
    GSK_CHECK(t.press_owner.size() == npf && t.press_slot0.size() == npf && t.press_nslots.size() == npf &&
                  t.press_eps2.size() == npf && t.press_to_face.size() == 3 * npf && t.slot_bface.size() == nsl &&
                  t.slot_periodic.size() == nsl && t.slot_lsq.size() == 3 * nsl && t.slot_offset.size() == 3 * nsl,
              "pressure-face tables are inconsistent");
    std::vector<dev::CtrlFace> faces(nb);
    for (std::size_t b = 0; b < nb; ++b) {
      GSK_CHECK(t.face_owner[b] >= 0 && t.face_owner[b] < nc_ && t.face_patch[b] >= 0, "boundary face table out of range");
      ctl_max_patch_ = std::max(ctl_max_patch_, t.face_patch[b]);
      faces[b] = dev::CtrlFace{t.face_owner[b], t.face_patch[b], t.face_area[b], t.face_normal[3 * b],
                               t.face_normal[3 * b + 1], t.face_normal[3 * b + 2]};
    }
    std::vector<dev::CtrlPressureFace> pf(npf);
    for (std::size_t i = 0; i < npf; ++i) {
      GSK_CHECK(t.press_face[i] >= 0 && t.press_face[i] < nb_ && t.press_owner[i] >= 0 && t.press_owner[i] < nc_ &&
                    t.press_slot0[i] >= 0 && t.press_nslots[i] >= 0 &&
                    static_cast<std::size_t>(t.press_slot0[i]) + t.press_nslots[i] <= nsl,
                "pressure-face table out of range");
      pf[i] = dev::CtrlPressureFace{t.press_face[i], t.press_owner[i], t.press_slot0[i], t.press_nslots[i], t.press_eps2[i],
                                    t.press_to_face[3 * i], t.press_to_face[3 * i + 1], t.press_to_face[3 * i + 2]};
    }
    std::vector<dev::CtrlSlot> slots(nsl);
    for (std::size_t k = 0; k < nsl; ++k) {
      GSK_CHECK(t.slot_cell[k] >= -1 && t.slot_cell[k] < nc_ &&
                    (t.slot_cell[k] >= 0 || (t.slot_bface[k] >= 0 && t.slot_bface[k] < nb_)),
                "stencil slot out of range");
      slots[k] = dev::CtrlSlot{t.slot_cell[k], t.slot_bface[k], t.slot_periodic[k], 0,
                               t.slot_lsq[3 * k], t.slot_lsq[3 * k + 1], t.slot_lsq[3 * k + 2],
                               t.slot_offset[3 * k], t.slot_offset[3 * k + 1], t.slot_offset[3 * k + 2]};
    }
    ctl_faces_.upload(faces.data(), nb, stream_);
    ctl_pf_.upload(pf.data(), npf, stream_);
    ctl_pf_index_.upload(t.press_face.data(), npf, stream_);
    ctl_slots_.upload(slots.data(), nsl, stream_);
    ctl_volume_.upload(t.volume.data(), t.volume.size(), stream_);
    if (!t.wall_unit.empty()) {
      GSK_CHECK(t.wall_unit.size() == 10 * nb, "wall unit sums do not match the boundary faces");
      ctl_wall_unit_.upload(t.wall_unit.data(), t.wall_unit.size(), stream_);
    }
    n_pf_ = static_cast<int>(npf);
    ctl_blocks_ = std::max(1, std::min(1024, (nc_ + dev::kCtlBlock - 1) / dev::kCtlBlock));
    ctl_part_.alloc(static_cast<std::size_t>(ctl_blocks_) * std::max(dev::kCtlMassParts, dev::kCtlStatParts));
    prev_moments_.alloc(moments_.n);
    ctl_gamma_ = t.gamma;
    ctl_u_floor_ = t.u_floor;
    ctl_relax_ = t.relaxation;
    ctl_limiter_ = t.limiter;
    ctl_cell_mode_ = t.cell_reconstruction;
    ctl_closed_ = t.closed_domain;
    HIPCHK(hipStreamSynchronize(stream_));  // the host tables above are released on return
    ctl_ready_ = true;
    return true;
  }

  void control_prepare(ControlStep& st) override {
    GSK_CHECK(ctl_ready_, "the device prepare has no tables");
    GSK_CHECK(st.n_faces == static_cast<std::size_t>(nb_) && (nb_ == 0 || (st.inflow && st.face_state && st.flux_out &&
                                                                          st.flux_in)),
              "boundary state does not match the boundary faces");
    GSK_CHECK(st.n_patches > 0 && st.patches && st.patch_flow, "the device prepare needs the patches");
    const auto nb = static_cast<std::size_t>(nb_), np = st.n_patches;
    std::vector<dev::CtrlPatch> pp(np);
    for (std::size_t i = 0; i < np; ++i) {
      const ControlPatch& c = st.patches[i];
      GSK_CHECK(c.first >= 0 && c.n >= 0 && c.first + c.n <= nb_, "patch face range out of range");
      pp[i] = dev::CtrlPatch{c.kind, c.inlet, c.zero_tangential, c.first, c.n, c.p, c.T, c.accommodation};
    }
    GSK_CHECK(ctl_max_patch_ < static_cast<int>(np), "boundary face patch out of range");
    // one device record: control slots | states (6 per face) | face states (5) | F+ (5) | F- (5) |
    // patch flows | per-face flows
    const std::size_t o_state = dev::kCtlSize, o_fs = o_state + 6 * nb, o_fo = o_fs + 5 * nb, o_fi = o_fo + 5 * nb,
                      o_pf = o_fi + 5 * nb, o_ff = o_pf + np, total = o_ff + nb;
    if (ctl_out_.n != total) ctl_out_.alloc(total);
    ctl_host_.assign(o_pf, 0.0);
    ctl_host_[dev::kCtlTarget] = st.mass_target;
    ctl_host_[dev::kCtlHasTarget] = st.has_mass_target ? 1.0 : 0.0;
    if (nb > 0) {
      std::memcpy(&ctl_host_[o_state], st.inflow, nb * sizeof(gsk_inflow_state));
      std::memcpy(&ctl_host_[o_fs], st.face_state, 5 * nb * sizeof(double));
    }
    ctl_patches_.upload(pp.data(), np, stream_);
    HIPCHK(hipMemcpyAsync(ctl_out_.p, ctl_host_.data(), o_fo * sizeof(double), hipMemcpyHostToDevice, stream_));
    double* ctrl = ctl_out_.p;
    double* state = ctrl + o_state;
    // moments, positivity, mass, closed-domain rescaling, epsilon and the volume means
    launch_moments();
    begin(kOther);
    dev::k_ctrl_mass<<<ctl_blocks_, dev::kCtlBlock, 0, stream_>>>(moments_.p, ctl_volume_.p, nc_, ctl_part_.p);
    dev::k_ctrl_mass_reduce<<<1, 64, 0, stream_>>>(ctl_part_.p, ctl_blocks_, moments_.p, ctl_closed_ ? 1 : 0, ctrl);
    if (ctl_closed_) {
      dev::k_ctrl_scale_f<F><<<4096, 256, 0, stream_>>>(f_.p, static_cast<long long>(f_.n), ctrl);
      dev::k_ctrl_scale_moments<<<(nc_ + 255) / 256, 256, 0, stream_>>>(moments_.p, nc_, ctrl);
    }
    dev::k_ctrl_stats<<<ctl_blocks_, dev::kCtlBlock, 0, stream_>>>(moments_.p, st.has_prev ? prev_moments_.p : nullptr,
                                                                   ctl_volume_.p, nc_, ctl_part_.p);
    dev::k_ctrl_stats_reduce<<<1, 64, 0, stream_>>>(ctl_part_.p, ctl_blocks_, st.has_prev ? 1 : 0, ctl_u_floor_, ctrl);
    HIPCHK(hipGetLastError());
    end();
    download(ctl_host_.data(), ctrl, static_cast<std::size_t>(dev::kCtlSize));
    st.bad = static_cast<long long>(ctl_host_[dev::kCtlBad]);
    st.mass = ctl_host_[dev::kCtlMass];
    st.scale = ctl_host_[dev::kCtlScale];
    st.has_mass_target = ctl_host_[dev::kCtlHasTarget] != 0.0;
    st.mass_target = ctl_host_[dev::kCtlTarget];
    for (int k = 0; k < 3; ++k) st.eps[k] = ctl_host_[dev::kCtlEpsRho + k];
    for (int k = 0; k < 3; ++k) st.q_mean[k] = ctl_host_[dev::kCtlQ + k];
    for (int k = 0; k < 6; ++k) st.sigma_mean[k] = ctl_host_[dev::kCtlSigma + k];
    if (st.bad > 0) {  // the host throws before the cell parameters, as the host prepare does
      st.first_bad = static_cast<int>(ctl_host_[dev::kCtlFirstBad]);
      st.bad_rho = ctl_host_[dev::kCtlBadRho];
      st.bad_T = ctl_host_[dev::kCtlBadT];
      return;
    }
    update_cell_params();
    // boundary rounds
    const unsigned fb = static_cast<unsigned>((nb + 255) / 256);
    begin(kBoundary);
    if (nb > 0 && st.seed_walls)
      dev::k_ctrl_seed_walls<<<fb, 256, 0, stream_>>>(ctl_faces_.p, ctl_patches_.p, nb_, moments_.p, state);
    for (long long round = 0; round < st.rounds && nb > 0; ++round) {
      if (n_pf_ > 0) {
        dev::k_ctrl_pressure<<<static_cast<unsigned>((n_pf_ + 63) / 64), 64, 0, stream_>>>(
            ctl_pf_.p, n_pf_, ctl_slots_.p, ctl_faces_.p, ctl_patches_.p, moments_.p, ctrl + o_fs,
            (round > 0 || st.has_face_state) ? 1 : 0, ctl_limiter_ ? 1 : 0, ctl_cell_mode_ ? 1 : 0, ctl_gamma_, state);
        if (discrete_eq_)
          dev::k_inflow_correct<<<static_cast<unsigned>(std::min(n_pf_, 65535)), dev::kBlock, 0, stream_>>>(
              nv_, xd_.p, yd_.p, zd_.p, weight_, wqd_.p, view().deq, inflow_tol_, state, ctl_pf_index_.p, n_pf_);
      }
      dev::k_ctrl_inflow_table<F><<<fb, 256, 0, stream_>>>(state, nb_, inflow_.p, wall_scale_.p);
      launch_subset_sums(nullptr, nb);
      dev::k_ctrl_absorb<<<fb, 256, 0, stream_>>>(bsums_.p, ctl_faces_.p, ctl_patches_.p, ctl_wall_unit_.p, nb_,
                                                  ctl_relax_, state, ctrl + o_fo, ctrl + o_fi, ctrl + o_fs,
                                                  ctrl + o_ff, ctrl);
      dev::k_ctrl_patch_flow<<<static_cast<unsigned>(np), dev::kCtlBlock, 0, stream_>>>(ctrl + o_ff, ctl_patches_.p,
                                                                                        static_cast<int>(np), ctrl + o_pf);
      dev::k_ctrl_inflow_table<F><<<fb, 256, 0, stream_>>>(state, nb_, inflow_.p, wall_scale_.p);
    }
    HIPCHK(hipGetLastError());
    end();
    ctl_host_.resize(o_ff);
    download(ctl_host_.data(), ctrl, o_ff);
    st.maxwell_negative = ctl_host_[dev::kCtlMaxwellNegative] != 0.0;
    st.maxwell_no_incoming = ctl_host_[dev::kCtlMaxwellNoIncoming] != 0.0;
    if (nb > 0) {
      std::memcpy(st.inflow, &ctl_host_[o_state], nb * sizeof(gsk_inflow_state));
      std::memcpy(st.face_state, &ctl_host_[o_fs], 5 * nb * sizeof(double));
      std::memcpy(st.flux_out, &ctl_host_[o_fo], 5 * nb * sizeof(double));
      std::memcpy(st.flux_in, &ctl_host_[o_fi], 5 * nb * sizeof(double));
    }
    std::memcpy(st.patch_flow, &ctl_host_[o_pf], np * sizeof(double));
    wall_scale_host_.resize(nb);
    for (std::size_t b = 0; b < nb; ++b) wall_scale_host_[b] = st.inflow[b].scale;
  }

  void download_moments(double* out, std::size_t n) override {
    GSK_CHECK(n == moments_.n, "moments output length differs from 15 x cells");
    download(out, moments_.p, n);
  }
  void keep_previous_moments() override {
    if (prev_moments_.n != moments_.n) prev_moments_.alloc(moments_.n);
    HIPCHK(hipMemcpyAsync(prev_moments_.p, moments_.p, moments_.bytes(), hipMemcpyDeviceToDevice, stream_));
  }
  void set_previous_moments(const double* m, std::size_t n) override {
    GSK_CHECK(n == moments_.n, "previous moments length differs from 15 x cells");
    prev_moments_.upload(m, n, stream_);
    HIPCHK(hipStreamSynchronize(stream_));
  }
  void previous_moments(double* out, std::size_t n) override {
    GSK_CHECK(n == moments_.n && prev_moments_.n == n, "no previous moments on the device");
    download(out, prev_moments_.p, n);
  }

  // ---- adjoint operators (kinetic_adjoint.hpp), FP64 ----
  unsigned adj_grid(long long items) const { return static_cast<unsigned>(std::max(1LL, std::min(items, 1LL << 18))); }

  void adjoint_prepare() override {
    if constexpr (!std::is_same_v<F, double>) {
      throw Error("adjoint operators need the FP64 form ([kinetic].precision = \"fp64\")");
    } else {
      GSK_CHECK(lim_mode_ == 2, "adjoint operators need the limiter branches recorded at the linearisation state");
      const std::size_t nf = static_cast<std::size_t>(nc_) * static_cast<std::size_t>(nv_);
      if (adj_x_[0].n == 0) {
        std::size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        GSK_CHECK(8 * nf * sizeof(double) < free_b,
                  strf("the adjoint workspace needs %.2f GB but only %.2f GB of device memory is free",
                       8.0 * nf * sizeof(double) / 1e9, free_b / 1e9));
        for (auto* b : {&adj_x_[0], &adj_x_[1], &adj_x_[2], &adj_o_[0], &adj_o_[1], &adj_o_[2], &adj_o_[3], &adj_o_[4]})
          b->alloc(nf);
        adj_lp_.alloc(static_cast<std::size_t>(nc_) * dev::kAdjPStride);
        adj_jacp_.alloc(static_cast<std::size_t>(nc_) * dev::kAdjParams * dev::kMomentSums);
        adj_jacm_.alloc(static_cast<std::size_t>(nc_) * dev::kMoments * dev::kMomentSums);
        adj_lm_.alloc(static_cast<std::size_t>(nc_) * dev::kMoments);
        adj_lb_.alloc(static_cast<std::size_t>(nb_) * 6 + 1);
        adj_ls_.alloc(static_cast<std::size_t>(nb_) * dev::kBoundarySums + 1);
        adj_shift_.alloc(static_cast<std::size_t>(nc_) * 3);
        if (n_mw_ > 0) {
          adj_lspec_.alloc(static_cast<std::size_t>(n_mw_) * nv_);
          adj_lspec1_.alloc(static_cast<std::size_t>(n_mw_) * nv_);
        }
        adj_mate_.upload(adj_mate_host_.data(), adj_mate_host_.size(), stream_);
        std::vector<int> slot(adj_bslot_host_), owners;
        slot.push_back(0);
        adj_bslot_.upload(slot.data(), slot.size(), stream_);
        for (int b = 0; b < nb_; ++b) owners.push_back(host_bface_owner_[static_cast<std::size_t>(b)]);
        std::sort(owners.begin(), owners.end());
        owners.erase(std::unique(owners.begin(), owners.end()), owners.end());
        n_adj_cells_ = static_cast<int>(owners.size());
        owners.push_back(0);
        adj_cells_.upload(owners.data(), owners.size(), stream_);
      }
      HIPCHK(hipMemcpyAsync(adj_shift_.p, shift_.p, adj_shift_.bytes(), hipMemcpyDeviceToDevice, stream_));
      dev::k_adj_param_jacobian<<<grid_cells(), dev::kBlock, 0, stream_>>>(view(), adj_shift_.p, volume_.p, vdt_.p, mu0_,
                                                                          omega_, Pr_, 1e-5, adj_jacp_.p, adj_jacm_.p);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(stream_));
    }
  }

  void adjoint_stage1(const void* lam, void* out, double* inflow_adj, std::size_t n) override {
    if constexpr (!std::is_same_v<F, double>) {
      throw Error("adjoint operators need the FP64 form");
    } else {
      GSK_CHECK(adj_x_[0].n != 0, "adjoint operators are not prepared");
      GSK_CHECK(n == static_cast<std::size_t>(nb_) * 6, "inflow adjoint length differs from 6 x boundary faces");
      const dev::View<double> v = view();
      const double* l = static_cast<const double*>(lam);
      double* o = static_cast<double*>(out);
      const int tiles = (nv_ + dev::kAdjBlock - 1) / dev::kAdjBlock;
      const unsigned grid = adj_grid(static_cast<long long>(nc_) * tiles);
      if (n_mw_ > 0) {
        HIPCHK(hipMemsetAsync(adj_lspec_.p, 0, adj_lspec_.bytes(), stream_));
        HIPCHK(hipMemsetAsync(adj_lspec1_.p, 0, adj_lspec1_.bytes(), stream_));
      }
      dev::k_adj_transport<<<grid, dev::kAdjBlock, 0, stream_>>>(v, adj_mate_.p, l, o, adj_x_[0].p, adj_x_[1].p,
                                                                  adj_x_[2].p, adj_lspec_.p, tiles);
      dev::k_adj_collision<<<static_cast<unsigned>(std::min(nc_, 65535 * 4)), dev::kAdjBlock, 0, stream_>>>(v, l, adj_lp_.p);
      if (n_mw_ > 0)
        dev::k_adj_reflection<true><<<dim3((nv_ + 255) / 256, nb_), dev::kAdjBlock, 0, stream_>>>(
            v, adj_lspec_.p, o, adj_x_[0].p, adj_x_[1].p, adj_x_[2].p, adj_lp_.p);
      dev::k_adj_gradient<false><<<grid, dev::kAdjBlock, 0, stream_>>>(v, nullptr, 0, adj_x_[0].p, adj_x_[1].p,
                                                                       adj_x_[2].p, adj_o_[0].p, adj_o_[1].p, adj_o_[2].p,
                                                                       adj_o_[3].p, adj_o_[4].p, o, tiles);
      if (nb_ > 0)
        dev::k_adj_inflow<<<static_cast<unsigned>(std::min(nb_, 65535)), dev::kAdjBlock, 0, stream_>>>(
            v, adj_bslot_.p, l, adj_o_[0].p, adj_o_[1].p, adj_o_[2].p, adj_o_[3].p, adj_o_[4].p, adj_lb_.p);
      HIPCHK(hipGetLastError());
      if (nb_ > 0) download(inflow_adj, adj_lb_.p, n);
      HIPCHK(hipStreamSynchronize(stream_));
    }
  }

  void adjoint_stage2(const double* sums_adj, std::size_t n_sums, const double* moments_adj, std::size_t n_m,
                      void* out) override {
    if constexpr (!std::is_same_v<F, double>) {
      throw Error("adjoint operators need the FP64 form");
    } else {
      GSK_CHECK(adj_x_[0].n != 0, "adjoint operators are not prepared");
      GSK_CHECK(n_sums == static_cast<std::size_t>(nb_) * dev::kBoundarySums,
                "boundary-sum adjoint length differs from 20 x boundary faces");
      GSK_CHECK(moments_adj == nullptr || n_m == static_cast<std::size_t>(nc_) * dev::kMoments,
                "moment adjoint length differs from 15 x cells");
      const dev::View<double> v = view();
      double* o = static_cast<double*>(out);
      const int tiles = (nv_ + dev::kAdjBlock - 1) / dev::kAdjBlock;
      const unsigned grid = adj_grid(static_cast<long long>(nc_) * tiles);
      if (nb_ > 0) HIPCHK(hipMemcpyAsync(adj_ls_.p, sums_adj, n_sums * sizeof(double), hipMemcpyHostToDevice, stream_));
      if (moments_adj)
        HIPCHK(hipMemcpyAsync(adj_lm_.p, moments_adj, n_m * sizeof(double), hipMemcpyHostToDevice, stream_));
      for (auto* b : {&adj_x_[0], &adj_x_[1], &adj_x_[2]}) HIPCHK(hipMemsetAsync(b->p, 0, b->bytes(), stream_));
      if (n_mw_ > 0) {
        HIPCHK(hipMemsetAsync(adj_lspec_.p, 0, adj_lspec_.bytes(), stream_));
        HIPCHK(hipMemsetAsync(adj_lspec1_.p, 0, adj_lspec1_.bytes(), stream_));
      }
      if (nb_ > 0) {
        dev::k_adj_sums<<<static_cast<unsigned>(std::min(nb_, 65535)), dev::kAdjBlock, 0, stream_>>>(
            v, adj_ls_.p, o, adj_x_[0].p, adj_x_[1].p, adj_x_[2].p, adj_lspec_.p);
        if (n_mw_ > 0)
          dev::k_adj_reflection<true><<<dim3((nv_ + 255) / 256, nb_), dev::kAdjBlock, 0, stream_>>>(
              v, adj_lspec_.p, o, adj_x_[0].p, adj_x_[1].p, adj_x_[2].p, adj_lp_.p);
        dev::k_adj_gradient<true><<<adj_grid(static_cast<long long>(n_adj_cells_) * tiles), dev::kAdjBlock, 0, stream_>>>(
            v, adj_cells_.p, n_adj_cells_, adj_x_[0].p, adj_x_[1].p, adj_x_[2].p, adj_o_[0].p, adj_o_[1].p, adj_o_[2].p,
            adj_o_[3].p, adj_o_[4].p, o, tiles);
      }
      dev::k_adj_gather<<<grid, dev::kAdjBlock, 0, stream_>>>(v, adj_mate_.p, adj_o_[0].p, adj_o_[1].p, adj_o_[2].p,
                                                               adj_o_[3].p, adj_o_[4].p, o, adj_lspec1_.p, tiles);
      if (n_mw_ > 0)
        dev::k_adj_reflection<false><<<dim3((nv_ + 255) / 256, nb_), dev::kAdjBlock, 0, stream_>>>(
            v, adj_lspec1_.p, o, adj_x_[0].p, adj_x_[1].p, adj_x_[2].p, adj_lp_.p);
      dev::k_adj_moments<<<static_cast<unsigned>(std::min(nc_, 65535 * 4)), dev::kAdjBlock, 0, stream_>>>(
          v, adj_shift_.p, adj_jacp_.p, adj_jacm_.p, adj_lp_.p, moments_adj ? adj_lm_.p : nullptr, o);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(stream_));
    }
  }

  void precondition_device(const void* r, void* out) override {
    const F* ri = static_cast<const F*>(r);
    F* o = static_cast<F*>(out);
    const dev::View<F> v = view();
    for (int k0 = 0; k0 < nv_; k0 += chunk_) {
      const dev::Chunk<F> ch{k0, std::min(chunk_, nv_ - k0), chunk_, gx_.p, gy_.p, gz_.p};
      copy_chunk(rres_.p, static_cast<std::size_t>(chunk_), ri + k0, static_cast<std::size_t>(nv_), ch.kc);
      F* d = chunk_passes(v, ch, false);
      copy_chunk(o + k0, static_cast<std::size_t>(nv_), d, static_cast<std::size_t>(chunk_), ch.kc);
    }
    HIPCHK(hipStreamSynchronize(stream_));
  }

  void distribution_device(void* out) override {
    HIPCHK(hipMemcpyAsync(out, f_.p, f_.bytes(), hipMemcpyDeviceToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
  }

  void set_distribution_device(const void* in) override {
    HIPCHK(hipMemcpyAsync(f_.p, in, f_.bytes(), hipMemcpyDeviceToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
  }

  void set_mu0(double mu0) override { mu0_ = mu0; }

  void set_newton_tolerances(double cell_tol, double inflow_tol) override {
    deq_.tolerance = cell_tol;
    inflow_tol_ = inflow_tol;
  }

  void set_specular_fraction(const double* frac, std::size_t n) override {
    GSK_CHECK(n == static_cast<std::size_t>(nb_), "specular fractions length differs from the boundary-face count");
    GSK_CHECK(n_mw_ > 0, "the case has no Maxwell wall");
    std::vector<F> h(static_cast<std::size_t>(nb_));
    for (int b = 0; b < nb_; ++b) h[static_cast<std::size_t>(b)] = static_cast<F>(frac[b]);
    mw_frac_.upload(h.data(), h.size(), stream_);
    HIPCHK(hipStreamSynchronize(stream_));
  }

  long long gsis_correction(const double* w_half, const double* w_new, std::size_t n_w, int* flags,
                            std::size_t n_flags) override {
    GSK_CHECK(n_w == static_cast<std::size_t>(nc_) * 5, "GSIS correction states length differs from 5 x cells");
    GSK_CHECK(n_flags == 0 || n_flags == static_cast<std::size_t>(nc_), "flags length differs from the cell count");
    begin(kOther);
    DBuf<double> wh, wn;
    wh.upload(w_half, n_w, stream_);
    wn.upload(w_new, n_w, stream_);
    dev::k_gsis_correction<F><<<grid_cells(), dev::kBlock, 0, stream_>>>(view(), wh.p, wn.p, negative_mode_, shift_.p, flags_.p);
    HIPCHK(hipGetLastError());
    end();
    std::vector<int> h(static_cast<std::size_t>(nc_));
    download(h.data(), flags_.p, h.size());
    long long replaced = 0;
    for (std::size_t c = 0; c < h.size(); ++c) {
      if (h[c] != 0) ++replaced;
      if (n_flags) flags[c] = h[c];
    }
    return replaced;
  }

  // The same correction with w_half / w_new staged in the fixed scratch array instead of buffers
  // allocated per call; the replaced count is formed on the device. Nothing is allocated and no
  // timing event is recorded.
  long long gsis_correction_noalloc(const double* w_half, const double* w_new, std::size_t n_w, int* flags,
                                    std::size_t n_flags) override {
    GSK_CHECK(n_w == static_cast<std::size_t>(nc_) * 5, "GSIS correction states length differs from 5 x cells");
    GSK_CHECK(n_flags == 0 || n_flags == static_cast<std::size_t>(nc_), "flags length differs from the cell count");
    constexpr std::size_t kTail = 256;  // the count, in the last 256 bytes
    GSK_CHECK(2 * n_w * sizeof(double) + kTail <= GSK_NOALLOC_SCRATCH_BYTES,
              strf("%d cells exceed the fixed scratch of the non-allocating GSIS correction (%d bytes: at most %zu "
                   "cells)",
                   nc_, GSK_NOALLOC_SCRATCH_BYTES, (GSK_NOALLOC_SCRATCH_BYTES - kTail) / (2 * 5 * sizeof(double))));
    unsigned char* base = static_cast<unsigned char*>(scratch_);
    double* wh = reinterpret_cast<double*>(base);
    double* wn = wh + n_w;
    long long* count = reinterpret_cast<long long*>(base + GSK_NOALLOC_SCRATCH_BYTES - kTail);
    HIPCHK(hipMemcpyAsync(wh, w_half, n_w * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyAsync(wn, w_new, n_w * sizeof(double), hipMemcpyHostToDevice, stream_));
    dev::k_gsis_correction<F><<<grid_cells(), dev::kBlock, 0, stream_>>>(view(), wh, wn, negative_mode_, shift_.p, flags_.p);
    HIPCHK(hipGetLastError());
    dev::k_count_nonzero<<<1, dev::kBlock, 0, stream_>>>(flags_.p, nc_, count);
    HIPCHK(hipGetLastError());
    if (n_flags) HIPCHK(hipMemcpyAsync(flags, flags_.p, n_flags * sizeof(int), hipMemcpyDeviceToHost, stream_));
    long long replaced = 0;
    HIPCHK(hipMemcpyAsync(&replaced, count, sizeof(long long), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return replaced;
  }

  void internal_face_mass_flux(const int* cells, const int* entries, std::size_t n, double* out) override {
    if (n == 0) return;
    std::vector<int> off(static_cast<std::size_t>(nc_) + 1);
    HIPCHK(hipMemcpyAsync(off.data(), csr_off_.p, off.size() * sizeof(int), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (std::size_t i = 0; i < n; ++i) {
      GSK_CHECK(cells[i] >= 0 && cells[i] < nc_ && entries[i] >= off[static_cast<std::size_t>(cells[i])] &&
                    entries[i] < off[static_cast<std::size_t>(cells[i]) + 1],
                "internal face list refers to an entry outside its cell");
    }
    begin(kOther);
    DBuf<int> dc, de;
    DBuf<double> dout;
    dc.upload(cells, n, stream_);
    de.upload(entries, n, stream_);
    dout.alloc(n);
    dev::k_face_mass_flux<F><<<static_cast<unsigned>(std::min<std::size_t>(n, 65536)), dev::kBlock, 0, stream_>>>(
        view(), dc.p, de.p, static_cast<int>(n), dout.p);
    HIPCHK(hipGetLastError());
    end();
    download(out, dout.p, n);
  }

  gsk_timings take_timings() override {
    HIPCHK(hipStreamSynchronize(stream_));
    for (auto& e : events_) {
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
      timings_[e.phase] += 1e-3 * ms;
      free_events_.push_back(e.a);
      free_events_.push_back(e.b);
    }
    events_.clear();
    gsk_timings t{timings_[kMoments], timings_[kBoundary], timings_[kGradient], timings_[kResidual],
                  timings_[kForward], timings_[kBackward], timings_[kOther]};
    for (int i = 0; i < kNumPhases; ++i) timings_[i] = 0.0;
    return t;
  }

 private:
  // Dependency levels of every (direction, block) for the forward and the backward sweep, and the
  // storage positions of each direction's velocities (see k_sweep_forward).
  void build_levels(const gsk_setup_view& s) {
    n_dirs_eff_ = std::max(1, s.n_sweep_dirs);
    GSK_CHECK(n_dirs_eff_ <= dev::kMaxDirs, "too many sweep directions");
    const int nc = nc_, P = n_blocks_;
    std::vector<int> fw_cells, bw_cells, fw_ptr, bw_ptr, fw_gb, bw_gb;
    std::vector<dev::SweepRec<F>> fw_recs, bw_recs;
    GSK_CHECK(static_cast<long long>(nc_) * 4 < (1LL << 31), "cell count too large for sweep records");
    fw_cells.reserve(static_cast<std::size_t>(n_dirs_eff_) * nc);
    bw_cells.reserve(fw_cells.capacity());
    std::vector<int> lev(static_cast<std::size_t>(nc), 0), ident(static_cast<std::size_t>(nc));
    for (int c = 0; c < nc; ++c) ident[static_cast<std::size_t>(c)] = c;
    long long total_levels = 0;
    for (int g = 0; g < n_dirs_eff_; ++g) {
      const int* order = s.n_sweep_dirs > 0 ? s.sweep_order + static_cast<long long>(g) * nc : ident.data();
      const int* pos = s.n_sweep_dirs > 0 ? s.sweep_pos + static_cast<long long>(g) * nc : ident.data();
      for (int b = 0; b < P; ++b) {
        const int c0 = s.block_offsets[b], c1 = s.block_offsets[b + 1];
        for (int dirn = 0; dirn < 2; ++dirn) {  // 0 forward, 1 backward
          int nlev = 0;
          for (int kk = 0; kk < c1 - c0; ++kk) {
            const int k = dirn == 0 ? c0 + kk : c1 - 1 - kk;
            const int c = order[k];
            int L = 0;
            for (int q = s.csr_offsets[c]; q < s.csr_offsets[c + 1]; ++q) {
              const int j = s.csr[q].other;
              if (j < c0 || j >= c1) continue;
              const bool before = dirn == 0 ? pos[j] < pos[c] : pos[j] > pos[c];
              if (before) L = std::max(L, lev[static_cast<std::size_t>(j)] + 1);
            }
            lev[static_cast<std::size_t>(c)] = L;
            nlev = std::max(nlev, L + 1);
          }
          if (c1 == c0) nlev = 0;
          std::vector<int> count(static_cast<std::size_t>(nlev) + 1, 0);
          for (int k = c0; k < c1; ++k) ++count[static_cast<std::size_t>(lev[static_cast<std::size_t>(order[k])]) + 1];
          for (int l = 0; l < nlev; ++l) count[static_cast<std::size_t>(l) + 1] += count[static_cast<std::size_t>(l)];
          std::vector<int>& cells = dirn == 0 ? fw_cells : bw_cells;
          std::vector<int>& ptr = dirn == 0 ? fw_ptr : bw_ptr;
          std::vector<int>& gb = dirn == 0 ? fw_gb : bw_gb;
          const int base = static_cast<int>(cells.size());
          gb.push_back(static_cast<int>(ptr.size()));
          for (int l = 0; l <= nlev; ++l) ptr.push_back(base + count[static_cast<std::size_t>(l)]);
          cells.resize(cells.size() + static_cast<std::size_t>(c1 - c0));
          std::vector<int> fill(count.begin(), count.end() - 1);
          for (int kk = 0; kk < c1 - c0; ++kk) {
            const int k = dirn == 0 ? c0 + kk : c1 - 1 - kk;
            const int c = order[k];
            cells[static_cast<std::size_t>(base + fill[static_cast<std::size_t>(lev[static_cast<std::size_t>(c)])]++)] = c;
          }
          std::vector<dev::SweepRec<F>>& recs = dirn == 0 ? fw_recs : bw_recs;
          recs.resize(cells.size());
          for (std::size_t i = static_cast<std::size_t>(base); i < cells.size(); ++i) {
            const int c = cells[i];
            dev::SweepRec<F>& r = recs[i];
            std::memset(&r, 0, sizeof(r));
            r.cell = c;
            const int e0 = s.csr_offsets[c], e1 = s.csr_offsets[c + 1];
            if (e1 - e0 > dev::kGather) {
              r.nf = -1;
              continue;
            }
            r.row_w = static_cast<long long>(c) * chunk_ * static_cast<long long>(sizeof(F));
            r.row_f = static_cast<long long>(c) * nv_ * static_cast<long long>(sizeof(F));
            // slots: class 2 faces, then class 3, then class 1, each in CSR order; class 0 left out
            int k = 0;
            for (const int want : {2, 3, 1}) {
              for (int q = e0; q < e1; ++q) {
                const gsk_csr_entry& e = s.csr[q];
                int cls;
                if (e.other < 0)
                  cls = s.bfaces[e.bface].kind == GSK_BC_PERIODIC ? 0 : 1;
                else if (e.other < c0 || e.other >= c1)
                  cls = 3;
                else
                  cls = (dirn == 0 ? pos[e.other] < pos[c] : pos[e.other] > pos[c]) ? 2 : 1;
                if (cls != want) continue;
                if (cls == 2) ++r.n_in;
                if (cls == 3) ++r.n_out;
                r.nb_w[k] = e.other >= 0 ? static_cast<long long>(e.other) * chunk_ * static_cast<long long>(sizeof(F)) : 0;
                r.sx[k] = static_cast<F>(e.s[0]);
                r.sy[k] = static_cast<F>(e.s[1]);
                r.sz[k] = static_cast<F>(e.s[2]);
                ++k;
              }
            }
            r.nf = k;
          }
          if (dirn == 0) total_levels += nlev;
        }
      }
    }
    fw_gb.push_back(static_cast<int>(fw_ptr.size()));
    bw_gb.push_back(static_cast<int>(bw_ptr.size()));
    // waves per workgroup ~ mean level width (cells a level offers to run side by side)
    const double width = total_levels > 0 ? static_cast<double>(n_dirs_eff_) * nc / static_cast<double>(total_levels) : 1.0;
    int w = 1;
    while (w < 8 && 2 * w <= width) w *= 2;
    // tuning knob: GSK_SWEEP_WAVES overrides the waves per sweep workgroup (1..8)
    if (const char* env = std::getenv("GSK_SWEEP_WAVES")) {
      const int e = std::atoi(env);
      if (e >= 1 && e <= dev::kMaxSweepThreads / 64) w = e;
    }
    waves_ = w;
    mean_level_width_ = width;
    // LDS stages: each level of each (direction, block), cut into pieces of at most cap records
    const int cap = dev::stage_capacity<F>(64 * w);
    GSK_CHECK(cap >= 1, "sweep stage holds no record");
    for (int dirn = 0; dirn < 2; ++dirn) {
      const std::vector<int>& ptr = dirn == 0 ? fw_ptr : bw_ptr;
      const std::vector<int>& gb = dirn == 0 ? fw_gb : bw_gb;
      const std::vector<dev::SweepRec<F>>& recs = dirn == 0 ? fw_recs : bw_recs;
      std::vector<int> stage_off, group_stage;
      const int n_groups = static_cast<int>(gb.size()) - 1;
      GSK_CHECK(n_groups == n_dirs_eff_ * P, "sweep level tables do not cover every (direction, block)");
      for (int q = 0; q < n_groups; ++q) {
        group_stage.push_back(static_cast<int>(stage_off.size()));
        const int l0 = gb[static_cast<std::size_t>(q)], l1 = gb[static_cast<std::size_t>(q) + 1] - 1;
        for (int l = l0; l < l1; ++l)
          for (int i = ptr[static_cast<std::size_t>(l)]; i < ptr[static_cast<std::size_t>(l) + 1]; i += cap)
            stage_off.push_back(i);
      }
      group_stage.push_back(static_cast<int>(stage_off.size()));
      stage_off.push_back(static_cast<int>(recs.size()));
      (dirn == 0 ? fw_recs_ : bw_recs_).upload(recs.data(), recs.size(), stream_);
      (dirn == 0 ? fw_stage_ : bw_stage_).upload(stage_off.data(), stage_off.size(), stream_);
      (dirn == 0 ? fw_group_ : bw_group_).upload(group_stage.data(), group_stage.size(), stream_);
    }
    // velocities grouped by direction (ascending storage position inside a direction)
    vel_list_host_.clear();
    dir_off_.assign(1, 0);
    for (int g = 0; g < n_dirs_eff_; ++g) {
      for (int p = 0; p < nv_; ++p)
        if (s.n_sweep_dirs == 0 || s.sweep_dir_of[p] == g) vel_list_host_.push_back(p);
      dir_off_.push_back(static_cast<int>(vel_list_host_.size()));
    }
    GSK_CHECK(static_cast<int>(vel_list_host_.size()) == nv_, "sweep direction table does not cover the velocities");
    vel_list_.upload(vel_list_host_.data(), vel_list_host_.size(), stream_);
  }

  struct EventPair {
    hipEvent_t a, b;
    int phase;
  };

  unsigned grid_cells() const { return static_cast<unsigned>(std::min(nc_, 1 << 20)); }
  std::size_t pinned_bytes() const {
    return std::max<std::size_t>({static_cast<std::size_t>(nc_) * dev::kMoments * sizeof(double),
                                  (static_cast<std::size_t>(nb_) * dev::kBoundarySums + 1) * sizeof(double),
                                  static_cast<std::size_t>(nc_) * sizeof(double), std::size_t(1) << 20});
  }

  template <class T>
  void download(T* dst, const T* src, std::size_t n) {
    const std::size_t bytes = n * sizeof(T);
    if (bytes <= pinned_bytes()) {
      HIPCHK(hipMemcpyAsync(pinned_, src, bytes, hipMemcpyDeviceToHost, stream_));
      HIPCHK(hipStreamSynchronize(stream_));
      std::memcpy(dst, pinned_, bytes);
    } else {
      HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_));
      HIPCHK(hipStreamSynchronize(stream_));
    }
  }

  void convert_in(const double* f, std::size_t n) {
    const std::size_t step = std::size_t(1) << 24;
    DBuf<double> stage;
    stage.alloc(std::min(n, step));
    for (std::size_t i = 0; i < n; i += step) {
      const std::size_t m = std::min(step, n - i);
      HIPCHK(hipMemcpyAsync(stage.p, f + i, m * sizeof(double), hipMemcpyHostToDevice, stream_));
      dev::k_from_double<F><<<1024, 256, 0, stream_>>>(stage.p, f_.p + i, static_cast<long long>(m));
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(stream_));
    }
  }
  void convert_out(double* f, std::size_t n) {
    const std::size_t step = std::size_t(1) << 24;
    DBuf<double> stage;
    stage.alloc(std::min(n, step));
    for (std::size_t i = 0; i < n; i += step) {
      const std::size_t m = std::min(step, n - i);
      dev::k_to_double<F><<<1024, 256, 0, stream_>>>(f_.p + i, stage.p, static_cast<long long>(m));
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(f + i, stage.p, m * sizeof(double), hipMemcpyDeviceToHost, stream_));
      HIPCHK(hipStreamSynchronize(stream_));
    }
  }

  void begin(int phase) {
    EventPair e;
    e.a = take_event();
    e.b = take_event();
    e.phase = phase;
    HIPCHK(hipEventRecord(e.a, stream_));
    events_.push_back(e);
  }
  void end() { HIPCHK(hipEventRecord(events_.back().b, stream_)); }
  hipEvent_t take_event() {
    if (!free_events_.empty()) {
      hipEvent_t e = free_events_.back();
      free_events_.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }

  // The velocity grid axis by axis (dev::AxisGrid) from the storage-order points: node i of an axis is
  // the point with index i on that axis and 0 on the others (as the f precision stores it); on a
  // stretched grid the weight factor of that node is the point's relative weight, and since the
  // relative weight of (i, j, k) is wx(i) wy(j) wz(k) / (wx(0) wy(0) wz(0))^2 times those of the
  // three points, the product of the three factors is scaled by 1 / w(0, 0, 0)^2.
  void upload_axes(const gsk_setup_view& s) {
    const int n0 = s.grid_n[0], n1 = s.grid_n[1], n2 = s.grid_n[2];
    GSK_CHECK(s.grid_to_storage != nullptr && static_cast<long long>(n0) * n1 * n2 == nv_,
              "velocity grid shape differs from the velocity count");
    auto point = [&](int i, int j, int k) { return s.grid_to_storage[i + n0 * (j + n1 * k)]; };
    std::vector<double> node, w;
    for (int i = 0; i < n0; ++i) node.push_back(static_cast<double>(static_cast<F>(s.xi_x[point(i, 0, 0)])));
    for (int j = 0; j < n1; ++j) node.push_back(static_cast<double>(static_cast<F>(s.xi_y[point(0, j, 0)])));
    for (int k = 0; k < n2; ++k) node.push_back(static_cast<double>(static_cast<F>(s.xi_z[point(0, 0, k)])));
    axes_.n[0] = n0;
    axes_.n[1] = n1;
    axes_.n[2] = n2;
    axes_.wscale = 1.0;
    axis_node_.upload(node.data(), node.size(), stream_);
    if (s.weight_rel != nullptr) {
      for (int i = 0; i < n0; ++i) w.push_back(s.weight_rel[point(i, 0, 0)]);
      for (int j = 0; j < n1; ++j) w.push_back(s.weight_rel[point(0, j, 0)]);
      for (int k = 0; k < n2; ++k) w.push_back(s.weight_rel[point(0, 0, k)]);
      const double w0 = s.weight_rel[point(0, 0, 0)];
      GSK_CHECK(w0 > 0.0, "velocity grid weight is not positive");
      axes_.wscale = 1.0 / (w0 * w0);
      axis_w_.upload(w.data(), w.size(), stream_);
    }
    axes_.node = axis_node_.p;
    axes_.w = axis_w_.p;
  }

  dev::View<F> view() {
    dev::View<F> v{};
    v.axes = axes_;
    v.n_cells = nc_;
    v.nv = nv_;
    v.n_bfaces = nb_;
    v.xi_x = xi_x_.p;
    v.xi_y = xi_y_.p;
    v.xi_z = xi_z_.p;
    v.weight = static_cast<F>(weight_);
    v.weight_d = weight_;
    v.wq = wq_.p;
    v.wqd = wqd_.p;
    v.nodes = nodes_.p;
    v.csr_off = csr_off_.p;
    v.csr = csr_.p;
    v.ell = ell_.p;
    v.eps2 = eps2_.p;
    v.limiter_mode = limiter_mode_;
    v.positivity = positivity_;
    v.diag_scale = static_cast<F>(diag_scale_);
    v.interface_damping = static_cast<F>(interface_damping_);
    v.bf = bfaces_.p;
    v.inflow = inflow_.p;
    v.cells = cells_.p;
    v.f = f_.p;
    v.mw_off = n_mw_ > 0 ? mw_off_.p : nullptr;
    v.mw_frac = n_mw_ > 0 ? mw_frac_.p : nullptr;
    v.specular = specular_.p;
    v.specular_first = specular_first_.p;
    v.grid_to_storage = grid_to_storage_.p;
    for (int a = 0; a < 3; ++a) v.grid_n[a] = grid_n_[a];
    v.n_sweep_dirs = n_sweep_dirs_;
    v.sweep_order = sweep_order_.p;
    v.sweep_pos = sweep_pos_.p;
    v.sweep_dir_of = sweep_dir_of_.p;
    v.deq = deq_;
    v.discrete_eq = discrete_eq_;
    v.deq_fallback = deq_fallback_;
    v.cold_bgk = cold_bgk_;
    v.lim_mode = lim_mode_;
    v.lim_branch = branch_.p;
    return v;
  }

  hipStream_t stream_ = nullptr;
  void* scratch_ = nullptr;  // this device's fixed scratch array (noalloc_scratch)
  int nc_ = 0, nv_ = 0, nb_ = 0, n_blocks_ = 1, block_passes_ = 1, negative_mode_ = 0;
  int chunk_ = 0, tiles_ = 0, n_sweep_dirs_ = 0, n_mw_ = 0;
  int grid_n_[3] = {0, 0, 0};
  double mu0_ = 0, omega_ = 0, Pr_ = 0, cold_bgk_ = 0, diag_scale_ = 1, interface_damping_ = 0, weight_ = 0;
  int discrete_eq_ = 0, deq_fallback_ = 0, limiter_mode_ = 1, positivity_ = 0;
  // adjoint workspace (adjoint_prepare)
  DBuf<double> adj_x_[3], adj_o_[5], adj_lp_, adj_jacp_, adj_jacm_, adj_lm_, adj_lb_, adj_ls_, adj_shift_, adj_lspec_,
      adj_lspec1_;
  DBuf<int> adj_mate_, adj_bslot_, adj_cells_;
  std::vector<int> adj_mate_host_, adj_bslot_host_, host_bface_owner_;
  int n_adj_cells_ = 0;
  int lim_mode_ = 0, max_faces_ = 0;  // limiter branches (set_limiter_branches); most faces of a cell
  DBuf<unsigned> branch_;
  dev::DeqGrid deq_{};
  DBuf<F> xi_x_, xi_y_, xi_z_;
  DBuf<double> xd_, yd_, zd_;
  DBuf<F> wq_;                   // stretched grid only (relative weights; empty: uniform)
  DBuf<double> wqd_, nodes_;
  dev::AxisGrid axes_{};
  DBuf<double> axis_node_, axis_w_;  // the grid axis by axis (upload_axes)
  DBuf<int> csr_off_;
  DBuf<dev::CsrE<F>> csr_;
  DBuf<dev::EllRow<F>> ell_;
  DBuf<int> irregular_;
  int n_irregular_ = 0;
  bool moments_mfma_ = false;
  double inflow_tol_ = 1e-10;  // discrete-equilibrium correction of inflow states
  long long grid_cap_ = 8LL << 20;  // workgroups per launch; GSK_GRID_CAP lowers it (test hook)
  DBuf<F> eps2_;
  DBuf<double> volume_, vdt_;
  DBuf<dev::BFace<F>> bfaces_;
  DBuf<dev::Inflow<F>> inflow_;
  DBuf<int> block_off_, sweep_order_, sweep_pos_;
  DBuf<unsigned char> sweep_dir_of_;
  DBuf<long long> mw_off_;
  DBuf<F> mw_frac_, specular_, specular_first_;
  DBuf<int> grid_to_storage_;
  DBuf<F> f_;
  DBuf<dev::CellP<F>> cells_;
  DBuf<double> moments_, shift_, residual_, bsums_, partial_;
  DBuf<int> flags_;
  DBuf<F> gx_, gy_, gz_, rres_, bufA_, bufB_;
  int n_dirs_eff_ = 1, waves_ = 1;
  double mean_level_width_ = 1.0;
  DBuf<int> fw_stage_, bw_stage_, fw_group_, bw_group_, vel_list_;
  DBuf<dev::SweepRec<F>> fw_recs_, bw_recs_;
  DBuf<int> subset_;
  std::vector<int> subset_host_;
  DBuf<int> wall_faces_, wall_kind_, wall_negative_;  // device wall refresh: faces, Maxwell flags, error flag
  DBuf<double> wall_unit0_, wall_scale_;             // unit-density sums s[0] of Maxwell walls; FP64 scales
  std::vector<int> wall_faces_host_, wall_kind_host_;
  std::vector<double> wall_unit0_host_, wall_scale_host_;
  // the device prepare (control_setup / control_prepare)
  bool ctl_ready_ = false, ctl_limiter_ = true, ctl_cell_mode_ = false, ctl_closed_ = false;
  int n_pf_ = 0, ctl_blocks_ = 1, ctl_max_patch_ = -1;
  double ctl_gamma_ = 5.0 / 3.0, ctl_u_floor_ = 0.0, ctl_relax_ = 1.0;
  DBuf<dev::CtrlFace> ctl_faces_;
  DBuf<dev::CtrlPatch> ctl_patches_;
  DBuf<dev::CtrlPressureFace> ctl_pf_;
  DBuf<dev::CtrlSlot> ctl_slots_;
  DBuf<int> ctl_pf_index_;
  DBuf<double> ctl_volume_, ctl_wall_unit_, ctl_part_, ctl_out_, prev_moments_;
  std::vector<double> ctl_host_;
  DBuf<double> bsum_part_, inflow_stage_;
  int bsum_chunks_ = 1;
  std::vector<int> vel_list_host_, dir_off_;
  std::vector<int> deq_flags_;
  std::vector<int> host_bface_kind_;
  void* pinned_ = nullptr;
  std::vector<EventPair> events_;
  std::vector<hipEvent_t> free_events_;
  double timings_[kNumPhases];
};

}  // namespace

std::unique_ptr<Backend> make_hip_backend(const gsk_setup_view& view) {
  int ndev = 0;
  const hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    throw Error("no HIP device is available; the kinetic backend runs on the GPU only");
  }
  if (view.precision_fp64) return std::make_unique<HipBackend<double>>(view);
  return std::make_unique<HipBackend<float>>(view);
}

}  // namespace gsk
