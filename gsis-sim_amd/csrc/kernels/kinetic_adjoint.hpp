#pragma once
// Adjoint (transposed Jacobian) of the steady residual R(f), of the boundary-face sums and of the
// moment sums, FP64, on the limiter branches recorded at the linearisation state f* (no counterpart
// in the reference, which has no derivative capability). The forward operators transposed here are
// the residual and limited gradient of kinetic_view.hpp (the reference's residual_value and
// limited_gradient, src/kinetic/kinetic_kernels.hpp:165-278), the Maxwell-wall reflections
// (reflected_value, :80-127), the boundary sums (k_boundary_sums) and the moment sums (k_moments).
//
// With lam the vector J^T is applied to, one (cell c, velocity p) of the residual reads
//   f and the limited gradient G = phi g of c (own outflow faces), of the upwind neighbour (inflow
//   faces), of c at the mirror velocity (symmetry faces), inflow values of boundary faces, the
//   reflected caches of Maxwell walls, and the cell parameters P_c (collision term);
// G reads f of c and of its face neighbours, first-order boundary values and, through phi, the
// recorded extreme values and projections. The kernels below gather these dependencies per (c, p)
// so that every output element has one writer, except the boundary scatters (reflections, boundary
// sums), which add with atomics. The small, non-linear maps between f and those inputs (moments ->
// cell parameters through the discrete-equilibrium Newton; the host's boundary model) enter as
// Jacobians tabulated once at f* (k_adj_param_jacobian here; the host for the boundary model).
//
// Workspace, [cell][velocity] unless stated:
//   X (3)      adjoint of G, accumulated before the gradient adjoint runs
//   O (5)      gradient adjoint: adjoint of the unlimited g (3), of the recorded f_max and f_min
//   lspec      [Maxwell face][velocity] adjoint of the reconstructed reflected cache (specular)
//   lspec1     the same for the first-order cache (specular_first)
//   lP         [cell][kAdjPStride] adjoint of the cell parameters (layout below)

namespace gsk::dev {

constexpr int kAdjBlock = 256;
// cell parameters as one vector: 0 A, 1 B, 2-4 equilibrium u, 5-7 u of the polynomial, 8-10 q,
// 11 theta, 12 coef, 13 g0, 14 g4, 15-17 g, 18 collision rate V / tau
constexpr int kAdjParams = 19;
constexpr int kAdjPStride = 20;

__device__ __forceinline__ void params_to_vector(const CellP<double>& cp, double (&P)[kAdjParams]) {
  P[0] = cp.s.eq.A;
  P[1] = cp.s.eq.B;
  P[2] = cp.s.eq.ux;
  P[3] = cp.s.eq.uy;
  P[4] = cp.s.eq.uz;
  P[5] = cp.s.ucx;
  P[6] = cp.s.ucy;
  P[7] = cp.s.ucz;
  P[8] = cp.s.qx;
  P[9] = cp.s.qy;
  P[10] = cp.s.qz;
  P[11] = cp.s.theta;
  P[12] = cp.s.coef;
  P[13] = cp.s.g0;
  P[14] = cp.s.g4;
  P[15] = cp.s.gx;
  P[16] = cp.s.gy;
  P[17] = cp.s.gz;
  P[18] = cp.coll;
}

// partial derivatives of the unclamped limiter value r = num / den (venkat_frozen, clamp 0)
__device__ __forceinline__ void venkat_partials(double d2, double d1, double eps2, double& r_d2, double& r_d1,
                                                double& r_eps2) {
  const double num = (d1 * d1 + eps2) + 2.0 * d1 * d2;
  const double den = d1 * d1 + 2.0 * d2 * d2 + d1 * d2 + eps2;
  const double inv = 1.0 / (den * den);
  r_d2 = (2.0 * d1 * den - num * (4.0 * d2 + d1)) * inv;
  r_d1 = ((2.0 * d1 + 2.0 * d2) * den - num * (2.0 * d1 + d2)) * inv;
  r_eps2 = (den - num) * inv;
}

// Adjoint of one limited gradient G = phi g at (cell, p) on its recorded branch: given lG = dL/dG,
// returns dL/dg (lg), dL/d f_max and dL/d f_min for the recorded extreme faces (lmx, lmn; zero when
// the extreme is the cell's own value, whose share is in lfi) and dL/d f_i through phi (lfi). The
// dependence of g on the face values, g = sum_k c_k (f_k - f_i), is left to the caller.
template <int N>
__device__ __forceinline__ void gradient_adjoint_n(const View<double>& v, int cell, int p, double x, double y, double z,
                                                   double fi, double lGx, double lGy, double lGz, double& lgx,
                                                   double& lgy, double& lgz, double& lmx, double& lmn, double& lfi) {
  lgx = lGx;
  lgy = lGy;
  lgz = lGz;
  lmx = lmn = lfi = 0.0;
  if (v.limiter_mode == kLimNone && v.positivity == 0) return;  // G = g
  const unsigned code = v.lim_branch[static_cast<i64>(cell) * v.nv + p];
  const unsigned pos = (code >> 22) & 3u;
  if (v.positivity != 0 && pos == 2u) {  // G = 0
    lgx = lgy = lgz = 0.0;
    return;
  }
  const int e0 = v.csr_off[cell], n = v.csr_off[cell + 1] - e0;
  double fk[N], dsx[N], dsy[N], dsz[N];
  double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    fk[k] = dsx[k] = dsy[k] = dsz[k] = 0.0;
    if (k >= n) continue;
    const CsrE<double>& e = v.csr[e0 + k];
    fk[k] = e.other >= 0 ? f_at(v, e.other, p) : boundary_value_first_order(v, e, cell, p, x, y, z, fi);
    const double d = fk[k] - fi;
    gx += e.cx * d;
    gy += e.cy * d;
    gz += e.cz * d;
    dsx[k] = e.dsx;
    dsy[k] = e.dsy;
    dsz[k] = e.dsz;
  }
  const unsigned kmx = code & 15u, kmn = (code >> 4) & 15u, imx = (code >> 8) & 15u, imn = (code >> 12) & 15u;
  const double fmx = imx != kNoCandidate ? pick(fk, imx) : fi;
  const double fmn = imn != kNoCandidate ? pick(fk, imn) : fi;
  const double pmx_x = kmx != kNoCandidate ? pick(dsx, kmx) : 0.0, pmx_y = kmx != kNoCandidate ? pick(dsy, kmx) : 0.0,
               pmx_z = kmx != kNoCandidate ? pick(dsz, kmx) : 0.0;
  const double pmn_x = kmn != kNoCandidate ? pick(dsx, kmn) : 0.0, pmn_y = kmn != kNoCandidate ? pick(dsy, kmn) : 0.0,
               pmn_z = kmn != kNoCandidate ? pick(dsz, kmn) : 0.0;
  const double proj_max = dot3(gx, gy, gz, pmx_x, pmx_y, pmx_z), proj_min = dot3(gx, gy, gz, pmn_x, pmn_y, pmn_z);
  const double eps_base = static_cast<double>(v.eps2[cell]);
  double eps2 = eps_base;
  const bool rel = v.limiter_mode == kLimRel, eps_from_max = ((code >> 16) & 1u) != 0u;
  if (rel) {
    const double s = eps_from_max ? dabs(fmx) : dabs(fmn);
    eps2 *= s * s;
  }
  const double dmx = fmx - fi, dmn = fmn - fi;
  const unsigned cmx = (code >> 17) & 3u, cmn = (code >> 19) & 3u, side = (code >> 21) & 1u;
  double phi = 1.0;
  if (v.limiter_mode != kLimNone) {
    const double vmx = kmx == kNoCandidate ? 1.0 : venkat_frozen(proj_max, dmx, eps2, cmx);
    const double vmn = kmn == kNoCandidate ? 1.0 : venkat_frozen(proj_min, dmn, eps2, cmn);
    phi = side ? vmn : vmx;
  }
  const bool capped = v.positivity != 0 && pos == 1u;
  if (capped) phi = fi / (-proj_min);
  const double lphi = dot3(gx, gy, gz, lGx, lGy, lGz);
  lgx = phi * lGx;
  lgy = phi * lGy;
  lgz = phi * lGz;
  double lpmax = 0.0, lpmin = 0.0, ldmx = 0.0, ldmn = 0.0, leps = 0.0;
  if (capped) {
    lfi += lphi / (-proj_min);
    lpmin += lphi * fi / (proj_min * proj_min);
  } else if (v.limiter_mode != kLimNone) {
    double a, b, c;
    if (!side) {
      if (kmx != kNoCandidate && cmx == 0u) {
        venkat_partials(proj_max, dmx, eps2, a, b, c);
        lpmax = lphi * a;
        ldmx = lphi * b;
        leps = lphi * c;
      }
    } else if (kmn != kNoCandidate && cmn == 0u) {
      venkat_partials(proj_min, dmn, eps2, a, b, c);
      lpmin = lphi * a;
      ldmn = lphi * b;
      leps = lphi * c;
    }
  }
  lgx += lpmax * pmx_x + lpmin * pmn_x;
  lgy += lpmax * pmx_y + lpmin * pmn_y;
  lgz += lpmax * pmx_z + lpmin * pmn_z;
  double lfmx = ldmx, lfmn = ldmn;
  lfi -= ldmx + ldmn;
  if (rel && leps != 0.0) {
    if (eps_from_max)
      lfmx += leps * 2.0 * eps_base * fmx;
    else
      lfmn += leps * 2.0 * eps_base * fmn;
  }
  if (imx == kNoCandidate) {
    lfi += lfmx;
    lfmx = 0.0;
  }
  if (imn == kNoCandidate) {
    lfi += lfmn;
    lfmn = 0.0;
  }
  lmx = lfmx;
  lmn = lfmn;
}

__device__ __attribute__((noinline)) void gradient_adjoint_wide(const View<double>& v, int cell, int p, double x, double y,
                                                               double z, double fi, double lGx, double lGy, double lGz,
                                                               double& lgx, double& lgy, double& lgz, double& lmx,
                                                               double& lmn, double& lfi) {
  gradient_adjoint_n<kLimMaxFaces>(v, cell, p, x, y, z, fi, lGx, lGy, lGz, lgx, lgy, lgz, lmx, lmn, lfi);
}

// dL/d f_k for slot k of (cell, p), from the gradient adjoint's outputs at (cell, p)
__device__ __forceinline__ double slot_adjoint(const View<double>& v, i64 ip, unsigned k, double cx, double cy, double cz,
                                               const double* Ogx, const double* Ogy, const double* Ogz,
                                               const double* Omx, const double* Omn) {
  double r = cx * Ogx[ip] + cy * Ogy[ip] + cz * Ogz[ip];
  if (v.limiter_mode != kLimNone || v.positivity != 0) {
    const unsigned code = v.lim_branch[ip];
    if (((code >> 8) & 15u) == k) r += Omx[ip];
    if (((code >> 12) & 15u) == k) r += Omn[ip];
  }
  return r;
}

// ---- 1. transport: the residual's direct dependence on f and G (gather per (c, p)) ----
// out = coll lam + sum over the faces whose flux reads f_{c,p}; X = sum over the faces whose flux
// reads G_{c,p}; lspec[Maxwell face][p] = the adjoint of the reflected cache in the flux.
__global__ __launch_bounds__(kAdjBlock) void k_adj_transport(View<double> v, const int* __restrict__ mate,
                                                             const double* __restrict__ lam, double* __restrict__ out,
                                                             double* __restrict__ Xx, double* __restrict__ Xy,
                                                             double* __restrict__ Xz, double* __restrict__ lspec,
                                                             int tiles) {
  const i64 items = static_cast<i64>(v.n_cells) * tiles;
  for (i64 it = blockIdx.x; it < items; it += gridDim.x) {
    const int c = static_cast<int>(it / tiles);
    const int p = static_cast<int>(it % tiles) * kAdjBlock + static_cast<int>(threadIdx.x);
    if (p >= v.nv) continue;
    const double x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
    const i64 ip = static_cast<i64>(c) * v.nv + p;
    const double lc = lam ? lam[ip] : 0.0;
    double df = v.cells[c].coll * lc, gx = 0.0, gy = 0.0, gz = 0.0;
    const int e0 = v.csr_off[c], e1 = v.csr_off[c + 1];
    for (int k = e0; k < e1; ++k) {
      const CsrE<double>& e = v.csr[k];
      int kind = -1;
      if (e.other < 0) {
        kind = v.bf[e.bface].kind;
        if (kind == kBcPeriodic) continue;
      }
      const double a = x * e.sx + y * e.sy + z * e.sz;
      if (a >= 0.0) {  // own outflow face: f_c + G_c . ds
        const double la = lc * a;
        df += la;
        gx += la * e.dsx;
        gy += la * e.dsy;
        gz += la * e.dsz;
      }
      if (e.other >= 0) {
        // the neighbour's flux over this face reads f_c + G_c . do when the face is its inflow face
        const CsrE<double>& en = v.csr[v.csr_off[e.other] + mate[k]];
        const double an = x * en.sx + y * en.sy + z * en.sz;
        if (an < 0.0 && lam) {
          const double la = lam[static_cast<i64>(e.other) * v.nv + p] * an;
          df += la;
          gx += la * en.dox;
          gy += la * en.doy;
          gz += la * en.doz;
        }
      } else if (kind == kBcSymmetry) {
        // the flux at the mirror velocity pm reads f_{c,p} + G_{c,p} . ds when it enters
        const int pm = p ^ v.bf[e.bface].mirror_xor;
        const double am = v.xi_x[pm] * e.sx + v.xi_y[pm] * e.sy + v.xi_z[pm] * e.sz;
        if (am < 0.0 && lam) {
          const double la = lam[static_cast<i64>(c) * v.nv + pm] * am;
          df += la;
          gx += la * e.dsx;
          gy += la * e.dsy;
          gz += la * e.dsz;
        }
      } else if (kind == kBcMaxwell && a < 0.0) {
        lspec[v.mw_off[e.bface] + p] = lc * a * v.mw_frac[e.bface];
      }
    }
    out[ip] = df;
    Xx[ip] = gx;
    Xy[ip] = gy;
    Xz[ip] = gz;
  }
}

// ---- 2. collision term -coll (f_S(P) - f): adjoint of the cell parameters (block per cell) ----
__global__ __launch_bounds__(kAdjBlock) void k_adj_collision(View<double> v, const double* __restrict__ lam,
                                                             double* __restrict__ lP) {
  __shared__ double lds[kWaves * kAdjParams];
  for (int c = blockIdx.x; c < v.n_cells; c += gridDim.x) {
    double acc[kAdjParams];
#pragma unroll
    for (int k = 0; k < kAdjParams; ++k) acc[k] = 0.0;
    const CellP<double> cp = v.cells[c];
    const Target<double>& s = cp.s;
    if (lam) {
      for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
        const i64 ip = static_cast<i64>(c) * v.nv + p;
        const double l = lam[ip];
        if (l == 0.0) continue;
        const double x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
        const double ex = x - s.eq.ux, ey = y - s.eq.uy, ez = z - s.eq.uz;
        const double r2 = ex * ex + ey * ey + ez * ez;
        const double E = ::exp(-s.eq.B * r2), M = s.eq.A * E;
        const double cx = x - s.ucx, cy = y - s.ucy, cz = z - s.ucz;
        const double c2 = cx * cx + cy * cy + cz * cz, cq = cx * s.qx + cy * s.qy + cz * s.qz;
        const double k5 = c2 / s.theta - 5.0;
        const double Q = 1.0 + s.coef * cq * k5 + (s.g0 + cx * s.gx + cy * s.gy + cz * s.gz) + 0.5 * s.g4 * c2;
        const double w = -l * cp.coll;
        const double MQ = M * Q;
        acc[0] += w * E * Q;
        acc[1] -= w * r2 * MQ;
        acc[2] += w * 2.0 * s.eq.B * ex * MQ;
        acc[3] += w * 2.0 * s.eq.B * ey * MQ;
        acc[4] += w * 2.0 * s.eq.B * ez * MQ;
        const double t1 = s.coef * k5, t2 = s.coef * cq * 2.0 / s.theta;
        acc[5] -= w * M * (t1 * s.qx + t2 * cx + s.gx + s.g4 * cx);
        acc[6] -= w * M * (t1 * s.qy + t2 * cy + s.gy + s.g4 * cy);
        acc[7] -= w * M * (t1 * s.qz + t2 * cz + s.gz + s.g4 * cz);
        acc[8] += w * M * t1 * cx;
        acc[9] += w * M * t1 * cy;
        acc[10] += w * M * t1 * cz;
        acc[11] -= w * M * s.coef * cq * c2 / (s.theta * s.theta);
        acc[12] += w * M * cq * k5;
        acc[13] += w * M;
        acc[14] += w * M * 0.5 * c2;
        acc[15] += w * M * cx;
        acc[16] += w * M * cy;
        acc[17] += w * M * cz;
        acc[18] -= l * (MQ - v.f[ip]);
      }
    }
    block_sum<kAdjParams>(acc, lds);
    if (threadIdx.x == 0)
#pragma unroll
      for (int k = 0; k < kAdjParams; ++k) lP[static_cast<i64>(c) * kAdjPStride + k] = acc[k];
  }
}

// ---- 3. Maxwell-wall reflections: scatter the adjoint of a reflected cache (block (velocities,
// face)); Reconstruct: the cache of reconstructed values (f + G . d), else of cell values. The
// Maxwellian-ratio weights read the owner's equilibrium (B, u), whose adjoint is added to lP. ----
template <bool Reconstruct>
__global__ __launch_bounds__(kAdjBlock) void k_adj_reflection(View<double> v, const double* __restrict__ lspec,
                                                              double* __restrict__ out, double* __restrict__ Xx,
                                                              double* __restrict__ Xy, double* __restrict__ Xz,
                                                              double* __restrict__ lP) {
  __shared__ double lds[kWaves * 4];
  const int b = blockIdx.y;
  if (b >= v.n_bfaces || v.mw_off[b] < 0) return;  // the whole block
  const BFace<double> bf = v.bf[b];
  const int c = bf.owner;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};  // owner's equilibrium B, ux, uy, uz
  const double L = p < v.nv ? lspec[v.mw_off[b] + p] : 0.0;
  const double x = p < v.nv ? v.xi_x[p] : 0.0, y = p < v.nv ? v.xi_y[p] : 0.0, z = p < v.nv ? v.xi_z[p] : 0.0;
  if (p < v.nv && L != 0.0 && (x * bf.nx + y * bf.ny + z * bf.nz) < 0.0) {
    const i64 row = static_cast<i64>(c) * v.nv;
    if (bf.mirror_xor != 0) {
      const int q = p ^ bf.mirror_xor;
      atomicAdd(&out[row + q], L);
      if (Reconstruct) {
        atomicAdd(&Xx[row + q], L * bf.dx);
        atomicAdd(&Xy[row + q], L * bf.dy);
        atomicAdd(&Xz[row + q], L * bf.dz);
      }
    } else {
      // the forward's interpolation (reflected_value): bracketing nodes and weights of the reflected point
      const double un = x * bf.nx + y * bf.ny + z * bf.nz;
      const double r[3] = {x - 2.0 * un * bf.nx, y - 2.0 * un * bf.ny, z - 2.0 * un * bf.nz};
      int base[3];
      double t[3];
      int node_off = 0;
      bool inside = true;
      for (int a = 0; a < 3 && inside; ++a) {
        if (v.nodes != nullptr) {
          const double* nd = v.nodes + node_off;
          const int n = v.grid_n[a];
          node_off += n;
          const double ra = r[a];
          if (ra < nd[0] - 1e-5 * (nd[1] - nd[0]) || ra > nd[n - 1] + 1e-5 * (nd[n - 1] - nd[n - 2])) {
            inside = false;
            break;
          }
          int lo = 0, hi = n - 1;
          while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (nd[mid] <= ra)
              lo = mid;
            else
              hi = mid;
          }
          base[a] = lo;
          t[a] = dmax(0.0, dmin(1.0, (ra - nd[lo]) / (nd[lo + 1] - nd[lo])));
          continue;
        }
        const double zz = (r[a] - v.deq.lo[a]) / v.deq.d[a] - 0.5;
        if (zz < -1e-5 || zz > v.grid_n[a] - 1.0 + 1e-5) {
          inside = false;
          break;
        }
        const double cl = dmax(0.0, dmin(zz, v.grid_n[a] - 1.0));
        base[a] = static_cast<int>(cl);
        if (base[a] >= v.grid_n[a] - 1) base[a] = v.grid_n[a] - 2;
        t[a] = cl - base[a];
      }
      if (inside) {
        const Maxw<double> ref = v.cells[c].s.eq;
        for (int k = 0; k < 8; ++k) {
          const int ix = base[0] + (k & 1), iy = base[1] + ((k >> 1) & 1), iz = base[2] + ((k >> 2) & 1);
          const double w = ((k & 1) ? t[0] : 1.0 - t[0]) * ((k & 2) ? t[1] : 1.0 - t[1]) * ((k & 4) ? t[2] : 1.0 - t[2]);
          if (w == 0.0) continue;
          const int q = v.grid_to_storage[ix + v.grid_n[0] * (iy + v.grid_n[1] * iz)];
          const double qx = v.xi_x[q], qy = v.xi_y[q], qz = v.xi_z[q];
          const double drx = r[0] - ref.ux, dry = r[1] - ref.uy, drz = r[2] - ref.uz;
          const double dqx = qx - ref.ux, dqy = qy - ref.uy, dqz = qz - ref.uz;
          const double dd = drx * drx + dry * dry + drz * drz - dqx * dqx - dqy * dqy - dqz * dqz;
          const double ratio = ::exp(-ref.B * dd);
          const double lw = L * w * ratio;
          atomicAdd(&out[row + q], lw);
          double value;
          if (Reconstruct) {
            atomicAdd(&Xx[row + q], lw * bf.dx);
            atomicAdd(&Xy[row + q], lw * bf.dy);
            atomicAdd(&Xz[row + q], lw * bf.dz);
            value = reconstruct_at_boundary<2>(v, bf, q, qx, qy, qz);
          } else {
            value = f_at(v, c, q);
          }
          const double lr = L * w * value * ratio;  // times d ratio / d(B, u) over ratio
          acc[0] -= lr * dd;
          acc[1] += lr * 2.0 * ref.B * (r[0] - qx);
          acc[2] += lr * 2.0 * ref.B * (r[1] - qy);
          acc[3] += lr * 2.0 * ref.B * (r[2] - qz);
        }
      }
    }
  }
  block_sum<4>(acc, lds);
  if (threadIdx.x == 0 && (acc[0] != 0.0 || acc[1] != 0.0 || acc[2] != 0.0 || acc[3] != 0.0)) {
    double* P = lP + static_cast<i64>(c) * kAdjPStride;
    atomicAdd(&P[1], acc[0]);
    atomicAdd(&P[2], acc[1]);
    atomicAdd(&P[3], acc[2]);
    atomicAdd(&P[4], acc[3]);
  }
}

// ---- 4. gradient adjoint per (c, p) from X: O = (lg, lmx, lmn) (Acc: added to O), out += lfi.
// cells: optional list (each cell once); else all cells. ----
template <bool Acc>
__global__ __launch_bounds__(kAdjBlock) void k_adj_gradient(View<double> v, const int* __restrict__ cells, int n_list,
                                                            const double* __restrict__ Xx, const double* __restrict__ Xy,
                                                            const double* __restrict__ Xz, double* __restrict__ Ogx,
                                                            double* __restrict__ Ogy, double* __restrict__ Ogz,
                                                            double* __restrict__ Omx, double* __restrict__ Omn,
                                                            double* __restrict__ out, int tiles) {
  const int ncell = cells ? n_list : v.n_cells;
  const i64 items = static_cast<i64>(ncell) * tiles;
  for (i64 it = blockIdx.x; it < items; it += gridDim.x) {
    const int ci = static_cast<int>(it / tiles);
    const int c = cells ? cells[ci] : ci;
    const int p = static_cast<int>(it % tiles) * kAdjBlock + static_cast<int>(threadIdx.x);
    if (p >= v.nv) continue;
    const i64 ip = static_cast<i64>(c) * v.nv + p;
    const double lGx = Xx[ip], lGy = Xy[ip], lGz = Xz[ip];
    double lgx = lGx, lgy = lGy, lgz = lGz, lmx = 0.0, lmn = 0.0, lfi = 0.0;
    if (lGx != 0.0 || lGy != 0.0 || lGz != 0.0) {
      const double x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p], fi = v.f[ip];
      if (v.csr_off[c + 1] - v.csr_off[c] <= kGather)
        gradient_adjoint_n<kGather>(v, c, p, x, y, z, fi, lGx, lGy, lGz, lgx, lgy, lgz, lmx, lmn, lfi);
      else
        gradient_adjoint_wide(v, c, p, x, y, z, fi, lGx, lGy, lGz, lgx, lgy, lgz, lmx, lmn, lfi);
    }
    if (Acc) {
      Ogx[ip] += lgx;
      Ogy[ip] += lgy;
      Ogz[ip] += lgz;
      Omx[ip] += lmx;
      Omn[ip] += lmn;
    } else {
      Ogx[ip] = lgx;
      Ogy[ip] = lgy;
      Ogz[ip] = lgz;
      Omx[ip] = lmx;
      Omn[ip] = lmn;
    }
    out[ip] += lfi;
  }
}

// ---- 5. the gradients' dependence on f, gathered per (c, p): -sum_k c_k lg (own value), the
// neighbours' lg, lmx, lmn through their slot for this face, and the cell's first-order boundary
// values (outgoing and periodic: f_i; symmetry: half of f_i and of the mirror value; Maxwell wall
// incoming: the first-order reflected cache, whose adjoint goes to lspec1). ----
__global__ __launch_bounds__(kAdjBlock) void k_adj_gather(View<double> v, const int* __restrict__ mate,
                                                          const double* __restrict__ Ogx, const double* __restrict__ Ogy,
                                                          const double* __restrict__ Ogz, const double* __restrict__ Omx,
                                                          const double* __restrict__ Omn, double* __restrict__ out,
                                                          double* __restrict__ lspec1, int tiles) {
  const i64 items = static_cast<i64>(v.n_cells) * tiles;
  for (i64 it = blockIdx.x; it < items; it += gridDim.x) {
    const int c = static_cast<int>(it / tiles);
    const int p = static_cast<int>(it % tiles) * kAdjBlock + static_cast<int>(threadIdx.x);
    if (p >= v.nv) continue;
    const double x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
    const i64 ip = static_cast<i64>(c) * v.nv + p;
    const double lgx = Ogx[ip], lgy = Ogy[ip], lgz = Ogz[ip];
    double lf = 0.0;
    const int e0 = v.csr_off[c], e1 = v.csr_off[c + 1];
    for (int k = e0; k < e1; ++k) {
      const CsrE<double>& e = v.csr[k];
      if (e.other < 0 && v.bf[e.bface].kind == kBcPeriodic) continue;  // f_k = f_i: no dependence
      lf -= e.cx * lgx + e.cy * lgy + e.cz * lgz;
      const unsigned slot = static_cast<unsigned>(k - e0);
      if (e.other >= 0) {
        const unsigned m = static_cast<unsigned>(mate[k]);
        const CsrE<double>& en = v.csr[v.csr_off[e.other] + static_cast<int>(m)];
        lf += slot_adjoint(v, static_cast<i64>(e.other) * v.nv + p, m, en.cx, en.cy, en.cz, Ogx, Ogy, Ogz, Omx, Omn);
        continue;
      }
      const BFace<double>& bf = v.bf[e.bface];
      const double lfk = slot_adjoint(v, ip, slot, e.cx, e.cy, e.cz, Ogx, Ogy, Ogz, Omx, Omn);
      if (bf.kind == kBcSymmetry) {
        const i64 im = static_cast<i64>(c) * v.nv + (p ^ bf.mirror_xor);
        lf += 0.5 * (lfk + slot_adjoint(v, im, slot, e.cx, e.cy, e.cz, Ogx, Ogy, Ogz, Omx, Omn));
        continue;
      }
      const double a = x * e.sx + y * e.sy + z * e.sz;
      if (a >= 0.0)
        lf += lfk;
      else if (bf.kind == kBcMaxwell)
        lspec1[v.mw_off[e.bface] + p] = v.mw_frac[e.bface] * lfk;
    }
    out[ip] += lf;
  }
}

// ---- 6. adjoint of the inflow states' device parameters (scale, A, B, ux, uy, uz) per boundary
// face from the residual: the inflow flux a I(xi) and the first-order boundary value I(xi) in the
// owner's gradient (Maxwell walls: the scaled wall Maxwellian part). bslot: the face's slot in its
// owner's row. ----
__global__ __launch_bounds__(kAdjBlock) void k_adj_inflow(View<double> v, const int* __restrict__ bslot,
                                                          const double* __restrict__ lam, const double* __restrict__ Ogx,
                                                          const double* __restrict__ Ogy, const double* __restrict__ Ogz,
                                                          const double* __restrict__ Omx, const double* __restrict__ Omn,
                                                          double* __restrict__ lb) {
  __shared__ double lds[kWaves * 6];
  for (int j = blockIdx.x; j < v.n_bfaces; j += gridDim.x) {
    const BFace<double> bf = v.bf[j];
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (bf.kind != kBcSymmetry && bf.kind != kBcPeriodic) {
      const int c = bf.owner;
      const unsigned slot = static_cast<unsigned>(bslot[j]);
      const CsrE<double>& e = v.csr[v.csr_off[c] + bslot[j]];
      const Inflow<double> in = v.inflow[j];
      for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
        const double x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
        const double a = x * e.sx + y * e.sy + z * e.sz;
        if (a >= 0.0) continue;
        const i64 ip = static_cast<i64>(c) * v.nv + p;
        const double l = (lam ? lam[ip] * a : 0.0) + slot_adjoint(v, ip, slot, e.cx, e.cy, e.cz, Ogx, Ogy, Ogz, Omx, Omn);
        if (l == 0.0) continue;
        const double ex = x - in.m.ux, ey = y - in.m.uy, ez = z - in.m.uz;
        const double r2 = ex * ex + ey * ey + ez * ez;
        const double E = ::exp(-in.m.B * r2);
        const double sAE = in.scale * in.m.A * E;
        acc[0] += l * in.m.A * E;
        acc[1] += l * in.scale * E;
        acc[2] -= l * sAE * r2;
        acc[3] += l * 2.0 * in.m.B * sAE * ex;
        acc[4] += l * 2.0 * in.m.B * sAE * ey;
        acc[5] += l * 2.0 * in.m.B * sAE * ez;
      }
    }
    block_sum<6>(acc, lds);
    if (threadIdx.x == 0)
#pragma unroll
      for (int k = 0; k < 6; ++k) lb[static_cast<i64>(j) * 6 + k] = acc[k];
  }
}

// ---- 7. boundary sums: their dependence on f (block per face). ls: adjoint of the 20 sums per face
// (slots as k_boundary_reduce writes them). Outgoing velocities read the reconstructed value f + G . d
// of the owner; incoming ones of symmetry faces the same at the mirror velocity; incoming ones of
// Maxwell walls the reflected cache (its adjoint to lspec). ----
__global__ __launch_bounds__(kAdjBlock) void k_adj_sums(View<double> v, const double* __restrict__ ls,
                                                        double* __restrict__ out, double* __restrict__ Xx,
                                                        double* __restrict__ Xy, double* __restrict__ Xz,
                                                        double* __restrict__ lspec) {
  for (int j = blockIdx.x; j < v.n_bfaces; j += gridDim.x) {
    const BFace<double> bf = v.bf[j];
    double L[kBoundarySums];
#pragma unroll
    for (int k = 0; k < kBoundarySums; ++k) L[k] = ls[static_cast<i64>(j) * kBoundarySums + k];
    const int c = bf.owner;
    const i64 row = static_cast<i64>(c) * v.nv;
    for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
      const double x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
      const double an = x * bf.nx + y * bf.ny + z * bf.nz;
      const bool outgoing = an >= 0.0 || bf.kind == kBcPeriodic;
      const int fl = outgoing ? 0 : 5, mo = outgoing ? 10 : 15;
      const double e = 0.5 * (x * x + y * y + z * z);
      double ww = v.weight_d;
      if (v.wqd != nullptr) ww *= v.wqd[p];
      const double flux = L[fl] + L[fl + 1] * x + L[fl + 2] * y + L[fl + 3] * z + L[fl + 4] * e;
      const double mom = L[mo] + L[mo + 1] * x + L[mo + 2] * y + L[mo + 3] * z + L[mo + 4] * e;
      const double lfb = ww * (an * flux + mom);
      if (lfb == 0.0) continue;
      if (outgoing) {
        atomicAdd(&out[row + p], lfb);
        if (bf.kind != kBcPeriodic) {
          atomicAdd(&Xx[row + p], lfb * bf.dx);
          atomicAdd(&Xy[row + p], lfb * bf.dy);
          atomicAdd(&Xz[row + p], lfb * bf.dz);
        }
      } else if (bf.kind == kBcSymmetry) {
        const int pm = p ^ bf.mirror_xor;
        atomicAdd(&out[row + pm], lfb);
        atomicAdd(&Xx[row + pm], lfb * bf.dx);
        atomicAdd(&Xy[row + pm], lfb * bf.dy);
        atomicAdd(&Xz[row + pm], lfb * bf.dz);
      } else if (bf.kind == kBcMaxwell) {
        lspec[v.mw_off[j] + p] = v.mw_frac[j] * lfb;
      }
    }
  }
}

// ---- 8. moment sums: out += d S / d f^T (jacP^T lP + jacM^T lm) per cell (block per cell). S are
// the 13 raw sums about the shift u0 (moment_terms, weight wqd[p]); jacP [cell][19][13] and jacM
// [cell][15][13] their Jacobians tabulated at f*; lm (may be null): adjoint of the 15 moments. ----
__global__ __launch_bounds__(kAdjBlock) void k_adj_moments(View<double> v, const double* __restrict__ shift,
                                                           const double* __restrict__ jacP, const double* __restrict__ jacM,
                                                           const double* __restrict__ lP, const double* __restrict__ lm,
                                                           double* __restrict__ out) {
  __shared__ double sS[kMomentSums];
  for (int c = blockIdx.x; c < v.n_cells; c += gridDim.x) {
    if (threadIdx.x < kMomentSums) {
      const int k = static_cast<int>(threadIdx.x);
      double s = 0.0;
      for (int i = 0; i < kAdjParams; ++i)
        s += jacP[(static_cast<i64>(c) * kAdjParams + i) * kMomentSums + k] * lP[static_cast<i64>(c) * kAdjPStride + i];
      if (lm)
        for (int i = 0; i < kMoments; ++i)
          s += jacM[(static_cast<i64>(c) * kMoments + i) * kMomentSums + k] * lm[static_cast<i64>(c) * kMoments + i];
      sS[k] = s;
    }
    __syncthreads();
    const double u0x = shift[c * 3 + 0], u0y = shift[c * 3 + 1], u0z = shift[c * 3 + 2];
    for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
      double t[kMomentSums];
      moment_terms<double>(v.wqd != nullptr ? v.wqd[p] : 1.0, v.xi_x[p] - u0x, v.xi_y[p] - u0y, v.xi_z[p] - u0z, t);
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < kMomentSums; ++k) s += sS[k] * t[k];
      out[static_cast<i64>(c) * v.nv + p] += s;
    }
    __syncthreads();
  }
}

// Jacobians of the cell parameters (19, the adjoint's vector layout) and of the 15 moments with
// respect to the 13 raw moment sums S about the shift (moment_terms, k_moments), per cell by central
// differences in S (step rel_step times the sum's natural scale S_0 T^(degree / 2)); FP64, for the
// adjoint operators. jacP [cell][19][13], jacM [cell][15][13].
__global__ __launch_bounds__(kBlock) void k_adj_param_jacobian(View<double> v, const double* __restrict__ shift,
                                                               const double* __restrict__ volume,
                                                               const double* __restrict__ vdt, double mu0, double omega,
                                                               double Pr, double rel_step, double* __restrict__ jacP,
                                                               double* __restrict__ jacM) {
  __shared__ double lds[kWaves * kDeqSums];
  for (int c = blockIdx.x; c < v.n_cells; c += gridDim.x) {
    const double u0x = shift[c * 3 + 0], u0y = shift[c * 3 + 1], u0z = shift[c * 3 + 2];
    double S[kMomentSums];
    for (int k = 0; k < kMomentSums; ++k) S[k] = 0.0;
    const double* fc = v.f + static_cast<i64>(c) * v.nv;
    for (int p = threadIdx.x; p < v.nv; p += blockDim.x) {
      double t[kMomentSums];
      double fp = fc[p];
      if (v.wqd != nullptr) fp *= v.wqd[p];
      moment_terms<double>(fp, v.xi_x[p] - u0x, v.xi_y[p] - u0y, v.xi_z[p] - u0z, t);
      for (int k = 0; k < kMomentSums; ++k) S[k] += t[k];
    }
    block_sum<kMomentSums>(S, lds);
    double m0[kMoments];
    finalize_moments(S, v.weight_d, u0x, u0y, u0z, m0);
    const double vt = ::sqrt(m0[4] > 0.0 ? m0[4] : 1.0), s0 = S[0] != 0.0 ? ::fabs(S[0]) : 1.0;
    for (int k = 0; k < kMomentSums; ++k) {
      const int deg = k == 0 ? 0 : (k < 4 ? 1 : (k < 10 ? 2 : 3));
      double h = rel_step * s0;
      for (int d = 0; d < deg; ++d) h *= vt;
      double Pd[kAdjParams], md[kMoments];
      for (int sgn = 0; sgn < 2; ++sgn) {
        double Sk[kMomentSums], m[kMoments], P[kAdjParams];
        for (int i = 0; i < kMomentSums; ++i) Sk[i] = S[i];
        Sk[k] += sgn == 0 ? h : -h;
        finalize_moments(Sk, v.weight_d, u0x, u0y, u0z, m);
        int r = 0;
        const CellP<double> cp = cell_params_wave(v, m, volume[c], vdt[c], mu0, omega, Pr, r);
        params_to_vector(cp, P);
        for (int i = 0; i < kAdjParams; ++i) Pd[i] = sgn == 0 ? P[i] : (Pd[i] - P[i]) / (2.0 * h);
        for (int i = 0; i < kMoments; ++i) md[i] = sgn == 0 ? m[i] : (md[i] - m[i]) / (2.0 * h);
      }
      if (threadIdx.x == 0) {
        for (int i = 0; i < kAdjParams; ++i) jacP[(static_cast<i64>(c) * kAdjParams + i) * kMomentSums + k] = Pd[i];
        for (int i = 0; i < kMoments; ++i) jacM[(static_cast<i64>(c) * kMoments + i) * kMomentSums + k] = md[i];
      }
    }
  }
}

}  // namespace gsk::dev
