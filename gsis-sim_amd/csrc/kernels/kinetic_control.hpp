#pragma once
// The per-iteration control of the CIS step on the device: everything the prepare of one CIS
// iteration computes between the moment sums and the sweeps, which the reference runs on the host
// (src/kinetic/kinetic_solver.cpp):
//   positivity check, mass, closed-domain rescaling        :365-404 (prepare)
//   eq. (59) convergence measure                            :338-363 (compute_epsilon)
//   volume means of q and sigma reported per iteration      (prepare's statistics)
//   wall densities from the cells (first prepare)           :156-164
//   pressure-face states: limited reconstruction and        :172-260 (update_pressure_inflow)
//     characteristic condition
//   boundary sums -> fluxes, face states, wall densities,   :280-335 (process_boundary_sums)
//     patch mass flows
// FP64 throughout, each formula in the operation order of this library's host orchestration
// (csrc/host/solver.cpp, csrc/host/boundary.cpp), which the CPU oracle drives. Sums over cells and
// faces run in a fixed order (contiguous cell ranges per block, a fixed tree within the block, blocks
// in order), so the results are deterministic and differ from the host's sequential sums by rounding
// only.

#include "kinetic_view.hpp"

namespace gsk::dev {

struct CtrlFace {  // per boundary face
  int owner, patch;
  double area;     // face area / L_ref^2
  double nx, ny, nz;
};
struct CtrlPatch {  // per patch, uploaded at every prepare (mass-flow inlets move p, sensitivity solves alpha)
  int kind, inlet, zero_tangential, first, n;
  double p, T, accommodation;
};
struct CtrlPressureFace {  // a pressure face and its owner's reconstruction stencil (slots [slot0, slot0 + n_slots))
  int b, owner, slot0, n_slots;
  double eps2;             // K^3 V of the owner
  double tx, ty, tz;       // face centre - owner centre
};
struct CtrlSlot {  // one CSR face of the owner, in CSR order: a neighbour cell, or a boundary face
  int cell, bface, periodic, pad;
  double lx, ly, lz;  // LSQ coefficients
  double ox, oy, oz;  // that face's centre - owner centre
};

// slots of the control record the host reads back once per prepare
enum : int {
  kCtlBad,            // cells with a non-positive or non-finite density or temperature
  kCtlFirstBad,       // the first such cell (internal numbering), its density and temperature
  kCtlBadRho,
  kCtlBadT,
  kCtlMass,           // sum rho V before rescaling
  kCtlScale,          // closed-domain mass scale (1 otherwise)
  kCtlTarget,         // closed-domain target mass, and whether it is set
  kCtlHasTarget,
  kCtlEpsRho,
  kCtlEpsU,
  kCtlEpsT,
  kCtlQ,              // volume means of q (3) and sigma (6)
  kCtlSigma = kCtlQ + 3,
  kCtlMaxwellNegative = kCtlSigma + 6,  // a Maxwell-wall diffuse density below -1e-8
  kCtlMaxwellNoIncoming,                // a Maxwell wall with no incoming velocity points
  kCtlSize
};
constexpr int kCtlBlock = 256;
constexpr int kCtlMassParts = 3;    // bad count, first bad (as a double), mass
constexpr int kCtlStatParts = 16;   // change (3), size (3), volume, V q (3), V sigma (6)

// contiguous cell range of block blk out of nblk
__device__ __forceinline__ void ctrl_range(int n, int blk, int nblk, int& c0, int& c1) {
  const long long per = (static_cast<long long>(n) + nblk - 1) / nblk;
  c0 = static_cast<int>(per * blk < n ? per * blk : n);
  c1 = static_cast<int>(per * (blk + 1) < n ? per * (blk + 1) : n);
}

// fixed-order block sum of N values (thread order within waves by a shuffle tree, waves in order)
template <int N>
__device__ __forceinline__ void ctrl_block_sum(double (&v)[N], double* lds) {
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
    for (int i = 0; i < kCtlBlock / 64; ++i) s += lds[i * N + k];
    v[k] = s;
  }
  __syncthreads();
}

// positivity and mass per block: partial [blk][3] = {bad count, first bad cell (1e300: none), sum rho V}
__global__ __launch_bounds__(kCtlBlock) void k_ctrl_mass(const double* __restrict__ moments,
                                                         const double* __restrict__ volume, int nc,
                                                         double* __restrict__ partial) {
  __shared__ double lds[(kCtlBlock / 64) * 2];
  int c0, c1;
  ctrl_range(nc, blockIdx.x, gridDim.x, c0, c1);
  double s[2] = {0.0, 0.0};  // bad count, mass
  double first = 1e300;
  for (int c = c0 + threadIdx.x; c < c1; c += kCtlBlock) {
    const double* m = moments + static_cast<i64>(c) * kMoments;
    const bool ok = m[0] > 0.0 && m[4] > 0.0 && isfinite(m[0]) && isfinite(m[4]);
    if (!ok) {
      s[0] += 1.0;
      first = first < c ? first : static_cast<double>(c);
    }
    s[1] += m[0] * volume[c];
  }
  ctrl_block_sum<2>(s, lds);
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(first, off, 64);
    first = o < first ? o : first;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = first;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kCtlBlock / 64; ++i) first = lds[i] < first ? lds[i] : first;
    double* o = partial + static_cast<i64>(blockIdx.x) * kCtlMassParts;
    o[0] = s[0];
    o[1] = first;
    o[2] = s[1];
  }
}

// the blocks' partials in order -> bad count, first bad cell, mass, and the closed-domain scale
// target / mass (the target is the mass of the first prepare after initialisation)
__global__ void k_ctrl_mass_reduce(const double* __restrict__ partial, int nblk, const double* __restrict__ moments,
                                   int closed, double* __restrict__ ctrl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double bad = 0.0, first = 1e300, mass = 0.0;
  for (int i = 0; i < nblk; ++i) {
    const double* p = partial + static_cast<i64>(i) * kCtlMassParts;
    bad += p[0];
    first = p[1] < first ? p[1] : first;
    mass += p[2];
  }
  ctrl[kCtlBad] = bad;
  ctrl[kCtlFirstBad] = bad > 0.0 ? first : -1.0;
  if (bad > 0.0) {
    const double* m = moments + static_cast<i64>(first) * kMoments;
    ctrl[kCtlBadRho] = m[0];
    ctrl[kCtlBadT] = m[4];
  }
  ctrl[kCtlMass] = mass;
  double scale = 1.0;
  if (closed) {
    if (ctrl[kCtlHasTarget] == 0.0) {
      ctrl[kCtlTarget] = mass;
      ctrl[kCtlHasTarget] = 1.0;
    }
    scale = ctrl[kCtlTarget] / mass;
  }
  ctrl[kCtlScale] = scale;
}

// closed-domain rescaling with the scale of the control record: f and the density-proportional
// moments (rho and p, stress, heat flux)
template <class F>
__global__ void k_ctrl_scale_f(F* __restrict__ f, i64 n, const double* __restrict__ ctrl) {
  const F s = static_cast<F>(ctrl[kCtlScale]);
  for (i64 i = blockIdx.x * static_cast<i64>(blockDim.x) + threadIdx.x; i < n; i += static_cast<i64>(gridDim.x) * blockDim.x)
    f[i] *= s;
}
__global__ void k_ctrl_scale_moments(double* __restrict__ moments, int nc, const double* __restrict__ ctrl) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const double s = ctrl[kCtlScale];
  double* m = moments + static_cast<i64>(c) * kMoments;
  m[0] *= s;
  for (int k = 5; k < kMoments; ++k) m[k] *= s;
}

// eq. (59) parts and the volume sums of q and sigma per block (partial [blk][16]): the change of
// (rho, u, T) since the previous iteration and its size, volume-weighted; prev may be null (first
// iteration: no epsilon)
__global__ __launch_bounds__(kCtlBlock) void k_ctrl_stats(const double* __restrict__ moments,
                                                          const double* __restrict__ prev,
                                                          const double* __restrict__ volume, int nc,
                                                          double* __restrict__ partial) {
  __shared__ double lds[(kCtlBlock / 64) * kCtlStatParts];
  int c0, c1;
  ctrl_range(nc, blockIdx.x, gridDim.x, c0, c1);
  double s[kCtlStatParts];
#pragma unroll
  for (int k = 0; k < kCtlStatParts; ++k) s[k] = 0.0;
  for (int c = c0 + threadIdx.x; c < c1; c += kCtlBlock) {
    const double V = volume[c];
    const double* now = moments + static_cast<i64>(c) * kMoments;
    if (prev != nullptr) {
      const double* was = prev + static_cast<i64>(c) * kMoments;
      const double dr = now[0] - was[0], dT = now[4] - was[4];
      const double du[3] = {now[1] - was[1], now[2] - was[2], now[3] - was[3]};
      s[0] += dr * dr * V;
      s[3] += was[0] * was[0] * V;
      s[1] += (du[0] * du[0] + du[1] * du[1] + du[2] * du[2]) * V;
      s[4] += (was[1] * was[1] + was[2] * was[2] + was[3] * was[3]) * V;
      s[2] += dT * dT * V;
      s[5] += was[4] * was[4] * V;
    }
    s[6] += V;
#pragma unroll
    for (int k = 0; k < 3; ++k) s[7 + k] += V * now[12 + k];
#pragma unroll
    for (int k = 0; k < 6; ++k) s[10 + k] += V * now[6 + k];
  }
  ctrl_block_sum<kCtlStatParts>(s, lds);
  if (threadIdx.x == 0) {
    double* o = partial + static_cast<i64>(blockIdx.x) * kCtlStatParts;
#pragma unroll
    for (int k = 0; k < kCtlStatParts; ++k) o[k] = s[k];
  }
}

// epsilon = relative L2 changes of (rho, u, T) (the velocity size floored at u_floor^2 times the
// volume), and the volume means of q and sigma
__global__ void k_ctrl_stats_reduce(const double* __restrict__ partial, int nblk, int has_prev, double u_floor,
                                    double* __restrict__ ctrl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s[kCtlStatParts];
  for (int k = 0; k < kCtlStatParts; ++k) s[k] = 0.0;
  for (int i = 0; i < nblk; ++i)
    for (int k = 0; k < kCtlStatParts; ++k) s[k] += partial[static_cast<i64>(i) * kCtlStatParts + k];
  const double vol = s[6];
  if (has_prev) {
    const double size_u = s[4] < u_floor * u_floor * vol ? u_floor * u_floor * vol : s[4];
    const double size[3] = {s[3], size_u, s[5]};
    for (int k = 0; k < 3; ++k) {
      const double d = s[k], n = size[k];
      ctrl[kCtlEpsRho + k] = n > 0.0 ? ::sqrt(d / n) : (d > 0.0 ? 1.0 : 0.0);
    }
  }
  for (int k = 0; k < 3; ++k) ctrl[kCtlQ + k] = s[7 + k] / vol;
  for (int k = 0; k < 6; ++k) ctrl[kCtlSigma + k] = s[10 + k] / vol;
}

// First prepare: wall faces take the density alpha rho_c sqrt(T_c / T_wall) of their owner cell.
__global__ void k_ctrl_seed_walls(const CtrlFace* __restrict__ faces, const CtrlPatch* __restrict__ patches, int nb,
                                  const double* __restrict__ moments, double* __restrict__ state) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const CtrlPatch& pc = patches[faces[b].patch];
  if (pc.kind != kBcDiffuse && pc.kind != kBcMaxwell) return;
  const double* m = moments + static_cast<i64>(faces[b].owner) * kMoments;
  state[static_cast<i64>(b) * 6] = pc.accommodation * m[0] * ::sqrt(m[4] / pc.T);
}

// New state of each pressure face: the owner's (p, u, T) plus the limited LSQ gradient along the
// face offset (neighbours: the adjacent cells; across boundary faces the previous round's face
// state when it has a positive density and temperature, else the owner's own value; Venkatakrishnan
// with the absolute eps^2 = K^3 V, periodic faces left out of the limiter; a reconstruction that is
// not positive, or cell_mode, keeps the owner's value), then the patch's pressure (and at inlets
// temperature) with the acoustic characteristic du = (p_i - p_b) / (rho_i c_s), c_s = sqrt(gamma
// T_i); inlets with zero tangential velocity keep only the normal component. state: {scale, rho,
// ux, uy, uz, T} per boundary face.
__global__ void k_ctrl_pressure(const CtrlPressureFace* __restrict__ pf, int n_pf, const CtrlSlot* __restrict__ slots,
                                const CtrlFace* __restrict__ faces, const CtrlPatch* __restrict__ patches,
                                const double* __restrict__ moments, const double* __restrict__ face_state,
                                int has_face_state, int limiter, int cell_mode, double gamma,
                                double* __restrict__ state) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pf) return;
  const CtrlPressureFace& P = pf[i];
  const double* mo = moments + static_cast<i64>(P.owner) * kMoments;
  const double own[5] = {mo[5], mo[1], mo[2], mo[3], mo[4]};
  double w[5] = {own[0], own[1], own[2], own[3], own[4]};
  if (!cell_mode) {
    double out[5];
    for (int q = 0; q < 5; ++q) {
      double gx = 0.0, gy = 0.0, gz = 0.0, hi = own[q], lo = own[q];
      for (int j = 0; j < P.n_slots; ++j) {
        const CtrlSlot& sl = slots[P.slot0 + j];
        double nb;
        if (sl.cell >= 0) {
          const double* m = moments + static_cast<i64>(sl.cell) * kMoments;
          const double cs[5] = {m[5], m[1], m[2], m[3], m[4]};
          nb = cs[q];
        } else {
          const double* fs = face_state + static_cast<i64>(sl.bface) * 5;
          const bool usable = has_face_state && !sl.periodic && fs[0] > 0.0 && fs[4] > 0.0;
          const double fsw[5] = {fs[0] * fs[4], fs[1], fs[2], fs[3], fs[4]};
          nb = usable ? fsw[q] : own[q];
        }
        const double d = nb - own[q];
        gx += d * sl.lx;
        gy += d * sl.ly;
        gz += d * sl.lz;
        if (hi < nb) hi = nb;
        if (nb < lo) lo = nb;
      }
      double phi = 1.0;
      if (limiter) {
        for (int j = 0; j < P.n_slots; ++j) {
          const CtrlSlot& sl = slots[P.slot0 + j];
          if (sl.periodic) continue;
          const double d2 = gx * sl.ox + gy * sl.oy + gz * sl.oz;
          unsigned clamp = 0;
          const double v = venkat_tracked<double>(d2, hi - own[q], lo - own[q], P.eps2, clamp);
          if (v < phi) phi = v;
        }
      }
      out[q] = own[q] + phi * (gx * P.tx + gy * P.ty + gz * P.tz);
    }
    if (out[0] > 0.0 && out[4] > 0.0)
      for (int q = 0; q < 5; ++q) w[q] = out[q];
  }
  const CtrlFace& fc = faces[P.b];
  const CtrlPatch& pc = patches[fc.patch];
  const double p_i = w[0], T_i = w[4];
  const double rho_i = p_i / T_i;
  const double c_s = ::sqrt(gamma * T_i);
  const double p_b = pc.p, T_b = pc.inlet ? pc.T : T_i;
  const double du = (p_i - p_b) / (rho_i * c_s);
  double u[3];
  if (pc.zero_tangential) {
    const double un = w[1] * fc.nx + w[2] * fc.ny + w[3] * fc.nz + du;
    u[0] = un * fc.nx;
    u[1] = un * fc.ny;
    u[2] = un * fc.nz;
  } else {
    u[0] = w[1] + du * fc.nx;
    u[1] = w[2] + du * fc.ny;
    u[2] = w[3] + du * fc.nz;
  }
  double* st = state + static_cast<i64>(P.b) * 6;
  st[0] = 1.0;
  st[1] = p_b / T_b;
  st[2] = u[0];
  st[3] = u[1];
  st[4] = u[2];
  st[5] = T_b;
}

// inflow table of the kernels from the FP64 states: scale x Maxwellian(rho, u, T) in F, and the
// FP64 scales the inner-sweep wall refresh updates
template <class F>
__global__ void k_ctrl_inflow_table(const double* __restrict__ state, int nb, Inflow<F>* __restrict__ inflow,
                                    double* __restrict__ scale) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double* st = state + static_cast<i64>(b) * 6;
  Inflow<F> in;
  in.m = maxw_from_state<F>(st[1], st[2], st[3], st[4], st[5]);
  in.scale = static_cast<F>(st[0]);
  inflow[b] = in;
  scale[b] = st[0];
}

// One round's sums (20 per face: outgoing / incoming flux, then outgoing / incoming moments) ->
// the face's fluxes F+ and F-, its state (rho, u, T) from the moments, its mass flow, and the wall
// densities they imply: diffuse walls rho = -j+ / j- of the unit-density wall Maxwellian (relaxed by
// relaxation), Maxwell walls the density plus the mass imbalance over the unit-density incoming
// flux, vacuum faces nothing incoming. wall_unit: [face][10] unit-density half-range sums of Maxwell
// walls.
__global__ void k_ctrl_absorb(const double* __restrict__ sums, const CtrlFace* __restrict__ faces,
                              const CtrlPatch* __restrict__ patches, const double* __restrict__ wall_unit, int nb,
                              double relaxation, double* __restrict__ state, double* __restrict__ flux_out,
                              double* __restrict__ flux_in, double* __restrict__ face_state,
                              double* __restrict__ face_flow, double* __restrict__ ctrl) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double* sm = sums + static_cast<i64>(b) * kBoundarySums;
  const CtrlFace& fc = faces[b];
  const CtrlPatch& pc = patches[fc.patch];
  double* st = state + static_cast<i64>(b) * 6;
  const double area = fc.area;
  double in_scale = 1.0, delta = 0.0;
  const bool maxwell = pc.kind == kBcMaxwell;
  const double* unit = maxwell ? wall_unit + static_cast<i64>(b) * 10 : nullptr;
  if (pc.kind == kBcDiffuse) {
    const double jm = sm[5];
    const double rho = jm < 0.0 ? dmax(0.0, -sm[0] / jm) : 0.0;
    st[0] = relaxation == 1.0 ? rho : relaxation * rho + (1.0 - relaxation) * st[0];
    in_scale = rho;
  } else if (pc.kind == kBcVacuum) {
    in_scale = 0.0;
  } else if (maxwell) {
    if (!(unit[0] < 0.0)) ctrl[kCtlMaxwellNoIncoming] = 1.0;
    delta = -(sm[0] + sm[5]) / unit[0];
    const double next = st[0] + delta;
    if (next < -1e-8) ctrl[kCtlMaxwellNegative] = 1.0;
    st[0] = dmax(0.0, next);
  }
  double mom[5];
  for (int k = 0; k < 5; ++k) {
    flux_out[static_cast<i64>(b) * 5 + k] = area * sm[k];
    if (maxwell) {
      flux_in[static_cast<i64>(b) * 5 + k] = area * (sm[5 + k] + delta * unit[k]);
      mom[k] = sm[10 + k] + sm[15 + k] + delta * unit[5 + k];
    } else {
      flux_in[static_cast<i64>(b) * 5 + k] = area * in_scale * sm[5 + k];
      mom[k] = sm[10 + k] + in_scale * sm[15 + k];
    }
  }
  double* fs = face_state + static_cast<i64>(b) * 5;
  for (int k = 0; k < 5; ++k) fs[k] = 0.0;
  if (mom[0] > 0.0) {
    const double ux = mom[1] / mom[0], uy = mom[2] / mom[0], uz = mom[3] / mom[0];
    fs[0] = mom[0];
    fs[1] = ux;
    fs[2] = uy;
    fs[3] = uz;
    fs[4] = (2.0 * mom[4] / mom[0] - (ux * ux + uy * uy + uz * uz)) / 3.0;
  }
  face_flow[b] = flux_out[static_cast<i64>(b) * 5] + flux_in[static_cast<i64>(b) * 5];
}

// patch mass flows: one block per patch, its faces strided over the threads, then a fixed-order
// block sum
__global__ __launch_bounds__(kCtlBlock) void k_ctrl_patch_flow(const double* __restrict__ face_flow,
                                                               const CtrlPatch* __restrict__ patches, int n_patches,
                                                               double* __restrict__ flow) {
  __shared__ double lds[kCtlBlock / 64];
  const int pi = blockIdx.x;
  if (pi >= n_patches) return;
  double s[1] = {0.0};
  const int b0 = patches[pi].first, b1 = patches[pi].first + patches[pi].n;
  for (int b = b0 + static_cast<int>(threadIdx.x); b < b1; b += kCtlBlock) s[0] += face_flow[b];
  ctrl_block_sum<1>(s, lds);
  if (threadIdx.x == 0) flow[pi] = s[0];
}

}  // namespace gsk::dev
