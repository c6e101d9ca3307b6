#pragma once
// Device-side tables and the per-(cell, velocity) operators that read neighbours: LSQ gradient
// with limiter, boundary values, reconstruction at faces, Maxwell-wall reflection, residual and
// the LU-SGS cell updates. Reference: src/kinetic/kinetic_kernels.hpp:60-372 (same formulas and
// operation order; see kinetic_math.hpp for the point-wise ones).

#include "kinetic_math.hpp"

namespace gsk::dev {

using i64 = long long;

template <class F>
struct CsrE {
  int other, bface;
  F sx, sy, sz;
  F dsx, dsy, dsz;
  F dox, doy, doz;
  F cx, cy, cz;
};

template <class F>
struct BFace {
  int owner, kind, mirror_xor, patch;
  F nx, ny, nz, area, dx, dy, dz;
};

template <class F>
struct Inflow {
  Maxw<F> m;
  F scale;
};

// All device pointers and scalars the kernels read; passed by value as a kernel argument.
template <class F>
struct View {
  int n_cells, nv, n_bfaces;
  const F* xi_x;
  const F* xi_y;
  const F* xi_z;
  F weight;
  double weight_d;
  const int* csr_off;
  const CsrE<F>* csr;
  const CsrE<F>* ell;  // [cell][kGather] copy of the CSR entries, padded with a no-op periodic face
  const F* eps2;
  int limiter_mode, positivity;
  F diag_scale, interface_damping;
  const BFace<F>* bf;
  const Inflow<F>* inflow;
  const CellP<F>* cells;
  F* f;
  const i64* mw_off;   // per boundary face, -1 when not a Maxwell wall
  const F* mw_frac;    // specular fraction per boundary face
  F* specular;         // [n_maxwell][nv] reflected reconstructed values
  F* specular_first;   // [n_maxwell][nv] reflected cell values
  const int* grid_to_storage;
  int grid_n[3];
  int n_sweep_dirs;
  const int* sweep_order;
  const int* sweep_pos;
  const unsigned char* sweep_dir_of;
  DeqGrid deq;
  int discrete_eq, deq_fallback;
  double cold_bgk;
};

// velocity chunk workspace [n_cells][stride], position t = p - k0
template <class F>
struct Chunk {
  int k0, kc, stride;
  F* gx;
  F* gy;
  F* gz;
};

template <class F>
__device__ __forceinline__ F f_at(const View<F>& v, int c, int p) {
  return v.f[static_cast<i64>(c) * v.nv + p];
}

template <class F>
__device__ __forceinline__ F inflow_value(const Inflow<F>& b, F x, F y, F z) {
  return b.scale * maxw_eval(b.m, x, y, z);
}

// first-order boundary value for gradient / limiter (kinetic_kernels.hpp:145-162)
template <class F>
__device__ F boundary_value_first_order(const View<F>& v, const CsrE<F>& e, int cell, int p, F x, F y, F z, F fi) {
  const BFace<F>& b = v.bf[e.bface];
  if (b.kind == kBcPeriodic) return fi;
  if (b.kind == kBcSymmetry) return F(0.5) * (fi + f_at(v, cell, p ^ b.mirror_xor));
  const F a = x * e.sx + y * e.sy + z * e.sz;
  if (a >= F(0)) return fi;
  if (b.kind == kBcMaxwell)
    return inflow_value(v.inflow[e.bface], x, y, z) + v.mw_frac[e.bface] * v.specular_first[v.mw_off[e.bface] + p];
  return inflow_value(v.inflow[e.bface], x, y, z);
}

// limited LSQ gradient phi * grad f at (cell, p) (kinetic_kernels.hpp:165-201).
// The neighbour values are gathered first (all loads in flight together), then combined in CSR order,
// so the arithmetic is the reference's while the wave waits on memory once instead of once per face.
constexpr int kGather = 6;  // faces held in registers (hex 6, prism/pyramid 5, tet 4); more: CSR path

template <class F>
__device__ __forceinline__ void limit_and_scale(const View<F>& v, int cell, const CsrE<F>* ents, int n, F fi, F fmx,
                                                F fmn, F gx, F gy, F gz, F& ogx, F& ogy, F& ogz) {
  if (v.limiter_mode == kLimNone && v.positivity == 0) {
    ogx = gx;
    ogy = gy;
    ogz = gz;
    return;
  }
  if (v.positivity != 0 && !(fi > F(0))) {
    ogx = ogy = ogz = F(0);
    return;
  }
  F eps2 = v.eps2[cell];
  if (v.limiter_mode == kLimRel) {
    const F s = dmax(dabs(fmx), dabs(fmn));
    eps2 *= s * s;
  }
  const F dmx = fmx - fi, dmn = fmn - fi;
  F phi = F(1);
  F proj_min = F(0);
  for (int k = 0; k < n; ++k) {
    const CsrE<F>& e = ents[k];
    if (e.other < 0 && v.bf[e.bface].kind == kBcPeriodic) continue;
    const F proj = gx * e.dsx + gy * e.dsy + gz * e.dsz;
    if (v.limiter_mode != kLimNone) phi = dmin(phi, venkat_face(proj, dmx, dmn, eps2));
    proj_min = dmin(proj_min, proj);
  }
  if (v.positivity != 0 && proj_min < F(0)) phi = fi > F(0) ? dmin(phi, fi / (-proj_min)) : F(0);
  ogx = phi * gx;
  ogy = phi * gy;
  ogz = phi * gz;
}

// gradient from at most kGather entries ents[0..n): neighbour values gathered first, then combined
// in entry order
template <class F>
__device__ __forceinline__ void gradient_gathered(const View<F>& v, int cell, int p, F x, F y, F z, F fi,
                                                  const CsrE<F>* ents, int n, F& ogx, F& ogy, F& ogz) {
  F gx = F(0), gy = F(0), gz = F(0);
  F fmx = fi, fmn = fi;
  F fk[kGather];
#pragma unroll
  for (int k = 0; k < kGather; ++k)
    if (k < n) {
      const int o = ents[k].other;
      fk[k] = o >= 0 ? f_at(v, o, p) : F(0);
    }
#pragma unroll
  for (int k = 0; k < kGather; ++k)
    if (k < n) {
      const CsrE<F>& e = ents[k];
      if (e.other < 0) fk[k] = boundary_value_first_order(v, e, cell, p, x, y, z, fi);
      const F d = fk[k] - fi;
      gx += e.cx * d;
      gy += e.cy * d;
      gz += e.cz * d;
      fmx = dmax(fmx, fk[k]);
      fmn = dmin(fmn, fk[k]);
    }
  limit_and_scale(v, cell, ents, n, fi, fmx, fmn, gx, gy, gz, ogx, ogy, ogz);
}

template <class F>
__device__ void limited_gradient(const View<F>& v, int cell, int p, F x, F y, F z, F fi, F& ogx, F& ogy, F& ogz) {
  const int e0 = v.csr_off[cell], e1 = v.csr_off[cell + 1];
  if (e1 - e0 <= kGather) {
    gradient_gathered(v, cell, p, x, y, z, fi, v.csr + e0, e1 - e0, ogx, ogy, ogz);
    return;
  }
  F gx = F(0), gy = F(0), gz = F(0);
  F fmx = fi, fmn = fi;
  for (int k = e0; k < e1; ++k) {
    const CsrE<F>& e = v.csr[k];
    const F fk = e.other >= 0 ? f_at(v, e.other, p) : boundary_value_first_order(v, e, cell, p, x, y, z, fi);
    const F d = fk - fi;
    gx += e.cx * d;
    gy += e.cy * d;
    gz += e.cz * d;
    fmx = dmax(fmx, fk);
    fmn = dmin(fmn, fk);
  }
  limit_and_scale(v, cell, v.csr + e0, e1 - e0, fi, fmx, fmn, gx, gy, gz, ogx, ogy, ogz);
}

// The same gradient from the cell's padded row of the ELL table: the entry count is the constant
// kGather, so the entry loads issue together with no branch on the face count. The padding entries
// name a periodic face: first-order value fi (d = 0, adds +0 to sums that start at +0 and so never
// hold -0; fmx / fmn unchanged), skipped by the limiter. Cells with more faces get a wrong value
// here and are recomputed by the CSR path afterwards.
template <class F>
__device__ __forceinline__ void limited_gradient_ell(const View<F>& v, int cell, int p, F x, F y, F z, F fi, F& ogx,
                                                     F& ogy, F& ogz) {
  gradient_gathered(v, cell, p, x, y, z, fi, v.ell + static_cast<i64>(cell) * kGather, kGather, ogx, ogy, ogz);
}

// outgoing reconstructed value at a boundary face (kinetic_kernels.hpp:204-211)
template <class F>
__device__ F reconstruct_at_boundary(const View<F>& v, const BFace<F>& b, int p, F x, F y, F z) {
  const F fi = f_at(v, b.owner, p);
  if (b.kind == kBcPeriodic) return fi;
  F gx, gy, gz;
  limited_gradient(v, b.owner, p, x, y, z, fi, gx, gy, gz);
  return fi + gx * b.dx + gy * b.dy + gz * b.dz;
}

// specular reflection about the face normal, trilinear in velocity with a Maxwellian-ratio weight
// (kinetic_kernels.hpp:80-127)
template <bool Reconstruct, class F>
__device__ F reflected_value(const View<F>& v, const BFace<F>& b, int p, F x, F y, F z) {
  if (b.mirror_xor != 0) {
    const int q = p ^ b.mirror_xor;
    if (Reconstruct) return reconstruct_at_boundary(v, b, q, v.xi_x[q], v.xi_y[q], v.xi_z[q]);
    return f_at(v, b.owner, q);
  }
  const F un = x * b.nx + y * b.ny + z * b.nz;
  const F r[3] = {x - F(2) * un * b.nx, y - F(2) * un * b.ny, z - F(2) * un * b.nz};
  int base[3];
  F t[3];
  for (int a = 0; a < 3; ++a) {
    const double zz = (static_cast<double>(r[a]) - v.deq.lo[a]) / v.deq.d[a] - 0.5;
    if (zz < -1e-5 || zz > v.grid_n[a] - 1.0 + 1e-5) return F(0);
    const double cl = dmax(0.0, dmin(zz, v.grid_n[a] - 1.0));
    base[a] = static_cast<int>(cl);
    if (base[a] >= v.grid_n[a] - 1) base[a] = v.grid_n[a] - 2;
    t[a] = static_cast<F>(cl - base[a]);
  }
  const Maxw<F> ref = v.cells[b.owner].s.eq;
  F result = F(0);
  for (int k = 0; k < 8; ++k) {
    const int ix = base[0] + (k & 1), iy = base[1] + ((k >> 1) & 1), iz = base[2] + ((k >> 2) & 1);
    const F w = ((k & 1) ? t[0] : F(1) - t[0]) * ((k & 2) ? t[1] : F(1) - t[1]) * ((k & 4) ? t[2] : F(1) - t[2]);
    if (w == F(0)) continue;
    const int q = v.grid_to_storage[ix + v.grid_n[0] * (iy + v.grid_n[1] * iz)];
    const F qx = v.xi_x[q], qy = v.xi_y[q], qz = v.xi_z[q];
    F value;
    if (Reconstruct)
      value = reconstruct_at_boundary(v, b, q, qx, qy, qz);
    else
      value = f_at(v, b.owner, q);
    const double drx = static_cast<double>(r[0]) - ref.ux, dry = static_cast<double>(r[1]) - ref.uy,
                 drz = static_cast<double>(r[2]) - ref.uz;
    const double dqx = static_cast<double>(qx) - ref.ux, dqy = static_cast<double>(qy) - ref.uy,
                 dqz = static_cast<double>(qz) - ref.uz;
    const double ratio =
        ::exp(-static_cast<double>(ref.B) * (drx * drx + dry * dry + drz * drz - dqx * dqx - dqy * dqy - dqz * dqz));
    result += w * static_cast<F>(static_cast<double>(value) * ratio);
  }
  return result;
}

template <class F>
__device__ __forceinline__ F maxwell_inflow(const View<F>& v, int b, int p, F x, F y, F z) {
  return inflow_value(v.inflow[b], x, y, z) + v.mw_frac[b] * v.specular[v.mw_off[b] + p];
}

// face flux a * f_face over one internal CSR entry (kinetic_kernels.hpp:216-232)
template <class F>
__device__ F internal_face_flux_term(const View<F>& v, int cell, const CsrE<F>& e, int p, F x, F y, F z) {
  const F a = x * e.sx + y * e.sy + z * e.sz;
  F ff;
  if (a >= F(0)) {
    const F fi = f_at(v, cell, p);
    F gx, gy, gz;
    limited_gradient(v, cell, p, x, y, z, fi, gx, gy, gz);
    ff = fi + gx * e.dsx + gy * e.dsy + gz * e.dsz;
  } else {
    const F fj = f_at(v, e.other, p);
    F gx, gy, gz;
    limited_gradient(v, e.other, p, x, y, z, fj, gx, gy, gz);
    ff = fj + gx * e.dox + gy * e.doy + gz * e.doz;
  }
  return a * ff;
}

// Flux sum over at most kGather entries ents[0..n) (kinetic_kernels.hpp:238-272): upwind neighbour
// values (f and its limited gradient) are gathered first, then the fluxes are added in entry order.
template <class F>
__device__ __forceinline__ F flux_gathered(const View<F>& v, const Chunk<F>& ch, int cell, int t, int p, F x, F y, F z,
                                           F fi, F gix, F giy, F giz, const CsrE<F>* ents, int n) {
  F R = F(0);
  F ak[kGather], fu[kGather], gux[kGather], guy[kGather], guz[kGather];
#pragma unroll
  for (int k = 0; k < kGather; ++k)
    if (k < n) {
      const CsrE<F>& e = ents[k];
      ak[k] = x * e.sx + y * e.sy + z * e.sz;
      fu[k] = gux[k] = guy[k] = guz[k] = F(0);
      if (ak[k] < F(0)) {
        int src_cell = -1, src_t = t;
        if (e.other >= 0) {
          src_cell = e.other;
        } else if (v.bf[e.bface].kind == kBcSymmetry) {
          src_cell = cell;
          src_t = t ^ v.bf[e.bface].mirror_xor;
        }
        if (src_cell >= 0) {
          const i64 wj = static_cast<i64>(src_cell) * ch.stride + src_t;
          fu[k] = f_at(v, src_cell, ch.k0 + src_t);
          gux[k] = ch.gx[wj];
          guy[k] = ch.gy[wj];
          guz[k] = ch.gz[wj];
        }
      }
    }
#pragma unroll
  for (int k = 0; k < kGather; ++k)
    if (k < n) {
      const CsrE<F>& e = ents[k];
      int kind = -1;
      if (e.other < 0) {
        kind = v.bf[e.bface].kind;
        if (kind == kBcPeriodic) continue;
      }
      const F a = ak[k];
      F ff;
      if (a >= F(0)) {
        ff = fi + gix * e.dsx + giy * e.dsy + giz * e.dsz;
      } else if (e.other >= 0) {
        ff = fu[k] + gux[k] * e.dox + guy[k] * e.doy + guz[k] * e.doz;
      } else if (kind == kBcSymmetry) {
        ff = fu[k] + gux[k] * e.dsx + guy[k] * e.dsy + guz[k] * e.dsz;
      } else if (kind == kBcMaxwell) {
        ff = maxwell_inflow(v, e.bface, p, x, y, z);
      } else {
        ff = inflow_value(v.inflow[e.bface], x, y, z);
      }
      R += a * ff;
    }
  return R;
}

// steady residual R = sum_e a f_e - coll (f_S - f_i) (kinetic_kernels.hpp:238-278)
template <class F>
__device__ F residual_value(const View<F>& v, const Chunk<F>& ch, int cell, int t) {
  const int p = ch.k0 + t;
  const F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
  const F fi = f_at(v, cell, p);
  const i64 wi = static_cast<i64>(cell) * ch.stride + t;
  const F gix = ch.gx[wi], giy = ch.gy[wi], giz = ch.gz[wi];
  F R = F(0);
  const int e0 = v.csr_off[cell], e1 = v.csr_off[cell + 1];
  if (e1 - e0 <= kGather) {
    R = flux_gathered(v, ch, cell, t, p, x, y, z, fi, gix, giy, giz, v.csr + e0, e1 - e0);
  } else {
    for (int k = e0; k < e1; ++k) {
      const CsrE<F>& e = v.csr[k];
      int kind = -1;
      if (e.other < 0) {
        kind = v.bf[e.bface].kind;
        if (kind == kBcPeriodic) continue;
      }
      const F a = x * e.sx + y * e.sy + z * e.sz;
      F ff;
      if (a >= F(0)) {
        ff = fi + gix * e.dsx + giy * e.dsy + giz * e.dsz;
      } else if (e.other >= 0) {
        const i64 wj = static_cast<i64>(e.other) * ch.stride + t;
        ff = f_at(v, e.other, p) + ch.gx[wj] * e.dox + ch.gy[wj] * e.doy + ch.gz[wj] * e.doz;
      } else if (kind == kBcSymmetry) {
        const int tm = t ^ v.bf[e.bface].mirror_xor;
        const i64 wm = static_cast<i64>(cell) * ch.stride + tm;
        ff = f_at(v, cell, ch.k0 + tm) + ch.gx[wm] * e.dsx + ch.gy[wm] * e.dsy + ch.gz[wm] * e.dsz;
      } else if (kind == kBcMaxwell) {
        ff = maxwell_inflow(v, e.bface, p, x, y, z);
      } else {
        ff = inflow_value(v.inflow[e.bface], x, y, z);
      }
      R += a * ff;
    }
  }
  const CellP<F>& cp = v.cells[cell];
  R -= cp.coll * (target_eval(cp.s, x, y, z) - fi);
  return R;
}

// The same residual from the cell's padded ELL row (see limited_gradient_ell); padding entries are
// periodic faces, which the flux sum skips.
template <class F>
__device__ F residual_value_ell(const View<F>& v, const Chunk<F>& ch, int cell, int t) {
  const int p = ch.k0 + t;
  const F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
  const F fi = f_at(v, cell, p);
  const i64 wi = static_cast<i64>(cell) * ch.stride + t;
  const F gix = ch.gx[wi], giy = ch.gy[wi], giz = ch.gz[wi];
  F R = flux_gathered(v, ch, cell, t, p, x, y, z, fi, gix, giy, giz, v.ell + static_cast<i64>(cell) * kGather, kGather);
  const CellP<F>& cp = v.cells[cell];
  R -= cp.coll * (target_eval(cp.s, x, y, z) - fi);
  return R;
}

}  // namespace gsk::dev
