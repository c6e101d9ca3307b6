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

// One cell's faces for the gradient and residual kernels, padded to kGather faces (ELL form). The
// integer part comes first and the per-face coefficients are stored field by field, so a wave reads
// a whole row in a few wide scalar loads and can issue all neighbour loads together; neighbour
// offsets are precomputed 64-bit element offsets, so no address needs a scalar multiply.
constexpr int kGather = 6;  // faces per row (hex 6, prism/pyramid 5, tet 4); cells with more: CSR path

template <class F>
struct EllRow {
  i64 nbf[kGather];    // neighbour row in f: other * nv (the cell's own row for boundary faces)
  i64 nbw[kGather];    // neighbour row in the chunk workspace: other * chunk stride (own row likewise)
  int bface[kGather];  // boundary face index, -1 for internal faces
  unsigned skip;       // bit k: face k is periodic or padding, a face no operator reads
  unsigned bnd;        // bit k: face k is a boundary face
  F sx[kGather], sy[kGather], sz[kGather];
  F dsx[kGather], dsy[kGather], dsz[kGather];
  F dox[kGather], doy[kGather], doz[kGather];
  F cx[kGather], cy[kGather], cz[kGather];
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

// The tensor velocity grid axis by axis, for sums in product form (deq_product_newton_sums): the
// node coordinates of axes x, y, z back to back (the values the f precision stores) and, on a
// stretched grid, one weight factor per node such that the relative weight of point (i, j, k) is
// w[i] * w[n0 + j] * w[n0 + n1 + k] * wscale (w null and wscale 1 on the uniform grid).
struct AxisGrid {
  const double* node;
  const double* w;
  double wscale;
  int n[3];
};

// All device pointers and scalars the kernels read; passed by value as a kernel argument.
template <class F>
struct View {
  AxisGrid axes;
  int n_cells, nv, n_bfaces;
  const F* xi_x;
  const F* xi_y;
  const F* xi_z;
  F weight;
  double weight_d;
  // stretched velocity grid (an extension; all three null on the reference's uniform grid): the
  // quadrature weight of velocity p is weight * wq[p] (wqd: the same in FP64), and nodes holds the
  // node coordinates of axes x, y, z back to back for the Maxwell-wall interpolation
  const F* wq;
  const double* wqd;
  const double* nodes;
  const int* csr_off;
  const CsrE<F>* csr;
  const EllRow<F>* ell;  // [cell] padded face rows (cells with more than kGather faces: CSR path)
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
  // limiter branches of sensitivity solves (lim_mode 0: off): 1 records, per (cell, velocity), the
  // branch every limited gradient takes; 2 evaluates every limited gradient on its recorded branch
  int lim_mode;
  unsigned* lim_branch;  // [cell][velocity]
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
__device__ F boundary_value_first_order(const View<F>& v, int bface, F sx, F sy, F sz, int cell, int p, F x, F y, F z,
                                        F fi) {
  const BFace<F>& b = v.bf[bface];
  if (b.kind == kBcPeriodic) return fi;
  if (b.kind == kBcSymmetry) return F(0.5) * (fi + f_at(v, cell, p ^ b.mirror_xor));
  const F a = x * sx + y * sy + z * sz;
  if (a >= F(0)) return fi;
  if (b.kind == kBcMaxwell)
    return inflow_value(v.inflow[bface], x, y, z) + v.mw_frac[bface] * v.specular_first[v.mw_off[bface] + p];
  return inflow_value(v.inflow[bface], x, y, z);
}

template <class F>
__device__ __forceinline__ F boundary_value_first_order(const View<F>& v, const CsrE<F>& e, int cell, int p, F x, F y,
                                                        F z, F fi) {
  return boundary_value_first_order(v, e.bface, e.sx, e.sy, e.sz, cell, p, x, y, z, fi);
}

// limited LSQ gradient phi * grad f at (cell, p) (kinetic_kernels.hpp:165-201).
// The neighbour values are gathered first (all loads in flight together), then combined in CSR order,
// so the arithmetic is the reference's while the wave waits on memory once instead of once per face.

// phi * grad from the largest and the most negative face projection grad . ds (over the faces the
// limiter reads). The reference takes phi = min over faces of venkat_face(proj_k). For proj > 0
// (d1 = dmx >= 0) the clamped limiter function is 1 up to proj = dmx / 2 and strictly decreasing
// beyond (its derivative has the sign of (d1^2 + eps2)(d1 - 4 proj) - 4 d1 proj^2 < 0 there), so it
// is non-increasing in proj; the mirror statement holds for proj < 0 with d1 = dmn <= 0. The minimum
// over faces is therefore attained at the two extreme projections: two divisions per (cell,
// velocity) instead of one per face, with the same value up to rounding.
template <class F>
__device__ __forceinline__ void limit_from_extremes(const View<F>& v, int cell, F fi, F fmx, F fmn, F proj_max,
                                                    F proj_min, F gx, F gy, F gz, F& ogx, F& ogy, F& ogz) {
  F eps2 = v.eps2[cell];
  if (v.limiter_mode == kLimRel) {
    const F s = dmax(dabs(fmx), dabs(fmn));
    eps2 *= s * s;
  }
  const F dmx = fmx - fi, dmn = fmn - fi;
  F phi = F(1);
  if (v.limiter_mode != kLimNone) phi = dmin(venkat_face(proj_max, dmx, dmn, eps2), venkat_face(proj_min, dmx, dmn, eps2));
  if (v.positivity != 0 && proj_min < F(0)) phi = fi > F(0) ? dmin(phi, fi / (-proj_min)) : F(0);
  ogx = phi * gx;
  ogy = phi * gy;
  ogz = phi * gz;
}

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
  F proj_max = F(0), proj_min = F(0);
  for (int k = 0; k < n; ++k) {
    const CsrE<F>& e = ents[k];
    if (e.other < 0 && v.bf[e.bface].kind == kBcPeriodic) continue;
    const F proj = gx * e.dsx + gy * e.dsy + gz * e.dsz;
    proj_max = dmax(proj_max, proj);
    proj_min = dmin(proj_min, proj);
  }
  limit_from_extremes(v, cell, fi, fmx, fmn, proj_max, proj_min, gx, gy, gz, ogx, ogy, ogz);
}

// ---- Limiter branches (sensitivity solves) ----
// The limited gradient is piecewise smooth in f: the limiter takes the largest and the most negative
// face projection, the largest and smallest face value, the larger of |f_max| and |f_min| (relative
// eps), the smaller of two limiter values, and clamps each to [0, 1]. A sensitivity solve needs
// R(f) differentiated on the branch the converged state sits on; crossing a branch inside a
// difference product makes it nonlinear in the direction. So at the converged state every choice is
// recorded (Mode 1), and later evaluations (Mode 2) take the recorded candidates and clamps, which
// makes R smooth around that state; differences of it then give the branch's Jacobian products to
// second order in the step. Record layout (one 32-bit word per (cell, velocity)):
//   [0,4) face of the largest projection   [4,8) face of the most negative one   (kNoCandidate: stayed 0)
//   [8,12) face of f_max   [12,16) face of f_min   (kNoCandidate: the cell's own value)
//   16 eps from |f_max| (else |f_min|)   [17,19) max-side clamp   [19,21) min-side clamp
//   (0 none, 1 to 0, 2 to 1)   21 phi from the min side   [22,24) positivity (0 none, 1 phi =
//   f_i / -proj_min, 2 no gradient)
constexpr unsigned kNoCandidate = 15;
constexpr int kLimMaxFaces = 15;

// Element k of a register array, with k a run-time value: an unrolled select, so the array stays in
// registers (an indexed access would place it in scratch memory).
template <int N, class F>
__device__ __forceinline__ F pick(const F (&a)[N], unsigned k) {
  F r = a[0];
#pragma unroll
  for (int i = 1; i < N; ++i)
    if (k == static_cast<unsigned>(i)) r = a[i];
  return r;
}

// phi * grad for one (cell, velocity) from its N face slots (fk the face values; faces whose bit is
// set in `use` count for the extremes) and face offsets ds, recording (Mode 1) or following (Mode 2)
// the branch.
template <int Mode, int N, class F>
__device__ void limit_on_branch(const View<F>& v, int cell, int p, F fi, const F (&fk)[N], unsigned use,
                                const F (&dsx)[N], const F (&dsy)[N], const F (&dsz)[N], F gx, F gy, F gz, F& ogx,
                                F& ogy, F& ogz) {
  static_assert(N <= kLimMaxFaces, "a branch record holds at most kLimMaxFaces faces");
  unsigned* slot = v.lim_branch + static_cast<i64>(cell) * v.nv + p;
  if (v.limiter_mode == kLimNone && v.positivity == 0) {
    ogx = gx;
    ogy = gy;
    ogz = gz;
    return;
  }
  const unsigned code = Mode == 2 ? *slot : 0u;
  if (v.positivity != 0 && (Mode == 2 ? ((code >> 22) & 3u) == 2u : !(fi > F(0)))) {
    ogx = ogy = ogz = F(0);
    if (Mode == 1) *slot = 2u << 22;
    return;
  }
  unsigned kmx = kNoCandidate, kmn = kNoCandidate, imx = kNoCandidate, imn = kNoCandidate;
  F fmx = fi, fmn = fi, proj_max = F(0), proj_min = F(0);
  if (Mode == 1) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if (!((use >> k) & 1u)) continue;
      if (fk[k] > fmx) {
        fmx = fk[k];
        imx = static_cast<unsigned>(k);
      }
      if (fk[k] < fmn) {
        fmn = fk[k];
        imn = static_cast<unsigned>(k);
      }
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if (!((use >> k) & 1u)) continue;
      const F proj = dot3(gx, gy, gz, dsx[k], dsy[k], dsz[k]);
      if (proj > proj_max) {
        proj_max = proj;
        kmx = static_cast<unsigned>(k);
      }
      if (proj < proj_min) {
        proj_min = proj;
        kmn = static_cast<unsigned>(k);
      }
    }
  } else {
    kmx = code & 15u;
    kmn = (code >> 4) & 15u;
    imx = (code >> 8) & 15u;
    imn = (code >> 12) & 15u;
    if (imx != kNoCandidate) fmx = pick(fk, imx);
    if (imn != kNoCandidate) fmn = pick(fk, imn);
    if (kmx != kNoCandidate) proj_max = dot3(gx, gy, gz, pick(dsx, kmx), pick(dsy, kmx), pick(dsz, kmx));
    if (kmn != kNoCandidate) proj_min = dot3(gx, gy, gz, pick(dsx, kmn), pick(dsy, kmn), pick(dsz, kmn));
  }
  F eps2 = v.eps2[cell];
  bool eps_from_max = false;
  if (v.limiter_mode == kLimRel) {
    eps_from_max = Mode == 2 ? ((code >> 16) & 1u) != 0u : !(dabs(fmx) < dabs(fmn));
    const F s = eps_from_max ? dabs(fmx) : dabs(fmn);
    eps2 *= s * s;
  }
  const F dmx = fmx - fi, dmn = fmn - fi;
  F phi = F(1);
  unsigned cmx = 2, cmn = 2, side = 0, pos = 0;
  if (v.limiter_mode != kLimNone) {
    F vmx, vmn;
    if (Mode == 1) {
      vmx = venkat_tracked(proj_max, dmx, dmn, eps2, cmx);
      vmn = venkat_tracked(proj_min, dmx, dmn, eps2, cmn);
      side = vmn < vmx ? 1u : 0u;
    } else {
      cmx = (code >> 17) & 3u;
      cmn = (code >> 19) & 3u;
      side = (code >> 21) & 1u;
      vmx = kmx == kNoCandidate ? F(1) : venkat_frozen(proj_max, dmx, eps2, cmx);
      vmn = kmn == kNoCandidate ? F(1) : venkat_frozen(proj_min, dmn, eps2, cmn);
    }
    phi = side ? vmn : vmx;
  }
  if (v.positivity != 0) {
    if (Mode == 1) {
      if (proj_min < F(0)) {
        const F cap = fi / (-proj_min);
        if (cap < phi) {
          phi = cap;
          pos = 1;
        }
      }
    } else if (((code >> 22) & 3u) == 1u) {
      phi = fi / (-proj_min);
    }
  }
  if (Mode == 1)
    *slot = kmx | (kmn << 4) | (imx << 8) | (imn << 12) | ((eps_from_max ? 1u : 0u) << 16) | (cmx << 17) |
            (cmn << 19) | (side << 21) | (pos << 22);
  ogx = phi * gx;
  ogy = phi * gy;
  ogz = phi * gz;
}

// The limited gradient from the CSR entries (cells of at most kGather faces in registers, larger
// ones up to kLimMaxFaces), recording or following the branch.
template <int Mode, int N, class F>
__device__ __forceinline__ void limited_gradient_on_branch_n(const View<F>& v, int cell, int p, F x, F y, F z, F fi,
                                                             F& ogx, F& ogy, F& ogz) {
  const int e0 = v.csr_off[cell], n = v.csr_off[cell + 1] - e0;
  F fk[N], dsx[N], dsy[N], dsz[N];
  unsigned use = 0;
  F gx = F(0), gy = F(0), gz = F(0);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    fk[k] = dsx[k] = dsy[k] = dsz[k] = F(0);
    if (k >= n) continue;
    const CsrE<F>& e = v.csr[e0 + k];
    fk[k] = e.other >= 0 ? f_at(v, e.other, p) : boundary_value_first_order(v, e, cell, p, x, y, z, fi);
    const F d = fk[k] - fi;
    gx += e.cx * d;
    gy += e.cy * d;
    gz += e.cz * d;
    dsx[k] = e.dsx;
    dsy[k] = e.dsy;
    dsz[k] = e.dsz;
    if (!(e.other < 0 && v.bf[e.bface].kind == kBcPeriodic)) use |= 1u << k;
  }
  limit_on_branch<Mode>(v, cell, p, fi, fk, use, dsx, dsy, dsz, gx, gy, gz, ogx, ogy, ogz);
}

// cells with more than kGather faces (none on hexahedral, prism or tetrahedral meshes): out of line,
// so that their wider arrays do not weigh on the common path's registers
template <int Mode, class F>
__device__ __attribute__((noinline)) void limited_gradient_on_branch_wide(const View<F>& v, int cell, int p, F x, F y,
                                                                           F z, F fi, F& ogx, F& ogy, F& ogz) {
  limited_gradient_on_branch_n<Mode, kLimMaxFaces>(v, cell, p, x, y, z, fi, ogx, ogy, ogz);
}

template <int Mode, class F>
__device__ void limited_gradient_on_branch(const View<F>& v, int cell, int p, F x, F y, F z, F fi, F& ogx, F& ogy,
                                           F& ogz) {
  if (v.csr_off[cell + 1] - v.csr_off[cell] <= kGather)
    limited_gradient_on_branch_n<Mode, kGather>(v, cell, p, x, y, z, fi, ogx, ogy, ogz);
  else
    limited_gradient_on_branch_wide<Mode>(v, cell, p, x, y, z, fi, ogx, ogy, ogz);
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

// Mode: limiter branches, a compile-time choice so that the normal kernels (Mode 0) carry none of
// the branch code: 1 records the branch, 2 follows the record (limited_gradient_on_branch).
template <int Mode = 0, class F>
__device__ void limited_gradient(const View<F>& v, int cell, int p, F x, F y, F z, F fi, F& ogx, F& ogy, F& ogz) {
  if constexpr (Mode != 0) {
    limited_gradient_on_branch<Mode>(v, cell, p, x, y, z, fi, ogx, ogy, ogz);
    return;
  }
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

// The same gradient from the cell's padded row. Periodic and padding faces are skipped: the
// reference's first-order value there is fi, so the face adds e.c * 0 = +-0 to sums that start at +0
// (and so never hold -0, leaving them unchanged) and leaves fmx / fmn as they are, and the limiter
// skips periodic faces. Cells with more faces get a wrong value here and are recomputed by the CSR
// path afterwards. Interior = true is the form for rows whose faces are all internal (no mask
// tests, so the row's loads merge and nothing branches per face). Distribution values are never -0
// (they start >= +0 and a sum is -0 only if both terms are), so nmax / nmin equal dmax / dmin on
// them; for the projections, +0 and -0 both give the limiter value 1 and fail proj_min < 0.
template <bool Interior, class F>
__device__ __forceinline__ void limited_gradient_row(const View<F>& v, const EllRow<F>& r, int cell, int p, F x, F y,
                                                     F z, F fi, F& ogx, F& ogy, F& ogz) {
  const F* fp = v.f + p;
  const unsigned skip = Interior ? 0u : r.skip, bnd = Interior ? 0u : r.bnd;
  F fk[kGather];
#pragma unroll
  for (int k = 0; k < kGather; ++k) fk[k] = fp[r.nbf[k]];  // boundary and padding faces: the cell's own row
  F gx = F(0), gy = F(0), gz = F(0);
  F fmx = fi, fmn = fi;
#pragma unroll
  for (int k = 0; k < kGather; ++k) {
    if ((skip >> k) & 1u) continue;
    if ((bnd >> k) & 1u) fk[k] = boundary_value_first_order(v, r.bface[k], r.sx[k], r.sy[k], r.sz[k], cell, p, x, y, z, fi);
    const F d = fk[k] - fi;
    gx = madd(gx, r.cx[k], d);
    gy = madd(gy, r.cy[k], d);
    gz = madd(gz, r.cz[k], d);
    fmx = nmax(fmx, fk[k]);
    fmn = nmin(fmn, fk[k]);
  }
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
  F proj_max = F(0), proj_min = F(0);
#pragma unroll
  for (int k = 0; k < kGather; ++k) {
    if ((skip >> k) & 1u) continue;
    const F proj = dot3(gx, gy, gz, r.dsx[k], r.dsy[k], r.dsz[k]);
    proj_max = nmax(proj_max, proj);
    proj_min = nmin(proj_min, proj);
  }
  limit_from_extremes(v, cell, fi, fmx, fmn, proj_max, proj_min, gx, gy, gz, ogx, ogy, ogz);
}

// The row form of limited_gradient_on_branch (the gradient kernel's path).
template <int Mode, class F>
__device__ void limited_gradient_row_on_branch(const View<F>& v, const EllRow<F>& r, int cell, int p, F x, F y, F z,
                                               F fi, F& ogx, F& ogy, F& ogz) {
  const F* fp = v.f + p;
  F fk[kGather];
  unsigned use = 0;
  F gx = F(0), gy = F(0), gz = F(0);
#pragma unroll
  for (int k = 0; k < kGather; ++k) fk[k] = fp[r.nbf[k]];
#pragma unroll
  for (int k = 0; k < kGather; ++k) {
    if ((r.skip >> k) & 1u) continue;
    if ((r.bnd >> k) & 1u) fk[k] = boundary_value_first_order(v, r.bface[k], r.sx[k], r.sy[k], r.sz[k], cell, p, x, y, z, fi);
    const F d = fk[k] - fi;
    gx = madd(gx, r.cx[k], d);
    gy = madd(gy, r.cy[k], d);
    gz = madd(gz, r.cz[k], d);
    use |= 1u << k;
  }
  limit_on_branch<Mode>(v, cell, p, fi, fk, use, r.dsx, r.dsy, r.dsz, gx, gy, gz, ogx, ogy, ogz);
}

// outgoing reconstructed value at a boundary face (kinetic_kernels.hpp:204-211)
template <int Mode = 0, class F>
__device__ F reconstruct_at_boundary(const View<F>& v, const BFace<F>& b, int p, F x, F y, F z) {
  const F fi = f_at(v, b.owner, p);
  if (b.kind == kBcPeriodic) return fi;
  F gx, gy, gz;
  limited_gradient<Mode>(v, b.owner, p, x, y, z, fi, gx, gy, gz);
  return fi + gx * b.dx + gy * b.dy + gz * b.dz;
}

// specular reflection about the face normal, trilinear in velocity with a Maxwellian-ratio weight
// (kinetic_kernels.hpp:80-127)
template <bool Reconstruct, int Mode = 0, class F>
__device__ F reflected_value(const View<F>& v, const BFace<F>& b, int p, F x, F y, F z) {
  if (b.mirror_xor != 0) {
    const int q = p ^ b.mirror_xor;
    if (Reconstruct) return reconstruct_at_boundary<Mode>(v, b, q, v.xi_x[q], v.xi_y[q], v.xi_z[q]);
    return f_at(v, b.owner, q);
  }
  const F un = x * b.nx + y * b.ny + z * b.nz;
  const F r[3] = {x - F(2) * un * b.nx, y - F(2) * un * b.ny, z - F(2) * un * b.nz};
  int base[3];
  F t[3];
  int node_off = 0;
  for (int a = 0; a < 3; ++a) {
    if (v.nodes != nullptr) {
      // stretched grid: the bracketing nodes by bisection, linear in xi between them; outside the
      // node range (by more than 1e-5 of the end spacing, the uniform form's slack) the value is 0
      const double* nd = v.nodes + node_off;
      const int n = v.grid_n[a];
      node_off += n;
      const double ra = static_cast<double>(r[a]);
      if (ra < nd[0] - 1e-5 * (nd[1] - nd[0]) || ra > nd[n - 1] + 1e-5 * (nd[n - 1] - nd[n - 2])) return F(0);
      int lo = 0, hi = n - 1;
      while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (nd[mid] <= ra)
          lo = mid;
        else
          hi = mid;
      }
      base[a] = lo;
      t[a] = static_cast<F>(dmax(0.0, dmin(1.0, (ra - nd[lo]) / (nd[lo + 1] - nd[lo]))));
      continue;
    }
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
      value = reconstruct_at_boundary<Mode>(v, b, q, qx, qy, qz);
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

// The same residual from the cell's padded row: upwind neighbour values (f and its limited gradient)
// are loaded by the lanes that need them, all faces at once, then the fluxes are added in face order.
// Periodic and padding faces are skipped, as the reference skips periodic faces. Interior = true:
// as in limited_gradient_row.
template <bool Interior, class F>
__device__ __forceinline__ F residual_value_row(const View<F>& v, const Chunk<F>& ch, const EllRow<F>& r, int cell,
                                                int t) {
  const int p = ch.k0 + t;
  const F x = v.xi_x[p], y = v.xi_y[p], z = v.xi_z[p];
  const F* fp = v.f + p;
  const F fi = fp[static_cast<i64>(cell) * v.nv];
  const i64 wi = static_cast<i64>(cell) * ch.stride + t;
  const F gix = ch.gx[wi], giy = ch.gy[wi], giz = ch.gz[wi];
  const unsigned skip = Interior ? 0u : r.skip, bnd = Interior ? 0u : r.bnd;
  // ak: face flux coefficient; fup: reconstructed upwind value of an internal face, set only in the
  // lanes that are downwind of it (ak < 0), the only lanes that read it
  F ak[kGather], fup[kGather];
  if constexpr (sizeof(F) == 4) {
    // FP32: every face's neighbour loads are issued before any is used (more registers; measured
    // 5% faster than the per-face form below)
    F fu[kGather], gux[kGather], guy[kGather], guz[kGather];
#pragma unroll
    for (int k = 0; k < kGather; ++k) {
      ak[k] = dot3(x, y, z, r.sx[k], r.sy[k], r.sz[k]);
      fu[k] = gux[k] = guy[k] = guz[k] = F(0);
      if (((skip | bnd) >> k) & 1u) continue;
      if (ak[k] < F(0)) {
        const i64 wj = r.nbw[k] + t;
        fu[k] = fp[r.nbf[k]];
        gux[k] = ch.gx[wj];
        guy[k] = ch.gy[wj];
        guz[k] = ch.gz[wj];
      }
    }
#pragma unroll
    for (int k = 0; k < kGather; ++k) fup[k] = madd(madd(madd(fu[k], gux[k], r.dox[k]), guy[k], r.doy[k]), guz[k], r.doz[k]);
  } else {
    // FP64: the face's value is formed where its loads land (58 VGPRs instead of 104, so twice the
    // waves per SIMD hide the latency; all-loads-first measured 40% slower here)
#pragma unroll
    for (int k = 0; k < kGather; ++k) {
      ak[k] = dot3(x, y, z, r.sx[k], r.sy[k], r.sz[k]);
      fup[k] = F(0);
      if (((skip | bnd) >> k) & 1u) continue;
      if (ak[k] < F(0)) {
        const i64 wj = r.nbw[k] + t;
        fup[k] = madd(madd(madd(fp[r.nbf[k]], ch.gx[wj], r.dox[k]), ch.gy[wj], r.doy[k]), ch.gz[wj], r.doz[k]);
      }
    }
  }
  F R = F(0);
#pragma unroll
  for (int k = 0; k < kGather; ++k) {
    if ((skip >> k) & 1u) continue;
    const F a = ak[k];
    F ff;
    if (a >= F(0)) {
      ff = madd(madd(madd(fi, gix, r.dsx[k]), giy, r.dsy[k]), giz, r.dsz[k]);
    } else if (!((bnd >> k) & 1u)) {
      ff = fup[k];
    } else {
      const int b = r.bface[k];
      const int kind = v.bf[b].kind;
      if (kind == kBcSymmetry) {
        const int tm = t ^ v.bf[b].mirror_xor;
        const i64 wm = static_cast<i64>(cell) * ch.stride + tm;
        ff = f_at(v, cell, ch.k0 + tm) + ch.gx[wm] * r.dsx[k] + ch.gy[wm] * r.dsy[k] + ch.gz[wm] * r.dsz[k];
      } else if (kind == kBcMaxwell) {
        ff = maxwell_inflow(v, b, p, x, y, z);
      } else {
        ff = inflow_value(v.inflow[b], x, y, z);
      }
    }
    R = madd(R, a, ff);
  }
  const CellP<F>& cp = v.cells[cell];
  R -= cp.coll * (target_eval(cp.s, x, y, z) - fi);
  return R;
}

}  // namespace gsk::dev
