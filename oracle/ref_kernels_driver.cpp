// TEST INFRASTRUCTURE ONLY -- builds oracle/_ref/libref_kernels.so from the reference's own per-point kernel
// headers, which are compiled here straight from the reference tree (-I <reference>/src; nothing is
// copied). Used only by tests/ to pin the oracle; never shipped, never loaded by the library.
//
// Only the header-only kernel layer of the reference compiles with this toolchain
// (kinetic_kernels.hpp, discrete_equilibrium.hpp, kinetic_types.hpp, physics/*.hpp); the backend
// translation unit src/kinetic/kinetic_backend_cpu.cpp needs <format>, which libstdc++ 11 lacks.
// So the per-method loops below restate kinetic_backend_cpu.cpp (:64-432) around the reference's
// own per-point functions (limited_gradient, residual_value, sweep_*_cell, moment_terms,
// finalize_moments, make_cell_params, discrete_maxwell_seq, reflected_value, add_boundary_terms,
// internal_face_flux_term and the GSIS-correction helpers). Exposes the gsk_backend_ops table.

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../include/gsis_kinetic.h"
#include "kinetic/kinetic_kernels.hpp"

namespace refdrv {

using namespace gsis;
using namespace gsis::kinetic;

template <class F>
class RefBackend {
 public:
  explicit RefBackend(const gsk_setup_view& s) : sv_(s) {
    nc_ = s.n_cells;
    nv_ = s.nv;
    nb_ = s.n_bfaces;
    for (int p = 0; p < nv_; ++p) {
      xd_.push_back(s.xi_x[p]);
      yd_.push_back(s.xi_y[p]);
      zd_.push_back(s.xi_z[p]);
      xx_.push_back(static_cast<F>(s.xi_x[p]));
      xy_.push_back(static_cast<F>(s.xi_y[p]));
      xz_.push_back(static_cast<F>(s.xi_z[p]));
    }
    off_.assign(s.csr_offsets, s.csr_offsets + nc_ + 1);
    for (int k = 0; k < off_[nc_]; ++k) {
      const gsk_csr_entry& e = s.csr[k];
      CsrEntry<F> o;
      o.other = e.other;
      o.bface = e.bface;
      o.sx = F(e.s[0]), o.sy = F(e.s[1]), o.sz = F(e.s[2]);
      o.dsx = F(e.ds[0]), o.dsy = F(e.ds[1]), o.dsz = F(e.ds[2]);
      o.dox = F(e.dother[0]), o.doy = F(e.dother[1]), o.doz = F(e.dother[2]);
      o.cx = F(e.c[0]), o.cy = F(e.c[1]), o.cz = F(e.c[2]);
      csr_.push_back(o);
    }
    for (int c = 0; c < nc_; ++c) eps2_.push_back(F(s.eps2[c]));
    for (int b = 0; b < nb_; ++b) {
      const gsk_bface& i = s.bfaces[b];
      BoundaryFaceInfo<F> o;
      o.owner = i.owner;
      o.kind = i.kind;
      o.mirror_xor = i.mirror_xor;
      o.patch = i.patch;
      o.nx = F(i.n[0]), o.ny = F(i.n[1]), o.nz = F(i.n[2]);
      o.area = F(i.area);
      o.dx = F(i.d[0]), o.dy = F(i.d[1]), o.dz = F(i.d[2]);
      bfaces_.push_back(o);
    }
    binflow_.assign(nb_, BoundaryInflow<F>{});
    if (s.n_maxwell_faces > 0) {
      for (int b = 0; b < nb_; ++b) maxwell_.push_back(MaxwellFace{s.maxwell_offset[b], s.maxwell_specular[b]});
      specular_.assign(static_cast<std::size_t>(s.n_maxwell_faces) * nv_, F(0));
      specular_first_.assign(specular_.size(), F(0));
    }
    g2s_.assign(s.grid_to_storage, s.grid_to_storage + nv_);
    blocks_.assign(s.block_offsets, s.block_offsets + s.n_blocks + 1);
    if (s.n_sweep_dirs > 0) {
      order_.assign(s.sweep_order, s.sweep_order + static_cast<long long>(s.n_sweep_dirs) * nc_);
      pos_.assign(s.sweep_pos, s.sweep_pos + static_cast<long long>(s.n_sweep_dirs) * nc_);
      dir_of_.assign(s.sweep_dir_of, s.sweep_dir_of + nv_);
    }
    for (int a = 0; a < 3; ++a) {
      deq_.lo[a] = s.deq_lo[a];
      deq_.hi[a] = s.deq_hi[a];
      deq_.d[a] = s.deq_d[a];
    }
    deq_.criterion = s.deq_criterion;
    deq_.max_iterations = s.deq_max_iterations;
    deq_.tolerance = s.deq_tolerance;
    long long k = s.velocity_chunk > 0 ? s.velocity_chunk : nv_;
    k = (k + s.orbit_size - 1) / s.orbit_size * s.orbit_size;
    chunk_ = static_cast<int>(std::min<long long>(k, nv_));
    cells_.assign(nc_, CellParams<F>{});
    f_.assign(static_cast<std::size_t>(nc_) * nv_, F(0));
    moments_.assign(static_cast<std::size_t>(nc_) * kMomentsPerCell, 0.0);
    shift_.assign(static_cast<std::size_t>(nc_) * 3, 0.0);
    const std::size_t work = static_cast<std::size_t>(nc_) * chunk_;
    gx_.assign(work, F(0));
    gy_.assign(work, F(0));
    gz_.assign(work, F(0));
    r_.assign(work, F(0));
  }

  int chunk() const { return chunk_; }

  KineticView<F> view() {
    KineticView<F> v;
    v.n_cells = nc_;
    v.nv = nv_;
    v.n_bfaces = nb_;
    v.xi_x = xx_.data();
    v.xi_y = xy_.data();
    v.xi_z = xz_.data();
    v.weight = static_cast<F>(sv_.weight);
    v.csr_offsets = off_.data();
    v.csr = csr_.data();
    v.eps2 = eps2_.data();
    v.limiter_mode = sv_.limiter_mode;
    v.positivity = sv_.positivity;
    v.diag_scale = static_cast<F>(sv_.diag_scale);
    v.interface_damping = static_cast<F>(sv_.interface_damping);
    v.bfaces = bfaces_.data();
    v.binflow = binflow_.data();
    v.cells = cells_.data();
    v.f = f_.data();
    if (sv_.n_maxwell_faces > 0) {
      v.maxwell = maxwell_.data();
      v.specular = specular_.data();
      v.specular_first = specular_first_.data();
      v.grid_to_storage = g2s_.data();
      for (int a = 0; a < 3; ++a) v.grid_n[a] = sv_.grid_n[a];
    }
    v.n_sweep_dirs = sv_.n_sweep_dirs;
    if (sv_.n_sweep_dirs > 0) {
      v.sweep_order = order_.data();
      v.sweep_pos = pos_.data();
      v.sweep_dir_of = dir_of_.data();
    }
    v.discrete_eq = sv_.discrete_eq;
    v.deq = deq_;
    v.deq_fallback = sv_.deq_fallback_collisionless;
    v.cold_bgk_temperature = sv_.cold_bgk_temperature;
    return v;
  }

  MaxwellParams<F> equilibrium_params(double rho, double ux, double uy, double uz, double T, int* result = nullptr) {
    if (result) *result = 0;
    if (!sv_.discrete_eq) return maxwell_params_from<F>(rho, ux, uy, uz, T);
    MaxwellParams<F> m;
    double gamma[5];
    const int r = discrete_maxwell_seq<F>(xx_.data(), xy_.data(), xz_.data(), nv_, sv_.weight, deq_, rho, ux, uy, uz, T,
                                          nullptr, m, gamma);
    if (result) *result = r;
    return m;
  }

  void set_equilibrium(const double* st) {
    const KineticView<F> v = view();
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc_; ++c) {
      const double* s = st + static_cast<std::size_t>(c) * 5;
      const MaxwellParams<F> m = equilibrium_params(s[0], s[1], s[2], s[3], s[4]);
      for (int p = 0; p < nv_; ++p) v.f[static_cast<i64>(c) * nv_ + p] = maxwellian(m, velocity_at(v, p));
      shift_[c * 3 + 0] = s[1];
      shift_[c * 3 + 1] = s[2];
      shift_[c * 3 + 2] = s[3];
    }
  }
  void set_distribution(const double* f) {
    for (std::size_t i = 0; i < f_.size(); ++i) f_[i] = static_cast<F>(f[i]);
  }
  void get_distribution(double* f) {
    for (std::size_t i = 0; i < f_.size(); ++i) f[i] = static_cast<double>(f_[i]);
  }
  void compute_moments(double* out) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc_; ++c) {
      const double u0x = shift_[c * 3], u0y = shift_[c * 3 + 1], u0z = shift_[c * 3 + 2];
      double sums[kMomentSums] = {};
      double t[kMomentSums];
      for (int p = 0; p < nv_; ++p) {
        moment_terms<double>(static_cast<double>(f_[static_cast<std::size_t>(c) * nv_ + p]), xd_[p] - u0x, yd_[p] - u0y,
                             zd_[p] - u0z, t);
        for (int k = 0; k < kMomentSums; ++k) sums[k] += t[k];
      }
      double* m = &moments_[static_cast<std::size_t>(c) * kMomentsPerCell];
      finalize_moments(sums, sv_.weight, u0x, u0y, u0z, m);
      if (m[0] > 0.0) {
        shift_[c * 3 + 0] = m[1];
        shift_[c * 3 + 1] = m[2];
        shift_[c * 3 + 2] = m[3];
      }
    }
    std::copy(moments_.begin(), moments_.end(), out);
  }
  void scale_distribution(double s) {
    for (F& x : f_) x *= static_cast<F>(s);
    for (int c = 0; c < nc_; ++c) {
      double* m = &moments_[static_cast<std::size_t>(c) * kMomentsPerCell];
      m[0] *= s;
      for (int k = 5; k < kMomentsPerCell; ++k) m[k] *= s;
    }
  }
  void update_cell_params() {
    flags_.assign(nc_, 0);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nc_; ++c) {
      const double* m = &moments_[static_cast<std::size_t>(c) * kMomentsPerCell];
      CellParams<F> cp = make_cell_params<F>(m, sv_.volume[c], sv_.vdt[c], sv_.mu0, sv_.omega, sv_.Pr,
                                             sv_.cold_bgk_temperature);
      if (sv_.discrete_eq) {
        MaxwellParams<F> eq;
        double gamma[5];
        const int r = discrete_maxwell_seq<F>(xx_.data(), xy_.data(), xz_.data(), nv_, sv_.weight, deq_, m[0], m[1], m[2],
                                              m[3], m[4], &cp.s, eq, gamma);
        if (r > 0) {
          cp.s.eq = eq;
          cp.s.g0 = static_cast<F>(gamma[0]);
          cp.s.g = Vec3T<F>{static_cast<F>(gamma[1]), static_cast<F>(gamma[2]), static_cast<F>(gamma[3])};
          cp.s.g4 = static_cast<F>(gamma[4]);
        } else if (r < 0 && sv_.deq_fallback_collisionless) {
          cp.coll = F(0);
          cp.diag = static_cast<F>(sv_.vdt[c]);
        }
        flags_[c] = r;
      }
      cells_[c] = cp;
    }
  }
  gsk_deq_counts deq_counts() {
    gsk_deq_counts o{0, 0, 0};
    for (int x : flags_) {
      if (x > 0) ++o.corrected;
      if (x < 0) ++o.failed;
      o.max_iterations = std::max(o.max_iterations, x);
    }
    return o;
  }
  void set_inflow(const gsk_inflow_state* in) {
    for (int b = 0; b < nb_; ++b) {
      binflow_[b].m = maxwell_params_from<F>(in[b].rho, in[b].ux, in[b].uy, in[b].uz, in[b].T);
      binflow_[b].scale = static_cast<F>(in[b].scale);
    }
  }
  void boundary_sums(double* out) {
    const KineticView<F> v = view();
    const F w = static_cast<F>(sv_.weight);
    if (sv_.n_maxwell_faces > 0) {
      for (int b = 0; b < nb_; ++b) {
        if (maxwell_[b].offset < 0) continue;
        const auto& bf = bfaces_[b];
        for (int p = 0; p < nv_; ++p) {
          const auto xi = velocity_at(v, p);
          specular_first_[maxwell_[b].offset + p] =
              dot(xi, Vec3T<F>{bf.nx, bf.ny, bf.nz}) < F(0) ? reflected_value<false>(v, bf, p, xi) : F(0);
        }
      }
      for (int b = 0; b < nb_; ++b) {
        if (maxwell_[b].offset < 0) continue;
        const auto& bf = bfaces_[b];
        for (int p = 0; p < nv_; ++p) {
          const auto xi = velocity_at(v, p);
          specular_[maxwell_[b].offset + p] =
              dot(xi, Vec3T<F>{bf.nx, bf.ny, bf.nz}) < F(0) ? reflected_value<true>(v, bf, p, xi) : F(0);
        }
      }
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int b = 0; b < nb_; ++b) {
      const BoundaryFaceInfo<F>& bf = bfaces_[b];
      Accum<double> acc[kBoundarySumsPerFace];
      for (int p = 0; p < nv_; ++p) {
        const Vec3T<F> xi = velocity_at(v, p);
        const F a_n = xi.x * bf.nx + xi.y * bf.ny + xi.z * bf.nz;
        if (a_n >= F(0) || bf.kind == kBcPeriodic) {
          const F fb = reconstruct_at_boundary(v, bf, p, xi);
          add_boundary_terms<F, Accum<double>>(w * fb, xi, a_n, acc + 0, acc + 10);
        } else {
          F fb;
          if (bf.kind == kBcSymmetry) {
            const int pm = p ^ bf.mirror_xor;
            fb = reconstruct_at_boundary(v, bf, pm, velocity_at(v, pm));
          } else if (bf.kind == kBcMaxwellWall) {
            fb = maxwell_inflow(v, b, p, xi);
          } else {
            fb = maxwellian(binflow_[b].m, xi);
          }
          add_boundary_terms<F, Accum<double>>(w * fb, xi, a_n, acc + 5, acc + 15);
        }
      }
      for (int k = 0; k < kBoundarySumsPerFace; ++k) out[static_cast<std::size_t>(b) * kBoundarySumsPerFace + k] = acc[k].value();
    }
  }
  void face_flux(const int* cells, const int* entries, std::size_t n, double* out) {
    const KineticView<F> v = view();
    for (std::size_t i = 0; i < n; ++i) {
      Accum<double> acc;
      for (int p = 0; p < nv_; ++p)
        acc.add(static_cast<double>(internal_face_flux_term(v, cells[i], csr_[entries[i]], p, velocity_at(v, p))));
      out[i] = static_cast<double>(v.weight) * acc.value();
    }
  }
  void sweep(double* rc) {
    std::fill(rc, rc + nc_, 0.0);
    const KineticView<F> v = view();
    const int n_blocks = static_cast<int>(blocks_.size()) - 1;
    for (int k0 = 0; k0 < nv_; k0 += chunk_) {
      ChunkView<F> ch;
      ch.k0 = k0;
      ch.kc = std::min(chunk_, nv_ - k0);
      ch.stride = chunk_;
      ch.gx = gx_.data();
      ch.gy = gy_.data();
      ch.gz = gz_.data();
      ch.r = r_.data();
#pragma omp parallel for schedule(static)
      for (int c = 0; c < nc_; ++c)
        for (int t = 0; t < ch.kc; ++t) {
          const int p = k0 + t;
          const Vec3T<F> g = limited_gradient(v, c, p, velocity_at(v, p), f_at(v, c, p));
          const i64 wi = static_cast<i64>(c) * ch.stride + t;
          ch.gx[wi] = g.x;
          ch.gy[wi] = g.y;
          ch.gz[wi] = g.z;
        }
#pragma omp parallel for schedule(static)
      for (int c = 0; c < nc_; ++c) {
        double r2 = 0.0;
        for (int t = 0; t < ch.kc; ++t) {
          const F R = residual_value(v, ch, c, t);
          ch.r[static_cast<i64>(c) * ch.stride + t] = R;
          r2 += static_cast<double>(R) * static_cast<double>(R);
        }
        rc[c] += r2;
      }
      const int passes = sv_.block_passes;
      if (passes > 1) std::copy(r_.begin(), r_.end(), gx_.begin());
      for (int pass = 0; pass < passes; ++pass) {
        ChunkView<F> cp = ch;
        if (pass > 0) {
          std::copy(r_.begin(), r_.end(), gy_.begin());
          cp.r0 = gx_.data();
          cp.dprev = gy_.data();
        }
        cp.update_f = pass + 1 == passes ? 1 : 0;
#pragma omp parallel for schedule(static)
        for (int lane = 0; lane < n_blocks * ch.kc; ++lane) {
          const int b = lane / ch.kc, t = lane % ch.kc;
          const Vec3T<F> xi = velocity_at(v, k0 + t);
          const int c0 = blocks_[b], c1 = blocks_[b + 1];
          const std::int32_t* order;
          const std::int32_t* pos;
          sweep_order_of(v, k0 + t, order, pos);
          for (int k = c0; k < c1; ++k) sweep_forward_cell(v, cp, order ? order[k] : k, t, c0, c1, xi, pos);
        }
#pragma omp parallel for schedule(static)
        for (int lane = 0; lane < n_blocks * ch.kc; ++lane) {
          const int b = lane / ch.kc, t = lane % ch.kc;
          const Vec3T<F> xi = velocity_at(v, k0 + t);
          const int c0 = blocks_[b], c1 = blocks_[b + 1];
          const std::int32_t* order;
          const std::int32_t* pos;
          sweep_order_of(v, k0 + t, order, pos);
          for (int k = c1 - 1; k >= c0; --k) sweep_backward_cell(v, cp, order ? order[k] : k, t, c0, c1, xi, pos);
        }
      }
    }
  }
  long long gsis_correction(const double* w_half, const double* w_new, int* flags, bool have_flags) {
    const KineticView<F> v = view();
    const int mode = sv_.negative_check;
    long long replaced = 0;
    for (int c = 0; c < nc_; ++c) {
      const double* wh = &w_half[static_cast<std::size_t>(c) * 5];
      const double* wn = &w_new[static_cast<std::size_t>(c) * 5];
      int rh = 0, rn = 0;
      const MaxwellParams<F> mh = equilibrium_params(wh[0], wh[1], wh[2], wh[3], wh[4], &rh);
      const MaxwellParams<F> mn = equilibrium_params(wn[0], wn[1], wn[2], wn[3], wn[4], &rn);
      if (sv_.deq_fallback_collisionless && (rh < 0 || rn < 0)) {
        ++replaced;
        if (have_flags) flags[c] = kCorrectionSkipped;
        shift_[c * 3 + 0] = wh[1];
        shift_[c * 3 + 1] = wh[2];
        shift_[c * 3 + 2] = wh[3];
        continue;
      }
      F* fc = v.f + static_cast<i64>(c) * nv_;
      const double threshold = -1e-8 * wn[0] / (wn[4] * std::sqrt(wn[4]));
      bool replace = false;
      if (mode == kNegativeLimit) {
        F alpha = F(1);
        for (int p = 0; p < nv_; ++p) {
          const Vec3T<F> xi = velocity_at(v, p);
          alpha = std::min(alpha, correction_alpha_limit(fc[p], maxwellian(mn, xi) - maxwellian(mh, xi),
                                                          static_cast<F>(threshold)));
        }
        for (int p = 0; p < nv_; ++p) {
          const Vec3T<F> xi = velocity_at(v, p);
          fc[p] += alpha * (maxwellian(mn, xi) - maxwellian(mh, xi));
        }
        replace = alpha < F(1);
      } else if (mode == kNegativeBlend) {
        F theta = F(0);
        for (int p = 0; p < nv_; ++p) {
          const Vec3T<F> xi = velocity_at(v, p);
          const F e = maxwellian(mn, xi);
          theta = std::max(theta, blend_theta(fc[p], e - maxwellian(mh, xi), e, static_cast<F>(threshold)));
        }
        for (int p = 0; p < nv_; ++p) {
          const Vec3T<F> xi = velocity_at(v, p);
          const F e = maxwellian(mn, xi);
          const F g = fc[p] + (e - maxwellian(mh, xi));
          fc[p] = (F(1) - theta) * g + theta * e;
        }
        replace = theta > F(0);
      } else {
        F fmin = F(0);
        bool first = true;
        for (int p = 0; p < nv_; ++p) {
          const Vec3T<F> xi = velocity_at(v, p);
          const F old = fc[p];
          const F x = old + (maxwellian(mn, xi) - maxwellian(mh, xi));
          fc[p] = x;
          const F chk = negative_check_value(old, x, mode == kNegativeCreated);
          fmin = first ? chk : std::min(fmin, chk);
          first = false;
        }
        replace = static_cast<double>(fmin) < threshold;
        if (replace)
          for (int p = 0; p < nv_; ++p) fc[p] = maxwellian(mn, velocity_at(v, p));
      }
      if (replace) ++replaced;
      if (have_flags) flags[c] = replace ? 1 : 0;
      shift_[c * 3 + 0] = wn[1];
      shift_[c * 3 + 1] = wn[2];
      shift_[c * 3 + 2] = wn[3];
    }
    return replaced;
  }

  std::size_t f_size() const { return f_.size(); }
  int n_cells() const { return nc_; }
  int n_bfaces() const { return nb_; }

 private:
  gsk_setup_view sv_;
  int nc_ = 0, nv_ = 0, nb_ = 0, chunk_ = 0;
  DeqGrid deq_{};
  std::vector<double> xd_, yd_, zd_, moments_, shift_;
  std::vector<F> xx_, xy_, xz_, eps2_, f_, gx_, gy_, gz_, r_, specular_, specular_first_;
  std::vector<std::int32_t> off_, blocks_, order_, pos_;
  std::vector<int> g2s_, flags_;
  std::vector<std::uint8_t> dir_of_;
  std::vector<CsrEntry<F>> csr_;
  std::vector<BoundaryFaceInfo<F>> bfaces_;
  std::vector<BoundaryInflow<F>> binflow_;
  std::vector<MaxwellFace> maxwell_;
  std::vector<CellParams<F>> cells_;
};

struct Handle {
  std::unique_ptr<RefBackend<double>> d;
  std::unique_ptr<RefBackend<float>> s;
  std::string err;
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

}  // namespace refdrv

extern "C" {

void* ref_create(const gsk_setup_view* view, int precision) {
  auto h = std::make_unique<refdrv::Handle>();
  const int fp64 = precision < 0 ? view->precision_fp64 : precision;
  if (fp64)
    h->d = std::make_unique<refdrv::RefBackend<double>>(*view);
  else
    h->s = std::make_unique<refdrv::RefBackend<float>>(*view);
  return h.release();
}
void ref_destroy(void* h) { delete static_cast<refdrv::Handle*>(h); }

static const char* r_name(void*) { return "reference-kernels"; }
static int r_chunk(void* ctx, int32_t* out) {
  return refdrv::call(ctx, [&](auto& b) { *out = b.chunk(); });
}
static int r_mem(void*, uint64_t* out) {
  *out = 0;
  return GSK_OK;
}
static int r_set_eq(void* ctx, const double* s, size_t n) {
  return refdrv::call(ctx, [&](auto& b) {
    if (n != static_cast<size_t>(b.n_cells()) * 5) throw std::runtime_error("states length");
    b.set_equilibrium(s);
  });
}
static int r_set_f(void* ctx, const double* f, size_t n) {
  return refdrv::call(ctx, [&](auto& b) {
    if (n != b.f_size()) throw std::runtime_error("distribution length");
    b.set_distribution(f);
  });
}
static int r_get_f(void* ctx, double* f, size_t n) {
  return refdrv::call(ctx, [&](auto& b) {
    if (n != b.f_size()) throw std::runtime_error("distribution length");
    b.get_distribution(f);
  });
}
static int r_moments(void* ctx, double* out, size_t) {
  return refdrv::call(ctx, [&](auto& b) { b.compute_moments(out); });
}
static int r_scale(void* ctx, double s) {
  return refdrv::call(ctx, [&](auto& b) { b.scale_distribution(s); });
}
static int r_params(void* ctx) {
  return refdrv::call(ctx, [&](auto& b) { b.update_cell_params(); });
}
static int r_inflow(void* ctx, const gsk_inflow_state* in, size_t) {
  return refdrv::call(ctx, [&](auto& b) { b.set_inflow(in); });
}
static int r_bsums(void* ctx, double* out, size_t) {
  return refdrv::call(ctx, [&](auto& b) { b.boundary_sums(out); });
}
static int r_sweep(void* ctx, double* r, size_t) {
  return refdrv::call(ctx, [&](auto& b) { b.sweep(r); });
}
static int r_gsis(void* ctx, const double* wh, const double* wn, size_t, int32_t* flags, size_t nf, long long* rep) {
  return refdrv::call(ctx, [&](auto& b) { *rep = b.gsis_correction(wh, wn, flags, flags != nullptr && nf > 0); });
}
static int r_flux(void* ctx, const int32_t* c, const int32_t* e, size_t n, double* out) {
  return refdrv::call(ctx, [&](auto& b) { b.face_flux(c, e, n, out); });
}
static int r_timings(void*, gsk_timings* out) {
  *out = gsk_timings{0, 0, 0, 0, 0, 0, 0};
  return GSK_OK;
}
static int r_deq(void* ctx, gsk_deq_counts* out) {
  return refdrv::call(ctx, [&](auto& b) { *out = b.deq_counts(); });
}
static const char* r_err(void* ctx) { return static_cast<refdrv::Handle*>(ctx)->err.c_str(); }

int ref_fill_ops(void* h, gsk_backend_ops* ops) {
  if (!h || !ops) return GSK_ERR_INVALID;
  *ops = gsk_backend_ops{h,        r_name,  r_chunk, r_mem,  r_set_eq, r_set_f,   r_get_f, r_moments, r_scale, r_params,
                         r_inflow, r_bsums, r_sweep, r_gsis, r_flux,   r_timings, r_deq,   r_err};
  return GSK_OK;
}

}  // extern "C"
