// Adjoint (transposed Jacobian) of the steady residual and of the patch mass flows, host part (no
// counterpart in the reference, which has no derivative capability).
//
// The residual R(f) is evaluated after prepare()'s boundary rounds, whose fixed point ties f to the
// boundary state z = (wall inflow scales, boundary face states):
//   pressure faces   b_P = H_P(moments, face states)  reconstruction, characteristic condition,
//                                                     discrete-equilibrium correction (host)
//   boundary sums    S   = Sigma(f, b)                device (b: inflow states of all faces)
//   absorb           z   = H_A(S, z)                  wall densities, face states, patch flows (host)
// and R reads f, the cell parameters (from the moments) and b. The device applies the transposes of
// R's direct dependence on f, of R's dependence on b (as the adjoint of each face's inflow
// parameters), of Sigma's dependence on f, and of the moments' chain. The host tabulates, once at the
// linearisation state, the Jacobians of H_A (per face), of H_P and of Sigma with respect to b, by
// central differences: perturbations that share a colour are applied together, and colours are
// chosen so that no output reads two perturbed inputs (faces of one owner cell get distinct colours;
// cells in one pressure-face stencil get distinct colours). Each product then solves the transposed
// fixed point  zeta = B_z^T (R_b^T lam + Sigma_b^T Lsum),  Lsum = H_A,S^T zeta + mass-flow weights,
// by the iteration that mirrors the forward rounds.

#include <algorithm>
#include <array>
#include <cmath>

#include "../kernels/kinetic_math.hpp"
#include "solver.hpp"

namespace gsk {

namespace {
constexpr double kStep = 1e-6;  // relative central-difference step of the boundary tables

double& state_component(gsk_inflow_state& s, int q) {
  switch (q) {
    case 0: return s.scale;
    case 1: return s.rho;
    case 2: return s.ux;
    case 3: return s.uy;
    case 4: return s.uz;
    default: return s.T;
  }
}
}  // namespace

struct KineticSolver::AdjointTables {
  // per boundary face: 0 no state, 1 pressure face (state rho, u, T), 2 wall (inflow scale)
  std::vector<int> kind, patch;
  std::vector<gsk_inflow_state> inflow0;
  // absorb, per face: d(scale', face state (5), patch-flow share) / d(sums (20), scale), [o * 21 + q]
  std::vector<std::array<double, 7 * 21>> absorb;
  // d sums_j / d b_{src, comp} (comp: 0 scale, 1 rho, 2-4 u, 5 T)
  struct SumCol {
    int j, src, comp;
    std::array<double, 20> d;
  };
  std::vector<SumCol> sums_b;
  // d b_P(j) (rho, ux, uy, uz, T) / d input: a moment (cell src, moment comp) or a face state
  // (face src, comp)
  struct PCol {
    int j, src, comp;
    std::array<double, 5> d;
  };
  std::vector<PCol> p_moments, p_faces;
};

void KineticSolver::adjoint_prepare() {
  GSK_CHECK(backend_->element_size() == 8, "adjoint operators need the FP64 form ([kinetic].precision = \"fp64\")");
  GSK_CHECK(freeze_ == Freeze::Follow,
            "adjoint operators need the limiter branches recorded at the linearisation state (freeze_limiters(true), "
            "then a residual evaluation there)");
  GSK_CHECK(!s_.closed_domain, "adjoint operators do not cover closed domains (mass rescaling)");
  for (const PatchCondition& pc : bnd_.patches())
    GSK_CHECK(pc.type != BoundaryType::MassFlowInlet, "adjoint operators do not cover mass-flow inlets");
  prepare();
  GSK_CHECK(!prepared_on_device_, "the adjoint tables need the host's boundary state of the last prepare");
  auto A = std::make_shared<AdjointTables>();
  const int nb = s_.n_bfaces(), nc = s_.n_cells();
  const auto unb = static_cast<std::size_t>(nb);
  A->kind.assign(unb, 0);
  A->patch.assign(unb, 0);
  for (int b = 0; b < nb; ++b) {
    const PatchCondition& pc = bnd_.patch_of(b);
    A->patch[static_cast<std::size_t>(b)] = s_.bface_patch[static_cast<std::size_t>(b)];
    if (pc.kind == kBcPressure) A->kind[static_cast<std::size_t>(b)] = 1;
    if (pc.kind == kBcDiffuseWall || pc.kind == kBcMaxwellWall) A->kind[static_cast<std::size_t>(b)] = 2;
  }
  A->inflow0 = bnd_.inflow;
  const BoundaryModel base(bnd_);
  const std::vector<double> sums0 = bsums_;  // the last round's sums

  // boundary faces of every owner cell and each face's rank among them (its colour)
  std::vector<int> owner(unb), colour(unb);
  std::vector<std::vector<int>> faces_of(static_cast<std::size_t>(nc));
  for (int b = 0; b < nb; ++b) {
    owner[static_cast<std::size_t>(b)] = s_.bfaces[static_cast<std::size_t>(b)].owner;
    auto& list = faces_of[static_cast<std::size_t>(owner[static_cast<std::size_t>(b)])];
    colour[static_cast<std::size_t>(b)] = static_cast<int>(list.size());
    list.push_back(b);
  }
  int n_face_colours = 0;
  for (const int c : colour) n_face_colours = std::max(n_face_colours, c + 1);
  auto face_of_colour = [&](int j, int col) {
    const auto& list = faces_of[static_cast<std::size_t>(owner[static_cast<std::size_t>(j)])];
    return col < static_cast<int>(list.size()) ? list[static_cast<std::size_t>(col)] : -1;
  };

  // ---- absorb: face-local, every face perturbed at once ----
  {
    A->absorb.assign(unb, {});
    std::vector<double> ref(unb);
    for (std::size_t j = 0; j < unb; ++j)
      ref[j] = std::max(std::fabs(sums0[j * 20 + 10]) + std::fabs(sums0[j * 20 + 15]), 1e-12);
    auto run = [&](int q, double sign, std::vector<std::array<double, 7>>& out, std::vector<double>& step) {
      BoundaryModel B(base);
      std::vector<double> sums = sums0;
      for (std::size_t j = 0; j < unb; ++j) {
        if (q < 20) {
          step[j] = kStep * (std::fabs(sums0[j * 20 + static_cast<std::size_t>(q)]) + ref[j]);
          sums[j * 20 + static_cast<std::size_t>(q)] += sign * step[j];
        } else {
          step[j] = kStep * std::max(std::fabs(base.inflow[j].scale), 1e-12);
          B.inflow[j].scale += sign * step[j];
        }
      }
      B.absorb(sums);
      for (std::size_t j = 0; j < unb; ++j)
        out[j] = {B.inflow[j].scale,   B.face_state[j][0], B.face_state[j][1], B.face_state[j][2],
                  B.face_state[j][3], B.face_state[j][4], B.flux_out[j][0] + B.flux_in[j][0]};
    };
    std::vector<std::array<double, 7>> plus(unb), minus(unb);
    std::vector<double> step(unb);
    for (int q = 0; q <= 20; ++q) {
      run(q, 1.0, plus, step);
      run(q, -1.0, minus, step);
      for (std::size_t j = 0; j < unb; ++j)
        for (int o = 0; o < 7; ++o)
          A->absorb[j][static_cast<std::size_t>(o * 21 + q)] =
              (plus[j][static_cast<std::size_t>(o)] - minus[j][static_cast<std::size_t>(o)]) / (2.0 * step[j]);
    }
  }

  // ---- pressure faces: H_P with respect to cell moments and face states ----
  std::vector<int> pfaces;
  {
    BoundaryModel B(base);
    for (const Index b : B.update_pressure_faces(moments_)) pfaces.push_back(static_cast<int>(b));
  }
  if (!pfaces.empty()) {
    auto evaluate = [&](const std::vector<double>& mom, const std::vector<std::array<double, 5>>& fs) {
      BoundaryModel B(base);
      B.face_state = fs;
      const std::vector<Index> set = B.update_pressure_faces(mom);
      std::vector<gsk_inflow_state> batch(set.size());
      for (std::size_t i = 0; i < set.size(); ++i) batch[i] = B.inflow[static_cast<std::size_t>(set[i])];
      if (s_.discrete_eq && !backend_->correct_inflow(batch.data(), batch.size()))
        for (auto& st : batch) discrete_inflow_state(s_, inflow_tol_, st);
      std::vector<std::array<double, 5>> out(batch.size());
      for (std::size_t i = 0; i < batch.size(); ++i) out[i] = {batch[i].rho, batch[i].ux, batch[i].uy, batch[i].uz, batch[i].T};
      return out;
    };
    // stencils: the owner and its internal neighbours
    std::vector<std::vector<int>> stencil(pfaces.size());
    for (std::size_t i = 0; i < pfaces.size(); ++i) {
      const int o = owner[static_cast<std::size_t>(pfaces[i])];
      stencil[i].push_back(o);
      for (int k = s_.csr_offsets[static_cast<std::size_t>(o)]; k < s_.csr_offsets[static_cast<std::size_t>(o) + 1]; ++k)
        if (s_.csr_entries[static_cast<std::size_t>(k)].other >= 0)
          stencil[i].push_back(s_.csr_entries[static_cast<std::size_t>(k)].other);
    }
    // greedy colouring of the stencil cells: no two cells of one stencil share a colour
    std::vector<int> cell_colour(static_cast<std::size_t>(nc), -1);
    std::vector<std::vector<int>> stencils_of(static_cast<std::size_t>(nc));
    for (std::size_t i = 0; i < stencil.size(); ++i)
      for (const int c : stencil[i]) stencils_of[static_cast<std::size_t>(c)].push_back(static_cast<int>(i));
    int n_cell_colours = 0;
    for (std::size_t i = 0; i < stencil.size(); ++i)
      for (const int c : stencil[i]) {
        if (cell_colour[static_cast<std::size_t>(c)] >= 0) continue;
        std::vector<char> used(static_cast<std::size_t>(n_cell_colours) + 1, 0);
        for (const int si : stencils_of[static_cast<std::size_t>(c)])
          for (const int o : stencil[static_cast<std::size_t>(si)])
            if (cell_colour[static_cast<std::size_t>(o)] >= 0) used[static_cast<std::size_t>(cell_colour[static_cast<std::size_t>(o)])] = 1;
        int col = 0;
        while (used[static_cast<std::size_t>(col)]) ++col;
        cell_colour[static_cast<std::size_t>(c)] = col;
        n_cell_colours = std::max(n_cell_colours, col + 1);
      }
    const std::vector<std::array<double, 5>> fs0 = base.face_state;
    static constexpr int kMomentComp[5] = {5, 1, 2, 3, 4};  // p, u, T: what the reconstruction reads
    for (int col = 0; col < n_cell_colours; ++col)
      for (const int k : kMomentComp) {
        std::vector<double> mp = moments_, mm = moments_, step(static_cast<std::size_t>(nc), 0.0);
        for (int c = 0; c < nc; ++c) {
          if (cell_colour[static_cast<std::size_t>(c)] != col) continue;
          const double* m = &moments_[static_cast<std::size_t>(c) * 15];
          const double T = std::max(std::fabs(m[4]), 1e-12);
          const double h = k == 5 ? kStep * std::max(std::fabs(m[5]), 1e-12) : (k == 4 ? kStep * T : kStep * std::sqrt(T));
          step[static_cast<std::size_t>(c)] = h;
          mp[static_cast<std::size_t>(c) * 15 + static_cast<std::size_t>(k)] += h;
          mm[static_cast<std::size_t>(c) * 15 + static_cast<std::size_t>(k)] -= h;
        }
        const auto ep = evaluate(mp, fs0), em = evaluate(mm, fs0);
        for (std::size_t i = 0; i < pfaces.size(); ++i)
          for (const int c : stencil[i]) {
            if (cell_colour[static_cast<std::size_t>(c)] != col) continue;
            AdjointTables::PCol pc{pfaces[i], c, k, {}};
            bool any = false;
            for (int r = 0; r < 5; ++r) {
              pc.d[static_cast<std::size_t>(r)] = (ep[i][static_cast<std::size_t>(r)] - em[i][static_cast<std::size_t>(r)]) /
                                                  (2.0 * step[static_cast<std::size_t>(c)]);
              any = any || pc.d[static_cast<std::size_t>(r)] != 0.0;
            }
            if (any) A->p_moments.push_back(pc);
          }
      }
    for (int col = 0; col < n_face_colours; ++col)
      for (int k = 0; k < 5; ++k) {
        std::vector<std::array<double, 5>> fp = fs0, fm = fs0;
        std::vector<double> step(unb, 0.0);
        for (std::size_t j = 0; j < unb; ++j) {
          if (colour[j] != col) continue;
          const double T = std::max(std::fabs(fs0[j][4]), 1e-12);
          const double h = k == 0 ? kStep * std::max(std::fabs(fs0[j][0]), 1e-12) : (k == 4 ? kStep * T : kStep * std::sqrt(T));
          step[j] = h;
          fp[j][static_cast<std::size_t>(k)] += h;
          fm[j][static_cast<std::size_t>(k)] -= h;
        }
        const auto ep = evaluate(moments_, fp), em = evaluate(moments_, fm);
        for (std::size_t i = 0; i < pfaces.size(); ++i) {
          const int src = face_of_colour(pfaces[i], col);
          if (src < 0) continue;
          AdjointTables::PCol pc{pfaces[i], src, k, {}};
          bool any = false;
          for (int r = 0; r < 5; ++r) {
            pc.d[static_cast<std::size_t>(r)] = (ep[i][static_cast<std::size_t>(r)] - em[i][static_cast<std::size_t>(r)]) /
                                                (2.0 * step[static_cast<std::size_t>(src)]);
            any = any || pc.d[static_cast<std::size_t>(r)] != 0.0;
          }
          if (any) A->p_faces.push_back(pc);
        }
      }
  }

  // ---- boundary sums with respect to the inflow states (device) ----
  {
    std::vector<double> sp(unb * 20), sm(unb * 20);
    for (int col = 0; col < n_face_colours; ++col)
      for (int q = 0; q < 6; ++q) {
        auto relevant = [&](int j) {
          const int kd = A->kind[static_cast<std::size_t>(j)];
          return (q == 0 && kd == 2) || (q > 0 && kd == 1);
        };
        std::vector<gsk_inflow_state> ip = A->inflow0, im = A->inflow0;
        std::vector<double> step(unb, 0.0);
        bool any = false;
        for (int j = 0; j < nb; ++j) {
          if (colour[static_cast<std::size_t>(j)] != col || !relevant(j)) continue;
          gsk_inflow_state& s0 = A->inflow0[static_cast<std::size_t>(j)];
          const double T = std::max(std::fabs(s0.T), 1e-12);
          const double x = state_component(s0, q);
          const double h = (q == 0 || q == 1 || q == 5) ? kStep * std::max(std::fabs(x), 1e-12) : kStep * std::sqrt(T);
          step[static_cast<std::size_t>(j)] = h;
          state_component(ip[static_cast<std::size_t>(j)], q) += h;
          state_component(im[static_cast<std::size_t>(j)], q) -= h;
          any = true;
        }
        if (!any) continue;
        backend_->set_inflow(ip.data(), ip.size());
        backend_->boundary_sums(sp.data(), sp.size());
        backend_->set_inflow(im.data(), im.size());
        backend_->boundary_sums(sm.data(), sm.size());
        for (int j = 0; j < nb; ++j) {
          const int src = face_of_colour(j, col);
          if (src < 0 || !relevant(src)) continue;
          AdjointTables::SumCol sc{j, src, q, {}};
          bool nz = false;
          for (std::size_t r = 0; r < 20; ++r) {
            sc.d[r] = (sp[static_cast<std::size_t>(j) * 20 + r] - sm[static_cast<std::size_t>(j) * 20 + r]) /
                      (2.0 * step[static_cast<std::size_t>(src)]);
            nz = nz || sc.d[r] != 0.0;
          }
          if (nz) A->sums_b.push_back(sc);
        }
      }
    // back to the linearisation state's inflow table and caches
    backend_->set_inflow(bnd_.inflow.data(), bnd_.inflow.size());
    std::vector<double> scratch(unb * 20);
    backend_->boundary_sums(scratch.data(), scratch.size());
  }

  backend_->adjoint_prepare();
  adj_ = A;
  adj_version_ = version_;
}

void KineticSolver::adjoint_apply_device(const void* lam, const double* patch_weights, std::size_t n_patches, void* out,
                                         std::size_t n) {
  GSK_CHECK(adj_ != nullptr && adj_version_ == version_,
            "adjoint operators are not prepared for the current state (adjoint_prepare after the last change)");
  GSK_CHECK(n == static_cast<std::size_t>(s_.n_cells()) * static_cast<std::size_t>(s_.nv()),
            "vector length differs from cells x velocities");
  GSK_CHECK(n_patches == bnd_.patches().size(), "one mass-flow weight per patch is needed");
  const AdjointTables& A = *adj_;
  const auto nb = static_cast<std::size_t>(s_.n_bfaces());
  std::vector<double> lb_dev(nb * 6);
  backend_->adjoint_stage1(lam, out, lb_dev.data(), lb_dev.size());
  // the device's inflow parameters (scale, A, B, u) -> the host states (scale, rho, u, T):
  // A = rho / (2 pi T)^(3/2), B = 1 / (2 T)
  std::vector<std::array<double, 6>> lb0(nb, std::array<double, 6>{});
  for (std::size_t j = 0; j < nb; ++j) {
    const double* l = &lb_dev[j * 6];
    if (A.kind[j] == 2) lb0[j][0] = l[0];
    if (A.kind[j] != 1) continue;
    const gsk_inflow_state& st = A.inflow0[j];
    const double tpt = 2.0 * dev::kPi * st.T, dA_drho = 1.0 / (tpt * std::sqrt(tpt)), Av = st.rho * dA_drho;
    lb0[j][1] = l[1] * dA_drho;
    lb0[j][5] = l[1] * (-1.5 * Av / st.T) + l[2] * (-0.5 / (st.T * st.T));
    lb0[j][2] = l[3];
    lb0[j][3] = l[4];
    lb0[j][4] = l[5];
  }
  // transposed boundary fixed point
  std::vector<double> zsc(nb, 0.0), zsc_new(nb);
  std::vector<std::array<double, 5>> zfs(nb, std::array<double, 5>{}), zfs_new(nb);
  std::vector<double> ls(nb * 20);
  std::vector<std::array<double, 6>> lb(nb);
  bool converged = false;
  for (int it = 0; it < 20000 && !converged; ++it) {
    for (std::size_t j = 0; j < nb; ++j) {
      const double dual[7] = {A.kind[j] == 2 ? zsc[j] : 0.0, zfs[j][0], zfs[j][1], zfs[j][2], zfs[j][3], zfs[j][4],
                              patch_weights[A.patch[j]]};
      for (int q = 0; q < 20; ++q) {
        double s = 0.0;
        for (int o = 0; o < 7; ++o) s += A.absorb[j][static_cast<std::size_t>(o * 21 + q)] * dual[o];
        ls[j * 20 + static_cast<std::size_t>(q)] = s;
      }
    }
    lb = lb0;
    for (const auto& c : A.sums_b) {
      double s = 0.0;
      for (std::size_t r = 0; r < 20; ++r) s += c.d[r] * ls[static_cast<std::size_t>(c.j) * 20 + r];
      lb[static_cast<std::size_t>(c.src)][static_cast<std::size_t>(c.comp)] += s;
    }
    for (std::size_t j = 0; j < nb; ++j) {
      zsc_new[j] = 0.0;
      zfs_new[j] = {};
      if (A.kind[j] != 2) continue;
      const double dual[7] = {zsc[j], zfs[j][0], zfs[j][1], zfs[j][2], zfs[j][3], zfs[j][4], patch_weights[A.patch[j]]};
      double s = lb[j][0];
      for (int o = 0; o < 7; ++o) s += A.absorb[j][static_cast<std::size_t>(o * 21 + 20)] * dual[o];
      zsc_new[j] = s;
    }
    for (const auto& c : A.p_faces) {
      double s = 0.0;
      for (std::size_t r = 0; r < 5; ++r) s += c.d[r] * lb[static_cast<std::size_t>(c.j)][r + 1];
      zfs_new[static_cast<std::size_t>(c.src)][static_cast<std::size_t>(c.comp)] += s;
    }
    double diff = 0.0, mag = 0.0;
    for (std::size_t j = 0; j < nb; ++j) {
      diff = std::max(diff, std::fabs(zsc_new[j] - zsc[j]));
      mag = std::max(mag, std::fabs(zsc_new[j]));
      for (std::size_t r = 0; r < 5; ++r) {
        diff = std::max(diff, std::fabs(zfs_new[j][r] - zfs[j][r]));
        mag = std::max(mag, std::fabs(zfs_new[j][r]));
      }
    }
    converged = diff <= 1e-15 * mag || mag == 0.0;
    zsc.swap(zsc_new);
    zfs.swap(zfs_new);
  }
  GSK_CHECK(converged, "the transposed boundary rounds did not converge");
  std::vector<double> lm(static_cast<std::size_t>(s_.n_cells()) * 15, 0.0);
  for (const auto& c : A.p_moments) {
    double s = 0.0;
    for (std::size_t r = 0; r < 5; ++r) s += c.d[r] * lb[static_cast<std::size_t>(c.j)][r + 1];
    lm[static_cast<std::size_t>(c.src) * 15 + static_cast<std::size_t>(c.comp)] += s;
  }
  backend_->adjoint_stage2(ls.data(), ls.size(), lm.data(), lm.size(), out);
}

}  // namespace gsk
