"""Per-point functions the kernels use (gsis-sim_amd/csrc/kernels/kinetic_math.hpp), compiled on the
host (the header is plain C++ outside hipcc) and checked against what the mathematics and the
reference's own tests fix:

- the Venkatakrishnan limiter's known values (tests/test_physics.cpp:174-200): 1 at a zero
  projection, 0 at a zero-width extremum without epsilon, 3/4 when the projection equals the
  neighbour maximum (either side, FP64 and FP32), above 0.99 for a large epsilon;
- the kernels' rewrite of phi = min over faces of venkat(proj_k) as min(venkat(max proj),
  venkat(min proj)) (limit_from_extremes in kinetic_view.hpp): bit-exact on 200k random face sets
  per precision, and within 4 ulp when the projections lie a few ulps apart, where the floating-
  point limiter is not monotone (measured up to 2.6 ulp);
- the Shakhov target (make_cell_params + target_eval) against its closed-form moments, and the
  cold-BGK switch;
- the curved Maxwell wall's reflected value (reflected_value in kinetic_view.hpp) against the closed
  form it must reproduce when f / M_eq is linear in the velocity, on the stretched and uniform grids."""

from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KERNELS = os.path.join(os.path.dirname(HERE), "gsis-sim_amd", "csrc", "kernels")

HARNESS = r"""
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include "kinetic_math.hpp"
using namespace gsk::dev;

template <class F>
void trial(long long n, unsigned seed, bool clustered, long long& bad, double& worst) {
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  bad = 0;
  worst = 0.0;
  for (long long t = 0; t < n; ++t) {
    const F dmax = F(std::pow(10.0, -3 + 4 * u(g))), dmin = -F(std::pow(10.0, -3 + 4 * u(g)));
    const F eps2 = u(g) < 0.25 ? F(0) : F(std::pow(10.0, -8 + 8 * u(g)));
    const int nf = clustered ? 6 : 1 + int(6 * u(g));
    const F base = F((u(g) < 0.5 ? -1 : 1) * std::pow(10.0, -4 + 5 * u(g)));
    F mn = F(1), pmax = F(0), pmin = F(0), p = base;
    for (int k = 0; k < nf; ++k) {
      if (!clustered) p = u(g) < 0.1 ? F(0) : F((u(g) < 0.5 ? -1 : 1) * std::pow(10.0, -4 + 5 * u(g)));
      const F v = venkat_face(p, dmax, dmin, eps2);
      mn = v < mn ? v : mn;
      pmax = pmax < p ? p : pmax;
      pmin = p < pmin ? p : pmin;
      if (clustered) p = std::nextafter(p, base > 0 ? F(INFINITY) : F(-INFINITY));
    }
    const F a = venkat_face(pmax, dmax, dmin, eps2), b = venkat_face(pmin, dmax, dmin, eps2);
    const F two = b < a ? b : a;
    if (std::memcmp(&two, &mn, sizeof(F)) != 0) {
      ++bad;
      const double r = std::fabs(double(two) - double(mn)) / std::fabs(double(mn));
      worst = r > worst ? r : worst;
    }
  }
}

int main() {
  std::printf("%.17g %.17g %.17g %.17g %.17g %.9g\n", venkat_face(0.0, 1.0, -1.0, 0.0), venkat_face(0.5, 0.0, -1.0, 0.0),
              venkat_face(1.0, 1.0, -1.0, 0.0), venkat_face(-1.0, 1.0, -1.0, 0.0), venkat_face(1.0, 1.0, -1.0, 100.0),
              double(venkat_face(1.0f, 1.0f, -1.0f, 0.0f)));
  long long b;
  double w;
  trial<double>(200000, 1, false, b, w);
  std::printf("%lld %.17g\n", b, w);
  trial<float>(200000, 2, false, b, w);
  std::printf("%lld %.17g\n", b, w);
  trial<double>(200000, 3, true, b, w);
  std::printf("%lld %.17g\n", b, w);
  trial<float>(200000, 4, true, b, w);
  std::printf("%lld %.17g\n", b, w);
}
"""


def test_limiter_values_and_two_extreme_rewrite(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available to compile the limiter harness")
    src = tmp_path / "lim.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "lim"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-I{KERNELS}", str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    v = [float(x) for x in lines[0].split()]
    assert v[0] == 1.0 and v[1] == 0.0
    assert v[2] == pytest.approx(0.75, abs=1e-15) and v[3] == pytest.approx(0.75, abs=1e-15)
    assert v[4] > 0.99 and v[5] == pytest.approx(0.75, abs=1e-7)
    sep64, sep32, cl64, cl32 = ([float(x) for x in ln.split()] for ln in lines[1:5])
    assert sep64[0] == 0 and sep32[0] == 0  # separated projections: bit for bit
    assert cl64[1] <= 4 * 2.0 ** -52 and cl32[1] <= 4 * 2.0 ** -23  # ulp-clustered: rounding only


SHAKHOV = r"""
#include <cmath>
#include <cstdio>
#include "kinetic_math.hpp"
using namespace gsk::dev;

int main() {
  // Shakhov target of (rho, u, T, q) with Pr = 2/3 on 48^3 points over [-9, 9]^3, moments in FP64
  double m[kMoments] = {};
  m[0] = 1.2; m[1] = 0.3; m[2] = -0.2; m[3] = 0.5; m[4] = 0.9; m[5] = m[0] * m[4];
  m[12] = 0.08; m[13] = -0.05; m[14] = 0.02;
  const CellP<double> cp = make_cell_params<double>(m, 1.0, 0.0, 1e-3, 0.66, 2.0 / 3.0, 0.0);
  const int n = 48;
  const double h = 18.0 / n, w = h * h * h;
  double s0 = 0, s1[3] = {0, 0, 0}, s2[3][3] = {{0}}, s3[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) {
        const double x[3] = {-9 + (i + 0.5) * h, -9 + (j + 0.5) * h, -9 + (k + 0.5) * h};
        const double f = target_eval(cp.s, x[0], x[1], x[2]) * w;
        s0 += f;
        for (int a = 0; a < 3; ++a) s1[a] += f * x[a];
      }
  const double u[3] = {s1[0] / s0, s1[1] / s0, s1[2] / s0};
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) {
        const double x[3] = {-9 + (i + 0.5) * h, -9 + (j + 0.5) * h, -9 + (k + 0.5) * h};
        const double f = target_eval(cp.s, x[0], x[1], x[2]) * w;
        const double c[3] = {x[0] - u[0], x[1] - u[1], x[2] - u[2]};
        const double c2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        for (int a = 0; a < 3; ++a) {
          s3[a] += 0.5 * f * c[a] * c2;
          for (int b = 0; b < 3; ++b) s2[a][b] += f * c[a] * c[b];
        }
      }
  const double T = (s2[0][0] + s2[1][1] + s2[2][2]) / (3.0 * s0);
  std::printf("%.17g %.17g %.17g %.17g %.17g\n", s0, u[0], u[1], u[2], T);
  const double p = s0 * T;
  std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", s2[0][0] - p, s2[1][1] - p, s2[2][2] - p, s2[0][1], s2[1][2], s2[0][2]);
  std::printf("%.17g %.17g %.17g\n", s3[0], s3[1], s3[2]);
  // q = 0: the target is the Maxwellian
  double m0[kMoments] = {};
  m0[0] = 1.2; m0[1] = 0.3; m0[2] = -0.2; m0[3] = 0.5; m0[4] = 0.9; m0[5] = m0[0] * m0[4];
  const CellP<double> c0 = make_cell_params<double>(m0, 1.0, 0.0, 1e-3, 0.66, 2.0 / 3.0, 0.0);
  const double xi[3] = {0.7, -1.1, 0.4};
  std::printf("%.17g %.17g\n", target_eval(c0.s, xi[0], xi[1], xi[2]), maxw_eval(c0.s.eq, xi[0], xi[1], xi[2]));
  // cold_bgk_temperature (tests/kinetic/test_kinetic_phase5.cpp:373-388)
  double mc[kMoments] = {};
  mc[0] = 0.02; mc[1] = 2.0; mc[4] = 0.05; mc[5] = mc[0] * mc[4]; mc[12] = 1e-4;
  const CellP<double> warm = make_cell_params<double>(mc, 1.0, 0.0, 1e-3, 0.66, 2.0 / 3.0, 0.0);
  const CellP<double> cold = make_cell_params<double>(mc, 1.0, 0.0, 1e-3, 0.66, 2.0 / 3.0, 0.1);
  const CellP<double> above = make_cell_params<double>(mc, 1.0, 0.0, 1e-3, 0.66, 2.0 / 3.0, 0.04);
  std::printf("%.17g %.17g %.17g %.17g %.17g\n", warm.s.coef, cold.s.coef, above.s.coef, cold.coll, warm.coll);
}
"""


def test_shakhov_target_moments_and_cold_bgk_switch(tmp_path):
    """The target the kernels evaluate (make_cell_params + target_eval): on 48^3 points over [-9, 9]^3
    its moments are (rho, u, T) with no stress and heat flux (1 - Pr) q (1e-10); with q = 0 it is the
    Maxwellian; below cold_bgk_temperature its Shakhov coefficient is 0 and the collision frequency
    unchanged (tests/test_physics.cpp:126-145, tests/kinetic/test_kinetic_phase5.cpp:373-388)."""
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available to compile the harness")
    src = tmp_path / "shakhov.cpp"
    src.write_text(SHAKHOV)
    exe = tmp_path / "shakhov"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-I{KERNELS}", str(src), "-o", str(exe)], check=True)
    rows = [[float(x) for x in ln.split()] for ln in
            subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.strip().split("\n")]
    rho, ux, uy, uz, T = rows[0]
    assert abs(rho / 1.2 - 1) < 1e-10 and abs(T / 0.9 - 1) < 1e-10
    assert max(abs(ux - 0.3), abs(uy + 0.2), abs(uz - 0.5)) < 1e-10
    assert max(abs(x) for x in rows[1]) < 1e-10
    q = (1 - 2.0 / 3.0) * np.array([0.08, -0.05, 0.02])
    assert np.max(np.abs(np.array(rows[2]) - q)) < 1e-10
    assert rows[3][0] == pytest.approx(rows[3][1], rel=1e-15)
    warm, cold, above, coll_cold, coll_warm = rows[4]
    assert warm > 0.0 and cold == 0.0 and above == warm and coll_cold == coll_warm


REFLECTION = r"""
#include <cmath>
#include <cstdio>
#include <vector>
#define __device__
#define __forceinline__ inline
#include "kinetic_view.hpp"
using namespace gsk::dev;

// g = f / M_eq is linear in the velocity, so the kernel's trilinear interpolation of g between the
// bracketing nodes must return M_eq(r) g(r) at the reflected velocity r, up to rounding.
static double lin(const double* r) { return 3.0 + 0.3 * r[0] - 0.2 * r[1] + 0.25 * r[2]; }

template <class F>
void run(bool stretched) {
  const int n[3] = {24, 20, 16};
  const double h = 6.0, b = 1.5;
  std::vector<double> nodes;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < n[a]; ++i) {
      const double s = -1.0 + (i + 0.5) * 2.0 / n[a];
      nodes.push_back(stretched ? (h / b) * std::atanh(s * std::tanh(b)) : -h + (i + 0.5) * 2.0 * h / n[a]);
    }
  const int nv = n[0] * n[1] * n[2];
  std::vector<F> xs(nv), ys(nv), zs(nv), f(nv);
  std::vector<int> g2s(nv);
  const double ux = 0.3, uy = -0.2, uz = 0.1, B = 1.0 / (2.0 * 1.1);
  double scale = 0.0;
  for (int iz = 0; iz < n[2]; ++iz)
    for (int iy = 0; iy < n[1]; ++iy)
      for (int ix = 0; ix < n[0]; ++ix) {
        const int g = ix + n[0] * (iy + n[1] * iz);
        const int q = nv - 1 - g;  // storage order reversed, so the grid-to-storage map is exercised
        g2s[g] = q;
        xs[q] = F(nodes[ix]);
        ys[q] = F(nodes[n[0] + iy]);
        zs[q] = F(nodes[n[0] + n[1] + iz]);
        const double r[3] = {double(xs[q]), double(ys[q]), double(zs[q])};
        const double c2 = (r[0] - ux) * (r[0] - ux) + (r[1] - uy) * (r[1] - uy) + (r[2] - uz) * (r[2] - uz);
        f[q] = F(std::exp(-B * c2) * lin(r));
        scale = std::fmax(scale, std::fabs(double(f[q])));
      }
  CellP<F> cell{};
  cell.s.eq = Maxw<F>{F(ux), F(uy), F(uz), F(1), F(B)};
  View<F> v{};
  v.nv = nv;
  v.xi_x = xs.data();
  v.xi_y = ys.data();
  v.xi_z = zs.data();
  v.nodes = stretched ? nodes.data() : nullptr;
  v.cells = &cell;
  v.f = f.data();
  v.grid_to_storage = g2s.data();
  for (int a = 0; a < 3; ++a) {
    v.grid_n[a] = n[a];
    v.deq.lo[a] = -h;
    v.deq.d[a] = 2.0 * h / n[a];
  }
  double worst = 0.0;
  long inside = 0, outside = 0, outside_nonzero = 0, hit_nodes = 0;
  unsigned long long state = 12345;
  for (int k = 0; k < 40; ++k) {
    double nrm[3];
    if (k == 0) {
      nrm[0] = nrm[1] = 1.0 / std::sqrt(2.0), nrm[2] = 0.0;  // x <-> y swap: nodes onto nodes (x, y share no count)
    } else {
      double l = 0.0;
      for (double& c : nrm) {
        state = state * 6364136223846793005ULL + 1442695040888963407ULL;
        c = double(state >> 11) / 9007199254740992.0 - 0.5;
        l += c * c;
      }
      for (double& c : nrm) c /= std::sqrt(l);
    }
    BFace<F> bf{};
    bf.nx = F(nrm[0]);
    bf.ny = F(nrm[1]);
    bf.nz = F(nrm[2]);
    for (int p = 0; p < nv; ++p) {
      const F x = xs[p], y = ys[p], z = zs[p];
      if (!(x * bf.nx + y * bf.ny + z * bf.nz < F(0))) continue;
      const F got = reflected_value<false, 0, F>(v, bf, p, x, y, z);
      const double un = double(x) * double(bf.nx) + double(y) * double(bf.ny) + double(z) * double(bf.nz);
      const double r[3] = {double(x) - 2.0 * un * double(bf.nx), double(y) - 2.0 * un * double(bf.ny),
                           double(z) - 2.0 * un * double(bf.nz)};
      int where = 0;  // 0 inside, 1 clearly outside, 2 within 1e-3 of the end spacing of the edge
      int off = 0;
      for (int a = 0; a < 3; ++a) {
        const double* nd = nodes.data() + off;
        const double lo = nd[0], hi = nd[n[a] - 1], d = nd[1] - nd[0];
        off += n[a];
        if (r[a] < lo - 1e-3 * d || r[a] > hi + 1e-3 * d)
          where = 1;
        else if (where == 0 && (r[a] < lo + 1e-3 * d || r[a] > hi - 1e-3 * d))
          where = 2;
      }
      if (where == 1) {
        ++outside;
        if (got != F(0)) ++outside_nonzero;
        continue;
      }
      if (where == 2) continue;
      ++inside;
      const double c2 = (r[0] - ux) * (r[0] - ux) + (r[1] - uy) * (r[1] - uy) + (r[2] - uz) * (r[2] - uz);
      const double expect = std::exp(-B * c2) * lin(r);
      worst = std::fmax(worst, std::fabs(double(got) - expect) / scale);
      if (k == 0) ++hit_nodes;
    }
  }
  std::printf("%.6g %ld %ld %ld %ld\n", worst, inside, outside, outside_nonzero, hit_nodes);
}

int main() {
  run<double>(true);
  run<double>(false);
  run<float>(true);
  run<float>(false);
}
"""


def test_maxwell_wall_reflection_interpolates_exactly_for_linear_ratio(tmp_path):
    """The curved Maxwell wall's reflected value (reflected_value in kinetic_view.hpp, the function
    the reflection kernel runs; kinetic_kernels.hpp:80-127 on the uniform grid) interpolates
    g = f / M_eq trilinearly between the bracketing nodes and multiplies by M_eq at the reflected
    velocity r. For f = M_eq times a linear function that is exact, so the result must be
    M_eq(r) g(r): on the tanh-stretched grid (24 x 20 x 16 nodes on [-6, 6], stretch 1.5, where
    the bracketing is a bisection over the node table) and on the uniform grid, for 40 wall normals
    including the x-y diagonal, with storage order reversed against grid order; and every reflected
    velocity outside the node range gives 0. A wrong bracket, axis, weight or Maxwellian ratio
    breaks the equality by far more than rounding."""
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available to compile the harness")
    src = tmp_path / "reflect.cpp"
    src.write_text(REFLECTION)
    exe = tmp_path / "reflect"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-I{KERNELS}", str(src), "-o", str(exe)], check=True)
    rows = [[float(x) for x in ln.split()] for ln in
            subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.strip().split("\n")]
    for (worst, inside, outside, outside_nonzero, diagonal), bound in zip(rows, (1e-13, 1e-13, 2e-6, 2e-6)):
        assert worst < bound, (worst, bound)
        assert inside > 100000 and outside > 1000 and outside_nonzero == 0 and diagonal > 1000


PRODUCT_SUMS = r"""
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
#include "kinetic_math.hpp"
using namespace gsk::dev;

// The discrete-equilibrium sums in product form (deq_axis_terms, deq_product_newton_sums,
// deq_product_shakhov_sums, with the axis tables built as the HIP backend builds them) against
// the point sums of deq_terms over every point of a tensor grid stored in a shuffled order.
int main() {
  const int n[3] = {24, 20, 16};
  const double lo[3] = {-6.0, -5.0, -7.0}, hi[3] = {6.0, 7.0, 5.0};
  std::mt19937_64 g(7);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  for (int stretched = 0; stretched < 2; ++stretched) {
    // nodes and per-axis weights (the tanh-stretched midpoint rule, or the uniform one)
    std::vector<double> node[3], aw[3];
    double weight = 1.0;
    for (int a = 0; a < 3; ++a) {
      const double h = 0.5 * (hi[a] - lo[a]), c = 0.5 * (hi[a] + lo[a]), b = 1.5, tb = std::tanh(b);
      weight *= (hi[a] - lo[a]) / n[a];
      for (int i = 0; i < n[a]; ++i) {
        const double s = (2.0 * i + 1.0 - n[a]) / n[a];
        if (!stretched) {
          node[a].push_back(c + s * h);
          aw[a].push_back((hi[a] - lo[a]) / n[a]);
        } else {
          node[a].push_back(c + (h / b) * std::atanh(s * tb));
          aw[a].push_back((2.0 / n[a]) * (h / b) * tb / (1.0 - s * tb * s * tb));
        }
      }
    }
    // points in a shuffled storage order, each with its relative weight
    const int nv = n[0] * n[1] * n[2];
    std::vector<int> g2s(nv);
    for (int k = 0; k < nv; ++k) g2s[k] = k;
    std::shuffle(g2s.begin(), g2s.end(), g);
    std::vector<double> xs(nv), ys(nv), zs(nv), wr(nv);
    for (int l = 0; l < n[2]; ++l)
      for (int j = 0; j < n[1]; ++j)
        for (int i = 0; i < n[0]; ++i) {
          const int p = g2s[i + n[0] * (j + n[1] * l)];
          xs[p] = node[0][i];
          ys[p] = node[1][j];
          zs[p] = node[2][l];
          wr[p] = aw[0][i] * aw[1][j] * aw[2][l] / weight;
        }
    // the backend's axis tables: the points on the three axes through node (0, 0, 0)
    auto point = [&](int i, int j, int l) { return g2s[i + n[0] * (j + n[1] * l)]; };
    std::vector<double> an, af;
    for (int i = 0; i < n[0]; ++i) { an.push_back(xs[point(i, 0, 0)]); af.push_back(wr[point(i, 0, 0)]); }
    for (int j = 0; j < n[1]; ++j) { an.push_back(ys[point(0, j, 0)]); af.push_back(wr[point(0, j, 0)]); }
    for (int l = 0; l < n[2]; ++l) { an.push_back(zs[point(0, 0, l)]); af.push_back(wr[point(0, 0, l)]); }
    const double w0 = wr[point(0, 0, 0)], wscale = stretched ? 1.0 / (w0 * w0) : 1.0;
    double worst = 0.0;
    for (int trial = 0; trial < 40; ++trial) {
      const double A = 0.2 + U(g), B = 0.3 + 1.5 * U(g);
      const double up[3] = {U(g) - 0.5, U(g) - 0.5, U(g) - 0.5}, u[3] = {U(g) - 0.5, U(g) - 0.5, U(g) - 0.5};
      Target<double> t{};
      t.coef = 0.1 + U(g);
      t.qx = U(g) - 0.5;
      t.qy = U(g) - 0.5;
      t.qz = U(g) - 0.5;
      t.theta = 0.5 + U(g);
      double point_sum[kDeqSums] = {}, magnitude[kDeqSums] = {};
      for (int p = 0; p < nv; ++p) {
        const double cx = xs[p] - u[0], cy = ys[p] - u[1], cz = zs[p] - u[2];
        const double dx = xs[p] - up[0], dy = ys[p] - up[1], dz = zs[p] - up[2];
        const double e = A * std::exp(-B * (dx * dx + dy * dy + dz * dz)) * (stretched ? wr[p] : 1.0);
        const double pm = shakhov_poly(t, cx, cy, cz, cx * cx + cy * cy + cz * cz);
        double terms[kDeqSums];
        deq_terms<double>(e, pm, cx, cy, cz, terms);
        for (int k = 0; k < kDeqSums; ++k) {
          point_sum[k] += terms[k];
          magnitude[k] += std::fabs(terms[k]);
        }
      }
      double mu[3][kAxisPowers] = {};
      int off = 0;
      for (int a = 0; a < 3; ++a) {
        for (int i = 0; i < n[a]; ++i) {
          double terms[kAxisPowers];
          deq_axis_terms(an[off + i], up[a], u[a], B, stretched ? af[off + i] : 1.0, terms);
          for (int k = 0; k < kAxisPowers; ++k) mu[a][k] += terms[k];
        }
        off += n[a];
      }
      double product_sum[kDeqSums];
      const double q[3] = {t.qx, t.qy, t.qz};
      deq_product_newton_sums(mu, A * wscale, product_sum);
      deq_product_shakhov_sums(mu, A * wscale, t.coef, q, t.theta, product_sum);
      for (int k = 0; k < kDeqSums; ++k)
        worst = std::fmax(worst, std::fabs(product_sum[k] - point_sum[k]) / magnitude[k]);
    }
    std::printf("%.17g\n", worst);
  }
}
"""


def test_discrete_equilibrium_sums_in_product_form(tmp_path):
    """The discrete-equilibrium Newton and Shakhov-correction sums the cell-parameter kernel takes in
    product form (one-axis sums of E_a c_a^k multiplied per monomial) equal the point sums of
    deq_terms over every point of a 24 x 20 x 16 tensor grid (shuffled storage order, asymmetric
    ranges), on the uniform grid and on the tanh-stretched one with the backend's factorisation of
    the relative weights, for 40 random trial Maxwellians and Shakhov targets each: within 1e-13 of
    the sum of the terms' magnitudes. A monomial with a wrong power or axis misses by O(1)."""
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available to compile the harness")
    src = tmp_path / "product.cpp"
    src.write_text(PRODUCT_SUMS.replace("#include <random>", "#include <random>\n#include <algorithm>"))
    exe = tmp_path / "product"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-I{KERNELS}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    uniform, stretched = (float(x) for x in out)
    assert uniform < 1e-13 and stretched < 1e-13, (uniform, stretched)
