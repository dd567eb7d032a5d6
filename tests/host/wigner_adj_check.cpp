// Host-side check of the reverse-mode Wigner sweep (ipg_wigner_adj.h), the
// core of the FP32 adjoint force kernel.  Test infrastructure: built by
// tests/test_wigner_adj.py with hipcc and compared there against the
// oracle's compute_uarray / compute_duarray restatement.
//
// stdin, one case per line: J rfac0 rmin0 rcut switch x y z, followed by
// idxu_max seed pairs (re im) in idxu order (left-half entries are used).
// stdout, per case: dE/dr from the sweep in FP64 (3 values), the same in
// FP32 (3 values), then the forward recursion's left-half entries in FP64
// (idxu_max pairs, zero outside the left half).
#include <cmath>
#include <cstdio>
#include <vector>

#include "ipg_wigner_adj.h"

using namespace ipg::wadj;

static int ub(int j) { return j * (j + 1) * (2 * j + 1) / 6; }

// Cayley-Klein parameters, their gradients and the switching function
// (compute_ui.jl:191-216, 328-347, 350-380).
struct Geo {
  double a[4], da[4][3], sfac, dsfac, uhat[3];
};
static Geo geometry(double x, double y, double z, double rcut, double rfac0, double rmin0, int sw) {
  Geo g;
  const double rsq = x * x + y * y + z * z, r = std::sqrt(rsq);
  const double rscale0 = rfac0 * M_PI / (rcut - rmin0);
  const double theta0 = (r - rmin0) * rscale0;
  const double z0 = r * std::cos(theta0) / std::sin(theta0);
  const double r0inv = 1.0 / std::sqrt(r * r + z0 * z0);
  g.a[0] = z0 * r0inv;
  g.a[1] = -z * r0inv;
  g.a[2] = y * r0inv;
  g.a[3] = -x * r0inv;
  const double dz0dr = z0 / r - (r * rscale0) * (rsq + z0 * z0) / rsq;
  const double dr0invdr = -r0inv * r0inv * r0inv * (r + z0 * dz0dr);
  const double u[3] = {x / r, y / r, z / r};
  for (int k = 0; k < 3; k++) {
    g.uhat[k] = u[k];
    const double dr0inv = dr0invdr * u[k], dz0 = dz0dr * u[k];
    g.da[0][k] = dz0 * r0inv + z0 * dr0inv;
    g.da[1][k] = -z * dr0inv;
    g.da[2][k] = y * dr0inv;
    g.da[3][k] = -x * dr0inv;
  }
  g.da[1][2] -= r0inv;
  g.da[3][0] -= r0inv;
  g.da[2][1] += r0inv;
  if (!sw || r <= rmin0) {
    g.sfac = 1.0;
    g.dsfac = 0.0;
  } else {
    const double f = M_PI / (rcut - rmin0), ang = (r - rmin0) * f;
    g.sfac = r > rcut ? 0.0 : 0.5 * (std::cos(ang) + 1.0);
    g.dsfac = -0.5 * std::sin(ang) * f;
  }
  return g;
}

// Seeds s_e read from the case's idxu-ordered list.
template <typename T>
struct Seed {
  static constexpr bool weighted = false;
  struct S2 {
    T x, y;
  };
  const std::vector<double>& s;
  int order(T) const { return 0; }
  T tie(T v, T) const { return v; }
  template <int j, int mb, int ma>
  S2 operator()(K<j>, K<mb>, K<ma>, int) const {
    const int e = ub(j) + (j + 1) * mb + ma;
    return S2{(T)s[2 * e], (T)s[2 * e + 1]};
  }
};

template <typename T, int J>
static void dedr(const Geo& g, const std::vector<double>& seed, double out[3]) {
  AB<T> c{(T)g.a[0], (T)g.a[1], (T)g.a[2], (T)g.a[3]};
  Seed<T> sd{seed};
  Acc acc;
  vjp<T, J>(c, sd, acc);
  for (int k = 0; k < 3; k++) {
    const double grad = acc.ar * g.da[0][k] + acc.ai * g.da[1][k] + acc.br * g.da[2][k] + acc.bi * g.da[3][k];
    out[k] = g.dsfac * acc.phi * g.uhat[k] + g.sfac * grad;
  }
}

template <int J>
static void run(const Geo& g, const std::vector<double>& seed) {
  double d64[3], d32[3];
  dedr<double, J>(g, seed, d64);
  dedr<float, J>(g, seed, d32);
  std::vector<double> u(2 * ub(J + 1), 0.0);
  AB<double> c{g.a[0], g.a[1], g.a[2], g.a[3]};
  auto em = [&](auto jc, auto mbc, auto mac, double ur, double ui) {
    constexpr int j = decltype(jc)::value, mb = decltype(mbc)::value, ma = decltype(mac)::value;
    const int e = ub(j) + (j + 1) * mb + ma;
    u[2 * e] = ur;
    u[2 * e + 1] = ui;
  };
  forward<double, J>(c, em);
  std::printf("%.17g %.17g %.17g %.17g %.17g %.17g", d64[0], d64[1], d64[2], d32[0], d32[1], d32[2]);
  for (double v : u) std::printf(" %.17g", v);
  std::printf("\n");
}

int main() {
  int J, sw;
  double rfac0, rmin0, rcut, x, y, z;
  while (std::scanf("%d %lf %lf %lf %d %lf %lf %lf", &J, &rfac0, &rmin0, &rcut, &sw, &x, &y, &z) == 8) {
    std::vector<double> seed(2 * ub(J + 1));
    for (double& s : seed)
      if (std::scanf("%lf", &s) != 1) return 2;
    const Geo g = geometry(x, y, z, rcut, rfac0, rmin0, sw);
    switch (J) {
      case 0: run<0>(g, seed); break;
      case 1: run<1>(g, seed); break;
      case 2: run<2>(g, seed); break;
      case 3: run<3>(g, seed); break;
      case 4: run<4>(g, seed); break;
      case 5: run<5>(g, seed); break;
      case 6: run<6>(g, seed); break;
      case 7: run<7>(g, seed); break;
      case 8: run<8>(g, seed); break;
      default: return 3;
    }
  }
  return 0;
}
