// 1D tables of the sum-factorised cell operator as compile-time constants.
//
// For degree p they depend on p alone: the GLL nodes x_a (the nodal basis),
// the p+3 Gauss-Legendre points g_q (the cell quadrature) and the p+1 Gauss
// points that integrate the 1D mass and stiffness exactly.  Computed here by
// bisection on the Legendre recurrence (constexpr, host and device), so that
// device code sees every entry as an immediate after unrolling; the host FE
// tables (fe.cpp, Newton iteration) agree with them to rounding, which the
// context checks when it is created.
#ifndef IDFT_SF_TABLES_H
#define IDFT_SF_TABLES_H

#include <hip/hip_runtime.h>

namespace idft {

struct LegendreVal {
  double P, dP;
};

// P_n(x) and P_n'(x) for |x| < 1.
__host__ __device__ constexpr LegendreVal legendre_c(int n, double x) {
  if (n == 0) return {1.0, 0.0};
  double p0 = 1.0, p1 = x;
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  return {p1, n * (x * p1 - p0) / (x * x - 1.0)};
}

// Roots in (-1, 1) of P_n (deriv = false) or P_n' (deriv = true), ascending:
// sign changes on an odd grid (so x = 0 is never a grid point), then
// bisection to the last representable midpoint.
template <int NR>
struct Roots {
  double x[NR > 0 ? NR : 1];
};
template <int NR>
__host__ __device__ constexpr Roots<NR> legendre_roots(int n, bool deriv) {
  Roots<NR> r{};
  constexpr int G = 1023;
  int found = 0;
  double xa = -1.0 + 1e-12;
  double fa = deriv ? legendre_c(n, xa).dP : legendre_c(n, xa).P;
  for (int i = 1; i <= G && found < NR; ++i) {
    const double xb = i == G ? 1.0 - 1e-12 : -1.0 + 2.0 * i / G;
    const double fb = deriv ? legendre_c(n, xb).dP : legendre_c(n, xb).P;
    if ((fa < 0.0) != (fb < 0.0)) {
      double a = xa, b = xb, f0 = fa;
      for (int it = 0; it < 200; ++it) {
        const double m = 0.5 * (a + b);
        if (m == a || m == b) break;
        const double fm = deriv ? legendre_c(n, m).dP : legendre_c(n, m).P;
        if ((fm < 0.0) == (f0 < 0.0)) {
          a = m;
          f0 = fm;
        } else {
          b = m;
        }
      }
      r.x[found++] = 0.5 * (a + b);
    }
    xa = xb;
    fa = fb;
  }
  if (NR % 2 == 1) r.x[NR / 2] = 0.0;  // odd count: the middle root is 0
  return r;
}

template <int P>
struct SfConst {
  static constexpr int N1 = P + 1, Q = P + 3;
  double L[Q * N1];   // L[q][a] = l_a(g_q) at the p+3 Gauss points
  double M[N1 * N1];  // exact 1D mass on [-1, 1]
  double D[N1 * N1];  // exact 1D stiffness on [-1, 1]
};

template <int P>
__host__ __device__ constexpr SfConst<P> make_sf_const() {
  constexpr int N1 = P + 1, Q = P + 3;
  SfConst<P> t{};
  double xn[N1] = {};
  xn[0] = -1.0;
  xn[P] = 1.0;
  const Roots<P - 1> in = legendre_roots<P - 1>(P, true);
  for (int i = 1; i < P; ++i) xn[i] = in.x[i - 1];
  if (P % 2 == 0) xn[P / 2] = 0.0;
  auto lag = [&](int a, double x) {
    double v = 1.0;
    for (int b = 0; b < N1; ++b)
      if (b != a) v *= (x - xn[b]) / (xn[a] - xn[b]);
    return v;
  };
  auto lag_d = [&](int a, double x) {
    double s = 0.0;
    for (int m = 0; m < N1; ++m) {
      if (m == a) continue;
      double v = 1.0 / (xn[a] - xn[m]);
      for (int b = 0; b < N1; ++b)
        if (b != a && b != m) v *= (x - xn[b]) / (xn[a] - xn[b]);
      s += v;
    }
    return s;
  };
  const Roots<Q> gq = legendre_roots<Q>(Q, false);
  for (int q = 0; q < Q; ++q)
    for (int a = 0; a < N1; ++a) t.L[q * N1 + a] = lag(a, gq.x[q]);
  const Roots<P + 1> ge = legendre_roots<P + 1>(P + 1, false);
  for (int g = 0; g < P + 1; ++g) {
    const double x = ge.x[g];
    const double d = legendre_c(P + 1, x).dP;
    const double w = 2.0 / ((1.0 - x * x) * d * d);
    for (int a = 0; a < N1; ++a)
      for (int b = 0; b < N1; ++b) {
        t.M[a * N1 + b] += w * lag(a, x) * lag(b, x);
        t.D[a * N1 + b] += w * lag_d(a, x) * lag_d(b, x);
      }
  }
  return t;
}

template <int P>
struct SfC {
  static constexpr SfConst<P> v = make_sf_const<P>();
};

}  // namespace idft

#endif  // IDFT_SF_TABLES_H
