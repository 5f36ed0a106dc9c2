// Host finite-element tables, geometry entry points and small dense helpers.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.h"

namespace idft {

namespace {

const double kBoltzmannHa = 3.166811563e-6;  // Hartree per Kelvin

// Legendre P_n and P_n' (interior points) by the three-term recurrence.
void legendre(int n, double x, double& P, double& dP) {
  double p0 = 1.0, p1 = x;
  if (n == 0) { P = 1.0; dP = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1; p1 = p2;
  }
  P = p1;
  dP = n * (x * p1 - p0) / (x * x - 1.0);
}

// Gauss-Legendre nodes (ascending) and weights by Newton iteration.
void gauss(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n); w.resize(n);
  for (int i = 0; i < n; ++i) {
    double z = -std::cos(M_PI * (i + 0.75) / (n + 0.5)), P, dP;
    for (int it = 0; it < 100; ++it) {
      legendre(n, z, P, dP);
      const double dz = P / dP;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    legendre(n, z, P, dP);
    x[i] = z; w[i] = 2.0 / ((1.0 - z * z) * dP * dP);
  }
  if (n % 2) x[n / 2] = 0.0;
}

// Gauss-Lobatto-Legendre nodes: +-1 and the roots of P_p'.
void gll(int p, std::vector<double>& x, std::vector<double>& w) {
  x.assign(p + 1, 0.0); w.assign(p + 1, 0.0);
  x[0] = -1.0; x[p] = 1.0;
  for (int i = 1; i < p; ++i) {
    double z = -std::cos(M_PI * i / p), P, dP;
    for (int it = 0; it < 100; ++it) {
      legendre(p, z, P, dP);
      const double d2P = (2.0 * z * dP - p * (p + 1.0) * P) / (1.0 - z * z);
      const double dz = dP / d2P;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = z;
  }
  if (p % 2 == 0) x[p / 2] = 0.0;
  for (int i = 0; i <= p; ++i) {
    double P = 1.0, dP;
    if (i > 0 && i < p) legendre(p, x[i], P, dP);
    w[i] = 2.0 / (p * (p + 1.0) * P * P);
  }
}

double lagrange(const std::vector<double>& xs, int a, double x) {
  double v = 1.0;
  for (size_t b = 0; b < xs.size(); ++b)
    if ((int)b != a) v *= (x - xs[b]) / (xs[a] - xs[b]);
  return v;
}

double lagrange_d(const std::vector<double>& xs, int a, double x) {
  double s = 0.0;
  for (size_t m = 0; m < xs.size(); ++m) {
    if ((int)m == a) continue;
    double t = 1.0 / (xs[a] - xs[m]);
    for (size_t b = 0; b < xs.size(); ++b)
      if ((int)b != a && b != m) t *= (x - xs[b]) / (xs[a] - xs[b]);
    s += t;
  }
  return s;
}

}  // namespace

bool desc_valid(const idft_desc* d) {
  return d && d->n_cells >= 1 && d->degree >= 2 && d->degree <= 5 && d->half_width > 0 &&
         d->n_eig >= 1 && d->n_eig <= kMaxCols &&
         d->relax_ratio >= 0.0;
}

void Tables::init(const idft_desc& d) {
  nc = d.n_cells; p = d.degree; n1 = p + 1; q1 = p + 3;
  ne = n1 * n1 * n1; nq = q1 * q1 * q1; N1 = nc * p + 1; NV1 = 2 * nc + 1;
  N = (int64_t)N1 * N1 * N1; ncell = (int64_t)nc * nc * nc; NV = (int64_t)NV1 * NV1 * NV1;
  L = d.half_width; h = 2.0 * L / nc;
  gll(p, gll_x, gll_w);
  gauss(q1, gq_x, gq_w);
  Lq.assign((size_t)q1 * n1, 0.0);
  for (int q = 0; q < q1; ++q)
    for (int a = 0; a < n1; ++a) Lq[q * n1 + a] = lagrange(gll_x, a, gq_x[q]);
  std::vector<double> ex, ew;
  gauss(p + 1, ex, ew);  // exact for the degree-2p integrands of M1 and D1
  M1.assign((size_t)n1 * n1, 0.0); D1.assign((size_t)n1 * n1, 0.0);
  for (int a = 0; a < n1; ++a)
    for (int b = 0; b < n1; ++b)
      for (size_t g = 0; g < ex.size(); ++g) {
        M1[a * n1 + b] += ew[g] * lagrange(gll_x, a, ex[g]) * lagrange(gll_x, b, ex[g]);
        D1[a * n1 + b] += ew[g] * lagrange_d(gll_x, a, ex[g]) * lagrange_d(gll_x, b, ex[g]);
      }
  // child Gauss-2 points: child cell ci covers [-1,0] (ci=0) or [0,1] (ci=1)
  const double g2 = 1.0 / std::sqrt(3.0), tl[2] = {-g2, g2};
  Lc.assign((size_t)4 * n1, 0.0);
  Wc.assign(4 * 3, 0.0);
  for (int ci = 0; ci < 2; ++ci)
    for (int k = 0; k < 2; ++k) {
      const int i = 2 * ci + k;
      const double xi = (ci == 0 ? -0.5 : 0.5) + 0.5 * tl[k];
      for (int a = 0; a < n1; ++a) Lc[i * n1 + a] = lagrange(gll_x, a, xi);
      Wc[i * 3 + ci] = 0.5 * (1.0 - tl[k]);
      Wc[i * 3 + ci + 1] = 0.5 * (1.0 + tl[k]);
    }
  Wq.assign((size_t)q1 * 3, 0.0);
  for (int q = 0; q < q1; ++q) {
    const double xi = gq_x[q];
    const int ci = xi < 0.0 ? 0 : 1;
    const double t = ci == 0 ? 2.0 * xi + 1.0 : 2.0 * xi - 1.0;
    Wq[q * 3 + ci] = 0.5 * (1.0 - t);
    Wq[q * 3 + ci + 1] = 0.5 * (1.0 + t);
  }
}

// ---- small dense helpers -----------------------------------------------------
void sym_eig(int n, std::vector<double> A, std::vector<double>& evals, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double a = A[(size_t)i * n + j] * A[(size_t)i * n + j];
        tot += a;
        if (i != j) off += a;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
  evals.assign(n, 0.0);
  std::vector<double> W((size_t)n * n);
  for (int i = 0; i < n; ++i) {
    evals[i] = A[(size_t)idx[i] * n + idx[i]];
    int big = 0;
    double bv = -1.0;
    for (int r = 0; r < n; ++r) {
      const double v = std::fabs(V[(size_t)r * n + idx[i]]);
      if (v > bv * (1.0 + 1e-12)) { bv = v; big = r; }
    }
    const double sg = V[(size_t)big * n + idx[i]] < 0 ? -1.0 : 1.0;
    for (int r = 0; r < n; ++r) W[(size_t)r * n + i] = sg * V[(size_t)r * n + idx[i]];
  }
  V.swap(W);
}

bool cholesky_lower(int n, std::vector<double>& S) {
  for (int j = 0; j < n; ++j) {
    double d = S[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= S[(size_t)j * n + k] * S[(size_t)j * n + k];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    S[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double v = S[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) v -= S[(size_t)i * n + k] * S[(size_t)j * n + k];
      S[(size_t)i * n + j] = v / d;
    }
    for (int i = 0; i < j; ++i) S[(size_t)i * n + j] = 0.0;
  }
  return true;
}

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
double rand_pm1(uint64_t& s) { return 2.0 * ((splitmix64(s) >> 11) * 0x1.0p-53) - 1.0; }

double fd_occ(double e, double mu, double T) {
  const double x = (e - mu) / (kBoltzmannHa * T);
  if (x >= 0) { const double t = std::exp(-x); return t / (1.0 + t); }
  return 1.0 / (1.0 + std::exp(x));
}
double fd_occ_der(double e, double mu, double T) {
  const double f = fd_occ(e, mu, T);
  return -f * (1.0 - f) / (kBoltzmannHa * T);
}
double fermi_level(const std::vector<double>& e, double ne, double T) {
  double lo = *std::min_element(e.begin(), e.end()) - 1.0;
  double hi = *std::max_element(e.begin(), e.end()) + 1.0;
  for (int it = 0; it < 400; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    double s = 0.0;
    for (double v : e) s += 2.0 * fd_occ(v, mid, T);
    if (s < ne) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

}  // namespace idft

// ---- geometry entry points (no context) --------------------------------------
using idft::Tables;

extern "C" {

void idft_desc_default(idft_desc* d) {
  std::memset(d, 0, sizeof(*d));
  d->n_cells = 4; d->degree = 4; d->half_width = 12.0; d->n_eig = 6; d->n_electrons = 4.0;
  d->precision = IDFT_FP64; d->temperature = 500.0; d->cheb_degree = 30;
  d->max_cheb_passes = 100; d->cheb_tol = 1e-10; d->initial_tol_cheb = 1e-6;
  d->adaptive_factor_cheb = 100.0; d->adaptive_factor_adjoint = 1000.0;
  d->adjoint_initial_tol = 1e-11; d->adjoint_max_iter = 5000; d->degeneracy_tol = 0.002;
  d->frac_occ_tol = 1e-8; d->additional_states = 0; d->block_size = 0;
  d->rho_tol_constraints = 1e-6; d->gradient_mode = IDFT_GRADIENT_REFERENCE;
  d->lanczos_steps = 12; d->psi_checkpoint = 0; d->check_every = 8; d->seed = 1;
  d->minres_graph = 1; d->relax_ratio = 1e3; d->minres_fused = 1;
}

int idft_geometry_sizes(const idft_desc* d, idft_sizes* s) {
  if (!idft::desc_valid(d) || !s) return IDFT_EINVAL;
  const int64_t N1 = (int64_t)d->n_cells * d->degree + 1, q = d->degree + 3,
                nv = 2 * (int64_t)d->n_cells + 1, n1 = d->degree + 1;
  s->n_nodes = N1 * N1 * N1;
  s->n_cells = (int64_t)d->n_cells * d->n_cells * d->n_cells;
  s->nodes_per_cell = n1 * n1 * n1;
  s->quad_per_cell = q * q * q;
  s->n_vxc = nv * nv * nv;
  s->child_quad_per_cell = 64;
  return IDFT_OK;
}

int idft_basis_tables(int32_t p, double* gx, double* gw, double* qx, double* qw) {
  if (p < 2 || p > 5 || !gx || !gw || !qx || !qw) return IDFT_EINVAL;
  idft_desc d;
  idft_desc_default(&d);
  d.degree = p;
  Tables t;
  t.init(d);
  std::copy(t.gll_x.begin(), t.gll_x.end(), gx);
  std::copy(t.gll_w.begin(), t.gll_w.end(), gw);
  std::copy(t.gq_x.begin(), t.gq_x.end(), qx);
  std::copy(t.gq_w.begin(), t.gq_w.end(), qw);
  return IDFT_OK;
}

int idft_geometry_quad_points(const idft_desc* d, double* xyz) {
  if (!idft::desc_valid(d) || !xyz) return IDFT_EINVAL;
  Tables t;
  t.init(*d);
  for (int64_t e = 0; e < t.ncell; ++e) {
    const int64_t cx = e % t.nc, cy = (e / t.nc) % t.nc, cz = e / ((int64_t)t.nc * t.nc);
    for (int qz = 0; qz < t.q1; ++qz)
      for (int qy = 0; qy < t.q1; ++qy)
        for (int qx = 0; qx < t.q1; ++qx) {
          const size_t i = (size_t)e * t.nq + (qz * t.q1 + qy) * t.q1 + qx;
          xyz[3 * i + 0] = -t.L + cx * t.h + 0.5 * t.h * (t.gq_x[qx] + 1.0);
          xyz[3 * i + 1] = -t.L + cy * t.h + 0.5 * t.h * (t.gq_x[qy] + 1.0);
          xyz[3 * i + 2] = -t.L + cz * t.h + 0.5 * t.h * (t.gq_x[qz] + 1.0);
        }
  }
  return IDFT_OK;
}

int idft_geometry_node_points(const idft_desc* d, double* xyz) {
  if (!idft::desc_valid(d) || !xyz) return IDFT_EINVAL;
  Tables t;
  t.init(*d);
  std::vector<double> x1(t.N1);
  for (int c = 0; c < t.nc; ++c)
    for (int a = 0; a < t.p; ++a) x1[c * t.p + a] = -t.L + c * t.h + 0.5 * t.h * (t.gll_x[a] + 1.0);
  x1[t.N1 - 1] = t.L;
  for (int64_t n = 0; n < t.N; ++n) {
    xyz[3 * n] = x1[n % t.N1];
    xyz[3 * n + 1] = x1[(n / t.N1) % t.N1];
    xyz[3 * n + 2] = x1[n / ((int64_t)t.N1 * t.N1)];
  }
  return IDFT_OK;
}

int idft_geometry_vxc_points(const idft_desc* d, double* xyz) {
  if (!idft::desc_valid(d) || !xyz) return IDFT_EINVAL;
  Tables t;
  t.init(*d);
  for (int64_t n = 0; n < t.NV; ++n) {
    const int64_t i = n % t.NV1, j = (n / t.NV1) % t.NV1, k = n / ((int64_t)t.NV1 * t.NV1);
    xyz[3 * n] = -t.L + 0.5 * t.h * i;
    xyz[3 * n + 1] = -t.L + 0.5 * t.h * j;
    xyz[3 * n + 2] = -t.L + 0.5 * t.h * k;
  }
  return IDFT_OK;
}

}  // extern "C"
