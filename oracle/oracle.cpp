/*
 * oracle.cpp -- FP64 CPU restatement of invDFT's per-evaluation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it.  The
 * product (invdft_amd/_lib/libidft.so) never links, loads or calls it.
 *
 * It implements include/idft.h with the arithmetic forms of the reference:
 *   - stored per-cell Hamiltonians H_e = K_e/2 + V_e applied by batched cell
 *     GEMMs (gather -> H_e X_e -> scatter-add), Loewdin-scaled by M^{-1/2};
 *   - the adjoint right-hand side with stored R_e = int N_i N_j u JxW
 *     (MultiVectorAdjointLinearSolverProblem.cpp:309-366), mu from quadrature
 *     values reduced cell-wise then over cells (.cpp:368-402, 912-986),
 *     rhs = -2 Mu^T (M^{1/2} psi) + 4 f M^{-1/2} R psi (.cpp:614-812);
 *   - vmult / Jacobi preconditioners (.cpp:540-612, 1345-1367);
 *   - distributeX with the eta correction and the null-space update
 *     (.cpp:1002-1141, 1143-1243, 246-305);
 *   - getForceVector / solveEigen (InverseDFTSolverFunction.cpp:657-1477),
 *     force integration on child Gauss points (.cpp:1609-1674);
 *   - the L-BFGS driver (BFGSInverseDFTSolver.cpp:66-627).
 * Parts that live in DFT-FE (not in the reference repository) are defined
 * here explicitly: the preconditioned MINRES recurrence (Paige-Saunders),
 * the ChFSI pass (Lanczos bound, Zhou-Saad scaled Chebyshev filter,
 * Cholesky-QR, Rayleigh-Ritz), the Fermi level (bisection on the electron
 * count) and the GLL/Gauss finite-element tables.
 */
#include "../include/idft.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

using std::vector;

namespace {

const double kBoltzmannHa = 3.166811563e-6;  // Hartree per Kelvin

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------- 1D tables
// Legendre P_n(x) and its derivative by the three-term recurrence.
void legendre(int n, double x, double& P, double& dP) {
  double p0 = 1.0, p1 = x;
  if (n == 0) { P = 1.0; dP = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1; p1 = p2;
  }
  P = p1;
  // P_n'(x) = n (x P_n - P_{n-1}) / (x^2 - 1), interior points only
  dP = n * (x * p1 - p0) / (x * x - 1.0);
}

void gauss_rule(int n, vector<double>& x, vector<double>& w) {
  x.assign(n, 0.0); w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double P, dP; legendre(n, z, P, dP);
      double dz = P / dP; z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double P, dP; legendre(n, z, P, dP);
    x[i] = z; w[i] = 2.0 / ((1.0 - z * z) * dP * dP);
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

void gll_rule(int p, vector<double>& x, vector<double>& w) {
  x.assign(p + 1, 0.0); w.assign(p + 1, 0.0);
  x[0] = -1.0; x[p] = 1.0;
  for (int i = 1; i < p; ++i) {
    double z = -std::cos(M_PI * i / p);
    for (int it = 0; it < 100; ++it) {
      double P, dP; legendre(p, z, P, dP);
      double d2P = (2.0 * z * dP - p * (p + 1.0) * P) / (1.0 - z * z);
      double dz = dP / d2P; z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = z;
  }
  if (p % 2 == 0) x[p / 2] = 0.0;
  for (int i = 0; i <= p; ++i) {
    double P, dP;
    if (i == 0 || i == p) P = (i == 0 && p % 2 == 1) ? -1.0 : 1.0;
    else legendre(p, x[i], P, dP);
    w[i] = 2.0 / (p * (p + 1.0) * P * P);
  }
}

double lagrange(const vector<double>& nodes, int a, double x) {
  double v = 1.0;
  for (size_t b = 0; b < nodes.size(); ++b)
    if ((int)b != a) v *= (x - nodes[b]) / (nodes[a] - nodes[b]);
  return v;
}

double lagrange_d(const vector<double>& nodes, int a, double x) {
  double s = 0.0;
  for (size_t m = 0; m < nodes.size(); ++m) {
    if ((int)m == a) continue;
    double t = 1.0 / (nodes[a] - nodes[m]);
    for (size_t b = 0; b < nodes.size(); ++b)
      if ((int)b != a && b != m) t *= (x - nodes[b]) / (nodes[a] - nodes[b]);
    s += t;
  }
  return s;
}

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
double rand_pm1(uint64_t& s) { return 2.0 * ((splitmix64(s) >> 11) * 0x1.0p-53) - 1.0; }

// ------------------------------------------------------------ small dense
// Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major).
// Eigenvalues ascending; column i of V (V[r*n+i]) is the eigenvector, signed
// so that its largest-magnitude entry is positive.
void sym_eig(int n, vector<double> A, vector<double>& evals, vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double a = A[(size_t)i * n + j] * A[(size_t)i * n + j];
        tot += a; if (i != j) off += a;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[(size_t)p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    return A[(size_t)a * n + a] < A[(size_t)b * n + b];
  });
  evals.assign(n, 0.0);
  vector<double> W((size_t)n * n);
  for (int i = 0; i < n; ++i) {
    evals[i] = A[(size_t)idx[i] * n + idx[i]];
    int big = 0; double bv = -1.0;
    for (int r = 0; r < n; ++r) {
      double v = std::fabs(V[(size_t)r * n + idx[i]]);
      if (v > bv * (1.0 + 1e-12)) { bv = v; big = r; }
    }
    double sg = V[(size_t)big * n + idx[i]] < 0 ? -1.0 : 1.0;
    for (int r = 0; r < n; ++r) W[(size_t)r * n + i] = sg * V[(size_t)r * n + idx[i]];
  }
  V.swap(W);
}

// Lower Cholesky in place; returns false on a non-positive pivot.
bool cholesky(int n, vector<double>& S) {
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

// ------------------------------------------------------------------ mesh
struct Mesh {
  int nc = 0, p = 0, n1 = 0, q1 = 0, ne = 0, nq = 0, N1 = 0, NV1 = 0;
  long N = 0, ncell = 0, NV = 0;
  double L = 0, h = 0;
  vector<double> gx, gw, qx, qw;
  vector<double> Lq, M1, D1;     // [q1][n1], [n1][n1], [n1][n1]
  vector<double> Nq;             // [nq][ne] shape values at parent quadrature
  vector<double> JxW;            // [nq]
  vector<double> Ke;             // [ne][ne]
  vector<long> cellNodes;        // [ncell][ne]
  vector<double> Lc;             // [4][n1] parent basis at child Gauss points
  double cw[4][2];               // child linear basis at the 4 child points / axis
  int ccell[4];                  // child cell (0/1) of each child point / axis
  // v_xc (child Q1 field) -> parent quadrature, per 1D parent point
  vector<int> qcc; vector<double> qlw0, qlw1;

  void init(const idft_desc& d) {
    nc = d.n_cells; p = d.degree; n1 = p + 1; q1 = p + 3;
    ne = n1 * n1 * n1; nq = q1 * q1 * q1; N1 = nc * p + 1;
    N = (long)N1 * N1 * N1; ncell = (long)nc * nc * nc;
    NV1 = 2 * nc + 1; NV = (long)NV1 * NV1 * NV1;
    L = d.half_width; h = 2.0 * L / nc;
    gll_rule(p, gx, gw);
    gauss_rule(q1, qx, qw);
    Lq.assign((size_t)q1 * n1, 0.0);
    for (int q = 0; q < q1; ++q)
      for (int a = 0; a < n1; ++a) Lq[(size_t)q * n1 + a] = lagrange(gx, a, qx[q]);
    vector<double> ex, ew; gauss_rule(p + 1, ex, ew);
    M1.assign((size_t)n1 * n1, 0.0); D1.assign((size_t)n1 * n1, 0.0);
    for (int a = 0; a < n1; ++a)
      for (int b = 0; b < n1; ++b)
        for (size_t g = 0; g < ex.size(); ++g) {
          M1[(size_t)a * n1 + b] += ew[g] * lagrange(gx, a, ex[g]) * lagrange(gx, b, ex[g]);
          D1[(size_t)a * n1 + b] += ew[g] * lagrange_d(gx, a, ex[g]) * lagrange_d(gx, b, ex[g]);
        }
    const double hh = 0.5 * h, J = hh * hh * hh;
    Nq.assign((size_t)nq * ne, 0.0); JxW.assign(nq, 0.0);
    for (int qz = 0; qz < q1; ++qz)
      for (int qy = 0; qy < q1; ++qy)
        for (int qxi = 0; qxi < q1; ++qxi) {
          int q = (qz * q1 + qy) * q1 + qxi;
          JxW[q] = J * qw[qxi] * qw[qy] * qw[qz];
          for (int c = 0; c < n1; ++c)
            for (int b = 0; b < n1; ++b)
              for (int a = 0; a < n1; ++a)
                Nq[(size_t)q * ne + (c * n1 + b) * n1 + a] =
                    Lq[qxi * n1 + a] * Lq[qy * n1 + b] * Lq[qz * n1 + c];
        }
    // K_e = (h/2) [D x M x M + M x D x M + M x M x D]
    Ke.assign((size_t)ne * ne, 0.0);
    for (int c = 0; c < n1; ++c) for (int b = 0; b < n1; ++b) for (int a = 0; a < n1; ++a)
      for (int c2 = 0; c2 < n1; ++c2) for (int b2 = 0; b2 < n1; ++b2) for (int a2 = 0; a2 < n1; ++a2) {
        double Ma = M1[a * n1 + a2], Mb = M1[b * n1 + b2], Mc = M1[c * n1 + c2];
        double Da = D1[a * n1 + a2], Db = D1[b * n1 + b2], Dc = D1[c * n1 + c2];
        Ke[(size_t)((c * n1 + b) * n1 + a) * ne + (c2 * n1 + b2) * n1 + a2] =
            hh * (Da * Mb * Mc + Ma * Db * Mc + Ma * Mb * Dc);
      }
    cellNodes.assign((size_t)ncell * ne, 0);
    for (long e = 0; e < ncell; ++e) {
      long cx = e % nc, cy = (e / nc) % nc, cz = e / ((long)nc * nc);
      for (int c = 0; c < n1; ++c) for (int b = 0; b < n1; ++b) for (int a = 0; a < n1; ++a)
        cellNodes[(size_t)e * ne + (c * n1 + b) * n1 + a] =
            ((cz * p + c) * N1 + (cy * p + b)) * (long)N1 + (cx * p + a);
    }
    // child Gauss points in parent reference coordinates
    const double g = 1.0 / std::sqrt(3.0);
    const double tloc[2] = {-g, g};
    vector<double> cp(4);
    for (int ci = 0; ci < 2; ++ci)
      for (int k = 0; k < 2; ++k) {
        int i = 2 * ci + k;
        cp[i] = (ci == 0 ? -0.5 : 0.5) + 0.5 * tloc[k];
        ccell[i] = ci;
        cw[i][0] = 0.5 * (1.0 - tloc[k]);
        cw[i][1] = 0.5 * (1.0 + tloc[k]);
      }
    Lc.assign((size_t)4 * n1, 0.0);
    for (int i = 0; i < 4; ++i)
      for (int a = 0; a < n1; ++a) Lc[(size_t)i * n1 + a] = lagrange(gx, a, cp[i]);
    qcc.assign(q1, 0); qlw0.assign(q1, 0.0); qlw1.assign(q1, 0.0);
    for (int q = 0; q < q1; ++q) {
      double xi = qx[q];
      int ci = xi < 0.0 ? 0 : 1;
      double t = ci == 0 ? 2.0 * xi + 1.0 : 2.0 * xi - 1.0;
      qcc[q] = ci; qlw0[q] = 0.5 * (1.0 - t); qlw1[q] = 0.5 * (1.0 + t);
    }
  }
  bool boundary(long n) const {
    long i = n % N1, j = (n / N1) % N1, k = n / ((long)N1 * N1);
    return i == 0 || j == 0 || k == 0 || i == N1 - 1 || j == N1 - 1 || k == N1 - 1;
  }
  int colour(long e) const {
    long cx = e % nc, cy = (e / nc) % nc, cz = e / ((long)nc * nc);
    return (int)((cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2));
  }
};

}  // namespace

// --------------------------------------------------------------- handles
struct idft_mv {
  int ncols = 0;
  vector<double> v;
};

namespace {

struct AdjointState {
  bool ready = false;
  int B = 0;
  idft_mv* psi = nullptr;
  vector<double> occ, occ4, negEig, eig;
  vector<int> pairs;
  vector<double> uJxW;
  double fermi = 0.0;
  idft_mv* x = nullptr;
  vector<double> Mu;
};

}  // namespace

struct idft_ctx {
  idft_desc d;
  Mesh m;
  std::string err;
  vector<double> potBase, rhoT, weight, rhoTnodal;
  vector<double> Mdiag, sqrtM, invSqrtM, S, Kdiag, Djac, DjacSqrt;
  vector<char> bnd, adjMask;
  vector<double> veff, He;    // [ncell*nq], [ncell*ne*ne]
  bool haveH = false;
  // eigen state: Loewdin vectors X~ [N][B]
  int B = 0;
  vector<double> X, eig, occ, resid;
  bool haveX = false, haveRitz = false, needRR = false;
  double fermi = 0.0, bup = 0.0, blow = 0.0;
  // inverse-problem state
  long counter = 0;
  double tolCheb = 1e-8, adjointTol = 1e-11, lossPrev = 0.0;
  vector<double> rho, u, rhoDiff;
  double loss = 0.0;
  bool haveForward = false;
  vector<double> pot0, potSol;
  idft_stats st{};
  AdjointState adj;
};

namespace {

int fail(idft_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// Fermi-Dirac occupancy and its derivative w.r.t. the eigenvalue.
double fd_occ(double e, double mu, double T) {
  double x = (e - mu) / (kBoltzmannHa * T);
  if (x >= 0) { double t = std::exp(-x); return t / (1.0 + t); }
  return 1.0 / (1.0 + std::exp(x));
}
double fd_occ_der(double e, double mu, double T) {
  double kt = kBoltzmannHa * T;
  double f = fd_occ(e, mu, T);
  return -f * (1.0 - f) / kt;
}

// Fermi level: bisection on sum_i 2 f(e_i) = N_e.
double fermi_level(const vector<double>& e, double ne, double T) {
  double lo = *std::min_element(e.begin(), e.end()) - 1.0;
  double hi = *std::max_element(e.begin(), e.end()) + 1.0;
  for (int it = 0; it < 400; ++it) {
    double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    double s = 0.0;
    for (double v : e) s += 2.0 * fd_occ(v, mid, T);
    if (s < ne) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

// ------------------------------------------------------- cell operators
// Y = S * sum_e P_e^T Hc_e P_e (S * X); Hc = per-cell matrices [ncell][ne][ne].
void apply_cells(idft_ctx* c, const vector<double>& Hc, const double* X, double* Y,
                 int B, bool lowdin) {
  const Mesh& m = c->m;
  if (lowdin) { c->st.applies_fp64++; c->st.apply_cols_fp64 += B; }
  const int ne = m.ne;
  std::fill(Y, Y + (size_t)m.N * B, 0.0);
  for (int col = 0; col < 8; ++col) {
#pragma omp parallel
    {
      vector<double> Xe((size_t)ne * B), Ye((size_t)ne * B);
#pragma omp for schedule(static)
      for (long e = 0; e < m.ncell; ++e) {
        if (m.colour(e) != col) continue;
        const long* nodes = &m.cellNodes[(size_t)e * ne];
        for (int l = 0; l < ne; ++l) {
          double s = lowdin ? c->S[nodes[l]] : 1.0;
          for (int j = 0; j < B; ++j) Xe[(size_t)l * B + j] = s * X[(size_t)nodes[l] * B + j];
        }
        const double* H = &Hc[(size_t)e * ne * ne];
        for (int a = 0; a < ne; ++a) {
          for (int j = 0; j < B; ++j) Ye[(size_t)a * B + j] = 0.0;
          for (int b = 0; b < ne; ++b) {
            double h = H[(size_t)a * ne + b];
            for (int j = 0; j < B; ++j) Ye[(size_t)a * B + j] += h * Xe[(size_t)b * B + j];
          }
        }
        for (int l = 0; l < ne; ++l)
          for (int j = 0; j < B; ++j) Y[(size_t)nodes[l] * B + j] += Ye[(size_t)l * B + j];
      }
    }
  }
  if (lowdin) {
#pragma omp parallel for schedule(static)
    for (long n = 0; n < m.N; ++n)
      for (int j = 0; j < B; ++j) Y[(size_t)n * B + j] *= c->S[n];
  }
}

void build_hamiltonian(idft_ctx* c) {
  const Mesh& m = c->m;
  const int ne = m.ne, nq = m.nq;
  c->He.assign((size_t)m.ncell * ne * ne, 0.0);
#pragma omp parallel
  {
    vector<double> T((size_t)nq * ne);
#pragma omp for schedule(static)
    for (long e = 0; e < m.ncell; ++e) {
      for (int q = 0; q < nq; ++q) {
        double w = c->veff[(size_t)e * nq + q] * m.JxW[q];
        for (int l = 0; l < ne; ++l) T[(size_t)q * ne + l] = w * m.Nq[(size_t)q * ne + l];
      }
      double* H = &c->He[(size_t)e * ne * ne];
      for (size_t i = 0; i < (size_t)ne * ne; ++i) H[i] = 0.5 * m.Ke[i];
      for (int q = 0; q < nq; ++q)
        for (int a = 0; a < ne; ++a) {
          const double na = m.Nq[(size_t)q * ne + a];
          const double* tq = &T[(size_t)q * ne];
          double* ha = &H[(size_t)a * ne];
          for (int b = 0; b < ne; ++b) ha[b] += na * tq[b];
        }
    }
  }
  c->haveH = true;
}

// values of a node-major multivector at the parent quadrature: out[e][q][j]
void interp_quad(const idft_ctx* c, const double* X, int B, vector<double>& out) {
  const Mesh& m = c->m;
  const int ne = m.ne, nq = m.nq;
  out.assign((size_t)m.ncell * nq * B, 0.0);
#pragma omp parallel
  {
    vector<double> Xe((size_t)ne * B);
#pragma omp for schedule(static)
    for (long e = 0; e < m.ncell; ++e) {
      const long* nodes = &m.cellNodes[(size_t)e * ne];
      for (int l = 0; l < ne; ++l)
        for (int j = 0; j < B; ++j) Xe[(size_t)l * B + j] = X[(size_t)nodes[l] * B + j];
      for (int q = 0; q < nq; ++q) {
        double* o = &out[((size_t)e * nq + q) * B];
        for (int l = 0; l < ne; ++l) {
          double nv = m.Nq[(size_t)q * ne + l];
          for (int j = 0; j < B; ++j) o[j] += nv * Xe[(size_t)l * B + j];
        }
      }
    }
  }
}

// values at the 64 child Gauss points of every parent cell: out[e][qc][j]
void interp_child(const idft_ctx* c, const double* X, int B, vector<double>& out) {
  const Mesh& m = c->m;
  const int ne = m.ne, n1 = m.n1;
  out.assign((size_t)m.ncell * 64 * B, 0.0);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < m.ncell; ++e) {
    const long* nodes = &m.cellNodes[(size_t)e * ne];
    for (int z = 0; z < 4; ++z) for (int y = 0; y < 4; ++y) for (int x = 0; x < 4; ++x) {
      double* o = &out[((size_t)e * 64 + (z * 4 + y) * 4 + x) * B];
      for (int cc = 0; cc < n1; ++cc) for (int b = 0; b < n1; ++b) for (int a = 0; a < n1; ++a) {
        double nv = m.Lc[x * n1 + a] * m.Lc[y * n1 + b] * m.Lc[z * n1 + cc];
        const double* xv = &X[(size_t)nodes[(cc * n1 + b) * n1 + a] * B];
        for (int j = 0; j < B; ++j) o[j] += nv * xv[j];
      }
    }
  }
}

// v_xc (child nodal) -> parent quadrature values
void vxc_to_quad(const idft_ctx* c, const double* pot, vector<double>& out) {
  const Mesh& m = c->m;
  const int q1 = m.q1, nq = m.nq;
  out.assign((size_t)m.ncell * nq, 0.0);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < m.ncell; ++e) {
    long cx = e % m.nc, cy = (e / m.nc) % m.nc, cz = e / ((long)m.nc * m.nc);
    for (int qz = 0; qz < q1; ++qz) for (int qy = 0; qy < q1; ++qy) for (int qx = 0; qx < q1; ++qx) {
      long ix = 2 * cx + m.qcc[qx], iy = 2 * cy + m.qcc[qy], iz = 2 * cz + m.qcc[qz];
      double wx[2] = {m.qlw0[qx], m.qlw1[qx]}, wy[2] = {m.qlw0[qy], m.qlw1[qy]},
             wz[2] = {m.qlw0[qz], m.qlw1[qz]};
      double v = 0.0;
      for (int dz = 0; dz < 2; ++dz) for (int dy = 0; dy < 2; ++dy) for (int dx = 0; dx < 2; ++dx)
        v += wx[dx] * wy[dy] * wz[dz] *
             pot[((iz + dz) * m.NV1 + (iy + dy)) * (long)m.NV1 + (ix + dx)];
      out[(size_t)e * nq + (qz * q1 + qy) * q1 + qx] = v;
    }
  }
}

// ------------------------------------------------------------- ChFSI
// Lanczos on the Loewdin Hamiltonian from a seeded random vector; returns
// the upper bound theta_max + |beta_k| and the lowest Ritz value.
void lanczos_bounds(idft_ctx* c, double& upper, double& lowest) {
  const Mesh& m = c->m;
  const int k = std::max(2, c->d.lanczos_steps);
  const long N = m.N;
  vector<double> v(N), vprev(N, 0.0), w(N);
  uint64_t s = c->d.seed ^ 0x5bd1e995ULL;
  double nrm = 0.0;
  for (long n = 0; n < N; ++n) { double r = rand_pm1(s); v[n] = c->bnd[n] ? 0.0 : r; nrm += v[n] * v[n]; }
  nrm = std::sqrt(nrm);
  for (long n = 0; n < N; ++n) v[n] /= nrm;
  vector<double> al, be;
  double beta = 0.0;
  for (int j = 0; j < k; ++j) {
    apply_cells(c, c->He, v.data(), w.data(), 1, true);
    double a = 0.0;
    for (long n = 0; n < N; ++n) a += w[n] * v[n];
    for (long n = 0; n < N; ++n) w[n] -= a * v[n] + beta * vprev[n];
    double b = 0.0;
    for (long n = 0; n < N; ++n) b += w[n] * w[n];
    b = std::sqrt(b);
    al.push_back(a); be.push_back(b);
    beta = b;
    if (b == 0.0) break;
    for (long n = 0; n < N; ++n) { vprev[n] = v[n]; v[n] = w[n] / b; }
  }
  int kk = (int)al.size();
  vector<double> T((size_t)kk * kk, 0.0), ev, V;
  for (int i = 0; i < kk; ++i) {
    T[(size_t)i * kk + i] = al[i];
    if (i + 1 < kk) { T[(size_t)i * kk + i + 1] = be[i]; T[(size_t)(i + 1) * kk + i] = be[i]; }
  }
  sym_eig(kk, T, ev, V);
  upper = ev[kk - 1] + std::fabs(be[kk - 1]);
  lowest = ev[0];
}

void chebyshev_filter(idft_ctx* c, double* X, int B, int deg, double a, double b, double a0) {
  const long NB = c->m.N * (long)B;
  const double e = 0.5 * (b - a), cc = 0.5 * (b + a);
  double sigma = e / (a0 - cc);
  const double sigma1 = sigma, tau = 2.0 / sigma1;
  vector<double> Y(NB), HY(NB);
  apply_cells(c, c->He, X, HY.data(), B, true);
  for (long i = 0; i < NB; ++i) Y[i] = (HY[i] - cc * X[i]) * (sigma1 / e);
  for (int k = 2; k <= deg; ++k) {
    double sigma2 = 1.0 / (tau - sigma);
    apply_cells(c, c->He, Y.data(), HY.data(), B, true);
    for (long i = 0; i < NB; ++i) {
      double yn = (HY[i] - cc * Y[i]) * (2.0 * sigma2 / e) - (sigma * sigma2) * X[i];
      X[i] = Y[i];
      Y[i] = yn;
    }
    sigma = sigma2;
  }
  std::copy(Y.begin(), Y.end(), X);
}

void gram(const idft_ctx* c, const double* X, const double* Y, int B, vector<double>& G) {
  const long N = c->m.N;
  G.assign((size_t)B * B, 0.0);
  for (int i = 0; i < B; ++i)
    for (int j = 0; j < B; ++j) {
      double s = 0.0;
      for (long n = 0; n < N; ++n) s += X[(size_t)n * B + i] * Y[(size_t)n * B + j];
      G[(size_t)i * B + j] = s;
    }
}

// X <- X Q (Q is B x B row-major)
void right_mult(const idft_ctx* c, double* X, int B, const vector<double>& Q) {
  const long N = c->m.N;
#pragma omp parallel
  {
    vector<double> t(B);
#pragma omp for schedule(static)
    for (long n = 0; n < N; ++n) {
      for (int j = 0; j < B; ++j) {
        double s = 0.0;
        for (int i = 0; i < B; ++i) s += X[(size_t)n * B + i] * Q[(size_t)i * B + j];
        t[j] = s;
      }
      for (int j = 0; j < B; ++j) X[(size_t)n * B + j] = t[j];
    }
  }
}

// Two orthonormalisation sweeps; each is Cholesky-QR, or SVQB (Stathopoulos &
// Wu 2002: X <- X V diag(max(lambda, tau))^{-1/2} with S = V diag(lambda) V^T)
// when the Gram matrix is not numerically positive definite.
void orthonormalize(idft_ctx* c, double* X, int B) {
  for (int pass = 0; pass < 2; ++pass) {
    vector<double> G, L;
    gram(c, X, X, B, G);
    L = G;
    vector<double> Q((size_t)B * B, 0.0);
    if (cholesky(B, L)) {
      for (int j = 0; j < B; ++j) {  // row j of L^{-1} -> column j of L^{-T}
        vector<double> y(B, 0.0);
        for (int i = 0; i < B; ++i) {
          double s = (i == j) ? 1.0 : 0.0;
          for (int k = 0; k < i; ++k) s -= L[(size_t)i * B + k] * y[k];
          y[i] = s / L[(size_t)i * B + i];
        }
        for (int i = 0; i < B; ++i) Q[(size_t)j * B + i] = y[i];
      }
    } else {
      vector<double> lam, V;
      sym_eig(B, G, lam, V);
      const double tau = 1e-14 * std::max(lam[B - 1], 1e-300);
      for (int i = 0; i < B; ++i)
        for (int j = 0; j < B; ++j) Q[(size_t)i * B + j] = V[(size_t)i * B + j] / std::sqrt(std::max(lam[j], tau));
    }
    right_mult(c, X, B, Q);
  }
}

void rayleigh_ritz(idft_ctx* c, double* X, int B, vector<double>& eval, vector<double>& res) {
  const long N = c->m.N;
  vector<double> HX((size_t)N * B), Hp, Q;
  apply_cells(c, c->He, X, HX.data(), B, true);
  gram(c, X, HX.data(), B, Hp);
  for (int i = 0; i < B; ++i)
    for (int j = i + 1; j < B; ++j) {
      double s = 0.5 * (Hp[(size_t)i * B + j] + Hp[(size_t)j * B + i]);
      Hp[(size_t)i * B + j] = Hp[(size_t)j * B + i] = s;
    }
  sym_eig(B, Hp, eval, Q);
  right_mult(c, X, B, Q);
  right_mult(c, HX.data(), B, Q);
  res.assign(B, 0.0);
  for (int j = 0; j < B; ++j) {
    double s = 0.0;
    for (long n = 0; n < N; ++n) {
      double r = HX[(size_t)n * B + j] - eval[j] * X[(size_t)n * B + j];
      s += r * r;
    }
    res[j] = std::sqrt(s);
  }
}

void init_subspace(idft_ctx* c) {
  const Mesh& m = c->m;
  c->X.assign((size_t)m.N * c->B, 0.0);
  uint64_t s = c->d.seed;
  for (long n = 0; n < m.N; ++n)
    for (int j = 0; j < c->B; ++j) {
      double r = rand_pm1(s);
      c->X[(size_t)n * c->B + j] = c->bnd[n] ? 0.0 : r;
    }
  c->haveX = true;
  c->haveRitz = false;
}

void update_occupancy(idft_ctx* c) {
  c->fermi = fermi_level(c->eig, c->d.n_electrons, c->d.temperature);
  c->occ.assign(c->B, 0.0);
  for (int i = 0; i < c->B; ++i) c->occ[i] = fd_occ(c->eig[i], c->fermi, c->d.temperature);
}

// One ChFSI pass.  Filter interval [a, b] scaled at a0: after a Rayleigh-Ritz
// a0 = lowest and a = highest Ritz value of the block; on the first pass from
// a random block a0 = lowest Lanczos Ritz value and a = a0 + (b - a0)/10.
int chfsi_pass(idft_ctx* c) {
  const int B = c->B;
  const double b = c->bup;
  double a0, a;
  if (c->haveRitz) { a0 = c->eig[0]; a = c->eig[B - 1]; }
  else { a0 = c->blow; a = a0 + 0.1 * (b - a0); }
  chebyshev_filter(c, c->X.data(), B, c->d.cheb_degree, a, b, a0);
  orthonormalize(c, c->X.data(), B);
  rayleigh_ritz(c, c->X.data(), B, c->eig, c->resid);
  for (double r : c->resid) if (!std::isfinite(r)) return fail(c, IDFT_ENAN, "non-finite Ritz residual");
  c->haveRitz = true;
  return IDFT_OK;
}

int solve_passes(idft_ctx* c, double tol, int maxPasses, int* passesOut, double* maxResOut) {
  if (!c->haveH) return fail(c, IDFT_ESTATE, "no Hamiltonian: call set_potential first");
  if (!c->haveX) init_subspace(c);
  lanczos_bounds(c, c->bup, c->blow);
  if (c->needRR) {  // caller-provided block: Ritz values give the filter bounds
    orthonormalize(c, c->X.data(), c->B);
    rayleigh_ritz(c, c->X.data(), c->B, c->eig, c->resid);
    c->haveRitz = true;
    c->needRR = false;
  }
  c->st.spectrum_upper_bound = c->bup;
  int pass = 0;
  double maxRes = 0.0;
  do {
    int rc = chfsi_pass(c);
    if (rc) return rc;
    update_occupancy(c);
    maxRes = 0.0;
    int homo = 0;
    for (int i = 0; i < c->B; ++i)
      if (c->eig[i] < c->fermi + 1e-3 && homo < i) homo = i;
    for (int i = 0; i < c->B; ++i)
      if (c->occ[i] > c->d.frac_occ_tol || i <= homo + c->d.additional_states)
        maxRes = std::max(maxRes, c->resid[i]);
    pass++;
  } while (maxRes > tol && pass < maxPasses);
  c->st.cheb_passes = pass;
  c->st.max_residual = maxRes;
  c->st.fermi_energy = c->fermi;
  if (passesOut) *passesOut = pass;
  if (maxResOut) *maxResOut = maxRes;
  return IDFT_OK;
}

void physical_psi(const idft_ctx* c, vector<double>& psi) {
  const long N = c->m.N;
  const int B = c->B;
  psi.assign((size_t)N * B, 0.0);
  for (long n = 0; n < N; ++n)
    for (int j = 0; j < B; ++j) psi[(size_t)n * B + j] = c->S[n] * c->X[(size_t)n * B + j];
}

void density(idft_ctx* c, vector<double>& rho) {
  vector<double> psi, pq;
  physical_psi(c, psi);
  interp_quad(c, psi.data(), c->B, pq);
  const long nQ = c->m.ncell * (long)c->m.nq;
  rho.assign(nQ, 0.0);
  for (long q = 0; q < nQ; ++q) {
    double s = 0.0;
    for (int j = 0; j < c->B; ++j) { double v = pq[(size_t)q * c->B + j]; s += 2.0 * c->occ[j] * v * v; }
    rho[q] = s;
  }
}

// -------------------------------------------------------- adjoint problem
void zero_boundary(const idft_ctx* c, vector<double>& X, int B) {
  for (long n = 0; n < c->m.N; ++n)
    if (c->bnd[n]) for (int j = 0; j < B; ++j) X[(size_t)n * B + j] = 0.0;
}

// computeMuMatrix (MultiVectorAdjointLinearSolverProblem.cpp:912-986)
void compute_mu(idft_ctx* c, const vector<double>& occ) {
  AdjointState& A = c->adj;
  const int B = A.B, nq = c->m.nq;
  const long ncell = c->m.ncell;
  vector<double> pq;
  interp_quad(c, A.psi->v.data(), B, pq);
  const int np = (int)A.pairs.size() / 2;
  vector<double> cellMu((size_t)ncell * np, 0.0);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < ncell; ++e)
    for (int k = 0; k < np; ++k) {
      int i = A.pairs[2 * k], j = A.pairs[2 * k + 1];
      double s = 0.0;
      for (int q = 0; q < nq; ++q)
        s += 2.0 * occ[i] * pq[((size_t)e * nq + q) * B + i] * pq[((size_t)e * nq + q) * B + j] *
             A.uJxW[(size_t)e * nq + q];
      cellMu[(size_t)e * np + k] = s;
    }
  A.Mu.assign((size_t)B * B, 0.0);
  for (int k = 0; k < np; ++k) {
    int i = A.pairs[2 * k], j = A.pairs[2 * k + 1];
    for (long e = 0; e < ncell; ++e) A.Mu[(size_t)i * B + j] += cellMu[(size_t)e * np + k];
  }
}

int adj_compute_rhs(idft_ctx* c, idft_mv* x, idft_mv* rhs) {
  AdjointState& A = c->adj;
  if (!A.ready) return fail(c, IDFT_ESTATE, "computeRhs before updateInputPsi");
  const Mesh& m = c->m;
  const int B = A.B, ne = m.ne, nq = m.nq;
  const long N = m.N;
  if (x->ncols != B || rhs->ncols != B) return fail(c, IDFT_EINVAL, "block size mismatch");
  A.x = x;
  vector<double> psiT = A.psi->v;
  zero_boundary(c, psiT, B);
  for (long n = 0; n < N; ++n) for (int j = 0; j < B; ++j) psiT[(size_t)n * B + j] *= c->sqrtM[n];
  // computeRMatrix (stored R_e, .cpp:309-366)
  vector<double> R((size_t)m.ncell * ne * ne, 0.0);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < m.ncell; ++e) {
    double* Re = &R[(size_t)e * ne * ne];
    for (int q = 0; q < nq; ++q) {
      const double w = A.uJxW[(size_t)e * nq + q];
      const double* nqv = &m.Nq[(size_t)q * ne];
      for (int a = 0; a < ne; ++a) {
        const double na = nqv[a] * w;
        double* ra = &Re[(size_t)a * ne];
        for (int b = 0; b < ne; ++b) ra[b] += na * nqv[b];
      }
    }
  }
  compute_mu(c, A.occ);
  // rhs = -2 Mu^T psiT  (xgemm at .cpp:722-725)
  rhs->v.assign((size_t)N * B, 0.0);
  for (long n = 0; n < N; ++n)
    for (int i = 0; i < B; ++i) {
      double s = 0.0;
      for (int k = 0; k < B; ++k) s += A.Mu[(size_t)k * B + i] * psiT[(size_t)n * B + k];
      rhs->v[(size_t)n * B + i] = -2.0 * s;
    }
  for (long n = 0; n < N; ++n) for (int j = 0; j < B; ++j) psiT[(size_t)n * B + j] *= c->invSqrtM[n];
  zero_boundary(c, psiT, B);
  vector<double> psiT2((size_t)N * B, 0.0);
  apply_cells(c, R, psiT.data(), psiT2.data(), B, false);
  zero_boundary(c, psiT2, B);
  for (long n = 0; n < N; ++n) for (int j = 0; j < B; ++j) psiT2[(size_t)n * B + j] *= c->invSqrtM[n];
  zero_boundary(c, psiT2, B);
  for (long n = 0; n < N; ++n)
    for (int j = 0; j < B; ++j) rhs->v[(size_t)n * B + j] += psiT2[(size_t)n * B + j] * A.occ4[j];
  zero_boundary(c, rhs->v, B);
  return IDFT_OK;
}

int adj_vmult(idft_ctx* c, idft_mv* Ax, idft_mv* x) {
  AdjointState& A = c->adj;
  if (!A.ready) return fail(c, IDFT_ESTATE, "vmult before updateInputPsi");
  const int B = A.B;
  if (x->ncols != B || Ax->ncols != B) return fail(c, IDFT_EINVAL, "block size mismatch");
  Ax->v.assign((size_t)c->m.N * B, 0.0);
  apply_cells(c, c->He, x->v.data(), Ax->v.data(), B, true);
  zero_boundary(c, x->v, B);
  zero_boundary(c, Ax->v, B);
  for (long n = 0; n < c->m.N; ++n)
    for (int j = 0; j < B; ++j) Ax->v[(size_t)n * B + j] += A.negEig[j] * x->v[(size_t)n * B + j];
  return IDFT_OK;
}

void precondition(const idft_ctx* c, const vector<double>& D, idft_mv* dst, const idft_mv* src) {
  const int B = src->ncols;
  dst->ncols = B;
  dst->v.resize((size_t)c->m.N * B);
  for (long n = 0; n < c->m.N; ++n)
    for (int j = 0; j < B; ++j) dst->v[(size_t)n * B + j] = D[n] * src->v[(size_t)n * B + j];
}

void dot_quad_wise(idft_ctx* c, const double* v1, const double* v2, vector<double>& out) {
  AdjointState& A = c->adj;
  const int B = A.B, nq = c->m.nq;
  vector<double> q1v, q2v;
  interp_quad(c, v1, B, q1v);
  interp_quad(c, v2, B, q2v);
  const int np = (int)A.pairs.size() / 2;
  out.assign(np, 0.0);
  for (int k = 0; k < np; ++k) {
    int i = A.pairs[2 * k], j = A.pairs[2 * k + 1];
    double s = 0.0;
    for (long e = 0; e < c->m.ncell; ++e)
      for (int q = 0; q < nq; ++q)
        s += c->m.JxW[q] * q1v[((size_t)e * nq + q) * B + i] * q2v[((size_t)e * nq + q) * B + j];
    out[k] = s;
  }
}

int adj_distribute_x(idft_ctx* c) {
  AdjointState& A = c->adj;
  if (!A.ready || !A.x) return fail(c, IDFT_ESTATE, "distributeX before computeRhs");
  const int B = A.B;
  const long N = c->m.N;
  vector<double>& x = A.x->v;
  for (long n = 0; n < N; ++n) for (int j = 0; j < B; ++j) x[(size_t)n * B + j] *= c->invSqrtM[n];
  zero_boundary(c, x, B);
  vector<double> dot;
  if (c->d.gradient_mode == IDFT_GRADIENT_EXACT) {
    const int np0 = (int)A.pairs.size() / 2;
    dot.assign(np0, 0.0);
    for (int k = 0; k < np0; ++k) {
      int i = A.pairs[2 * k], j = A.pairs[2 * k + 1];
      double s = 0.0;
      for (long n = 0; n < N; ++n) s += c->Mdiag[n] * x[(size_t)n * B + i] * A.psi->v[(size_t)n * B + j];
      dot[k] = s;
    }
  } else {
    dot_quad_wise(c, x.data(), A.psi->v.data(), dot);
  }
  vector<double> ones(B, 1.0);
  compute_mu(c, ones);
  double sumD = 0.0, etaR = 0.0;
  for (int i = 0; i < B; ++i) {
    double dF = -fd_occ_der(A.eig[i], A.fermi, c->d.temperature);
    sumD += dF;
    etaR += dF * A.Mu[(size_t)i * B + i];
  }
  double eta = std::fabs(etaR) < 1e-16 ? 0.0 : etaR / sumD;
  const int np = (int)A.pairs.size() / 2;
  for (int k = 0; k < np; ++k) {
    int i = A.pairs[2 * k], j = A.pairs[2 * k + 1];
    double dFi = fd_occ_der(A.eig[i], A.fermi, c->d.temperature);
    double corr = -dFi * A.Mu[(size_t)i * B + j];
    if (i == j) corr += dFi * eta;
    dot[k] = -dot[k] + corr;
  }
  for (long n = 0; n < N; ++n)
    for (int k = 0; k < np; ++k) {
      int i = A.pairs[2 * k], j = A.pairs[2 * k + 1];
      x[(size_t)n * B + i] += dot[k] * A.psi->v[(size_t)n * B + j];
    }
  return IDFT_OK;
}

// Preconditioned MINRES (Paige-Saunders), one independent recurrence per
// column, x0 = 0, stop when every column's phibar (= ||r||_D) <= tol.
int minres(idft_ctx* c, idft_mv* x, idft_mv* ndbc, double tol, int maxit, int distribute,
           int32_t* iters) {
  AdjointState& A = c->adj;
  const int B = A.B;
  const long N = c->m.N, NB = N * (long)B;
  idft_mv rhs; rhs.ncols = B;
  int rc = adj_compute_rhs(c, x, &rhs);
  if (rc) return rc;
  for (long i = 0; i < NB; ++i) if (!std::isfinite(rhs.v[i])) return fail(c, IDFT_ENAN, "non-finite adjoint rhs");
  x->v.assign(NB, 0.0);
  idft_mv r1, r2, y, v, w, w1, w2, Av;
  r1.ncols = r2.ncols = y.ncols = v.ncols = w.ncols = w1.ncols = w2.ncols = Av.ncols = B;
  r1.v = rhs.v; r2.v = rhs.v;
  precondition(c, c->Djac, &y, &r1);
  w.v.assign(NB, 0.0); w1.v.assign(NB, 0.0); w2.v.assign(NB, 0.0); v.v.assign(NB, 0.0);
  vector<double> beta(B), oldb(B, 0.0), dbar(B, 0.0), epsln(B, 0.0), phibar(B), cs(B, -1.0),
      sn(B, 0.0), alfa(B), tmp(B);
  vector<char> done(B, 0);
  vector<int> it(B, 0);
  for (int j = 0; j < B; ++j) {
    double s = 0.0;
    for (long n = 0; n < N; ++n) s += r1.v[(size_t)n * B + j] * y.v[(size_t)n * B + j];
    beta[j] = std::sqrt(std::max(s, 0.0));
    phibar[j] = beta[j];
    if (beta[j] <= tol) done[j] = 1;
  }
  int k = 0;
  for (; k < maxit; ++k) {
    bool all = true;
    for (int j = 0; j < B; ++j) all = all && done[j];
    if (all) break;
    for (long n = 0; n < N; ++n)
      for (int j = 0; j < B; ++j)
        v.v[(size_t)n * B + j] = done[j] ? 0.0 : y.v[(size_t)n * B + j] / beta[j];
    rc = adj_vmult(c, &Av, &v);
    if (rc) return rc;
    for (int j = 0; j < B; ++j) tmp[j] = (k >= 1 && !done[j]) ? beta[j] / oldb[j] : 0.0;
    for (long n = 0; n < N; ++n)
      for (int j = 0; j < B; ++j)
        y.v[(size_t)n * B + j] = Av.v[(size_t)n * B + j] - tmp[j] * r1.v[(size_t)n * B + j];
    for (int j = 0; j < B; ++j) {
      double s = 0.0;
      for (long n = 0; n < N; ++n) s += v.v[(size_t)n * B + j] * y.v[(size_t)n * B + j];
      alfa[j] = s;
    }
    for (long n = 0; n < N; ++n)
      for (int j = 0; j < B; ++j) {
        double yn = y.v[(size_t)n * B + j] - (done[j] ? 0.0 : alfa[j] / beta[j]) * r2.v[(size_t)n * B + j];
        r1.v[(size_t)n * B + j] = r2.v[(size_t)n * B + j];
        r2.v[(size_t)n * B + j] = yn;
      }
    precondition(c, c->Djac, &y, &r2);
    vector<double> gam(B), phi(B), delta(B), oldeps(B);
    for (int j = 0; j < B; ++j) {
      if (done[j]) continue;
      double s = 0.0;
      for (long n = 0; n < N; ++n) s += r2.v[(size_t)n * B + j] * y.v[(size_t)n * B + j];
      oldb[j] = beta[j];
      beta[j] = std::sqrt(std::max(s, 0.0));
      oldeps[j] = epsln[j];
      delta[j] = cs[j] * dbar[j] + sn[j] * alfa[j];
      double gbar = sn[j] * dbar[j] - cs[j] * alfa[j];
      epsln[j] = sn[j] * beta[j];
      dbar[j] = -cs[j] * beta[j];
      double g = std::sqrt(gbar * gbar + beta[j] * beta[j]);
      g = std::max(g, 1e-300);
      cs[j] = gbar / g; sn[j] = beta[j] / g;
      phi[j] = cs[j] * phibar[j];
      phibar[j] = sn[j] * phibar[j];
      gam[j] = g;
    }
    for (long n = 0; n < N; ++n)
      for (int j = 0; j < B; ++j) {
        if (done[j]) continue;
        size_t id = (size_t)n * B + j;
        double wn = (v.v[id] - oldeps[j] * w1.v[id] - delta[j] * w.v[id]) / gam[j];
        w2.v[id] = w1.v[id];
        w1.v[id] = w.v[id];
        w.v[id] = wn;
        x->v[id] += phi[j] * wn;
      }
    for (int j = 0; j < B; ++j) {
      if (done[j]) continue;
      it[j] = k + 1;
      if (phibar[j] <= tol || beta[j] == 0.0) done[j] = 1;
    }
  }
  (void)ndbc;
  long tot = 0; int mx = 0;
  for (int j = 0; j < B; ++j) { tot += it[j]; mx = std::max(mx, it[j]); if (iters) iters[j] = it[j]; }
  c->st.minres_iters_max = std::max(c->st.minres_iters_max, mx);
  c->st.minres_iters_total += tot;
  if (distribute) return adj_distribute_x(c);
  return IDFT_OK;
}

void degeneracy_map(const vector<double>& ev, double tol, vector<vector<int>>& out) {
  const int N = (int)ev.size();
  std::map<int, std::set<int>> groups;
  std::map<int, int> eigToGroup;
  int count = 0;
  for (int i = 0; i < N; ++i) {
    auto it = eigToGroup.find(i);
    int g;
    if (it != eigToGroup.end()) g = it->second;
    else { g = count++; groups[g].insert(i); eigToGroup[i] = g; }
    for (int j = 0; j < N; ++j)
      if (std::fabs(ev[i] - ev[j]) < tol) { groups[g].insert(j); eigToGroup[j] = g; }
  }
  out.assign(N, {});
  for (int i = 0; i < N; ++i) {
    const std::set<int>& s = groups[eigToGroup[i]];
    out[i].assign(s.begin(), s.end());
  }
}

int update_input_psi(idft_ctx* c, idft_mv* psi, const double* occ, const double* u_q,
                     const int32_t* pairs, int np, double fermi, double* eigv, int B) {
  AdjointState& A = c->adj;
  if (!psi || psi->ncols != B || B <= 0) return fail(c, IDFT_EINVAL, "psi block size mismatch");
  A.B = B; A.psi = psi; A.fermi = fermi;
  zero_boundary(c, psi->v, B);
  A.occ.assign(occ, occ + B);
  A.occ4.resize(B);
  for (int i = 0; i < B; ++i) A.occ4[i] = 4.0 * occ[i];
  A.pairs.assign(pairs, pairs + 2 * np);
  for (int k = 0; k < 2 * np; ++k)
    if (A.pairs[k] < 0 || A.pairs[k] >= B) return fail(c, IDFT_EINVAL, "pair index out of range");
  A.eig.assign(eigv, eigv + B);
  for (int i = 0; i < B; ++i) eigv[i] = -eigv[i];
  A.negEig.assign(eigv, eigv + B);
  const long nQ = c->m.ncell * (long)c->m.nq;
  A.uJxW.resize(nQ);
  for (long q = 0; q < nQ; ++q) A.uJxW[q] = u_q[q] * c->m.JxW[q % c->m.nq];
  A.x = nullptr;
  A.ready = true;
  return IDFT_OK;
}

// ------------------------------------------------------- getForceVector
int forward(idft_ctx* c, const double* pot, double* lossOut) {
  const Mesh& m = c->m;
  double t0 = now_ms();
  // solveEigen (InverseDFTSolverFunction.cpp:1313-1477)
  const double chebyTol = c->d.cheb_tol;
  if (c->counter > 3) {
    double prev = c->tolCheb;
    c->tolCheb = std::min(chebyTol, c->lossPrev / c->d.adaptive_factor_cheb);
    c->tolCheb = std::min(c->tolCheb, prev);
  } else {
    c->tolCheb = std::min(chebyTol, c->d.initial_tol_cheb);
  }
  vector<double> vq;
  vxc_to_quad(c, pot, vq);
  const long nQ = m.ncell * (long)m.nq;
  c->veff.resize(nQ);
  for (long q = 0; q < nQ; ++q) c->veff[q] = c->potBase[q] + vq[q];
  build_hamiltonian(c);
  int rc = solve_passes(c, c->tolCheb, c->d.max_cheb_passes, nullptr, nullptr);
  if (rc) return rc;
  c->st.cheb_tol_used = c->tolCheb;
  double t1 = now_ms();
  density(c, c->rho);
  double t2 = now_ms();
  c->rhoDiff.resize(nQ); c->u.resize(nQ);
  double loss = 0, lunw = 0, l1 = 0, irv = 0;
  for (long q = 0; q < nQ; ++q) {
    double jxw = m.JxW[q % m.nq];
    double rd = c->rhoT[q] - 0.5 * c->rho[q];
    c->rhoDiff[q] = rd;
    c->u[q] = rd * c->weight[q];
    l1 += std::fabs(rd * jxw);
    lunw += rd * rd * jxw;
    loss += rd * rd * c->weight[q] * jxw;
    irv += c->veff[q] * c->rho[q] * jxw;
  }
  c->loss = loss;
  c->st.loss_unweighted = lunw; c->st.l1_density_error = l1; c->st.int_rho_veff = irv;
  c->st.t_eigen_ms = t1 - t0; c->st.t_density_ms = t2 - t1;
  if (lossOut) *lossOut = loss;
  c->haveForward = true;
  return IDFT_OK;
}

int backward(idft_ctx* c, double* force) {
  if (!c->haveForward) return fail(c, IDFT_ESTATE, "backward before forward");
  const Mesh& m = c->m;
  const int Btot = c->B;
  const long N = m.N;
  double t0 = now_ms();
  vector<double> psiAll;
  physical_psi(c, psiAll);
  const bool childQ = c->d.gradient_mode == IDFT_GRADIENT_REFERENCE;
  const int npts = childQ ? 64 : m.nq;
  vector<double> Ssum((size_t)m.ncell * npts, 0.0);
  c->st.minres_iters_max = 0; c->st.minres_iters_total = 0;
  const int defBS = c->d.block_size > 0 ? c->d.block_size : Btot;
  int jvec = 0;
  c->st.adjoint_blocks = 0;
  while (jvec < Btot) {
    c->st.adjoint_blocks++;
    int bs = std::min(defBS, Btot - jvec);
    while (true) {
      int last = jvec + bs - 1;
      if (last + 1 != Btot && std::fabs(c->eig[last] - c->eig[last + 1]) < c->d.degeneracy_tol && bs > 1) bs--;
      else break;
    }
    idft_mv psiB; psiB.ncols = bs; psiB.v.assign((size_t)N * bs, 0.0);
    for (long n = 0; n < N; ++n)
      for (int j = 0; j < bs; ++j) psiB.v[(size_t)n * bs + j] = psiAll[(size_t)n * Btot + jvec + j];
    vector<double> occ(bs), shifts(bs);
    for (int j = 0; j < bs; ++j) { shifts[j] = c->eig[jvec + j]; occ[j] = c->occ[jvec + j]; }
    vector<vector<int>> dm;
    degeneracy_map(shifts, c->d.degeneracy_tol, dm);
    vector<int32_t> pairs;
    for (int i = 0; i < bs; ++i) for (int j : dm[i]) { pairs.push_back(i); pairs.push_back(j); }
    int rc = update_input_psi(c, &psiB, occ.data(), c->u.data(), pairs.data(),
                              (int)pairs.size() / 2, c->fermi, shifts.data(), bs);
    if (rc) return rc;
    c->adjointTol = std::min(c->adjointTol, c->tolCheb / c->d.adaptive_factor_adjoint);
    c->st.adjoint_tol_used = c->adjointTol;
    idft_mv x, ndbc; x.ncols = ndbc.ncols = bs;
    x.v.assign((size_t)N * bs, 0.0); ndbc.v.assign((size_t)N * bs, 0.0);
    rc = minres(c, &x, &ndbc, c->adjointTol, c->d.adjoint_max_iter, 1, nullptr);
    if (rc) return rc;
    // adjoint constraints: p = 0 where rho_target < rhoTol
    for (long n = 0; n < N; ++n)
      if (c->adjMask[n]) for (int j = 0; j < bs; ++j) x.v[(size_t)n * bs + j] = 0.0;
    vector<double> pq, xq;
    if (childQ) { interp_child(c, psiB.v.data(), bs, pq); interp_child(c, x.v.data(), bs, xq); }
    else { interp_quad(c, psiB.v.data(), bs, pq); interp_quad(c, x.v.data(), bs, xq); }
    for (size_t q = 0; q < Ssum.size(); ++q) {
      double s = 0.0;
      for (int j = 0; j < bs; ++j) s += pq[q * bs + j] * xq[q * bs + j];
      Ssum[q] += s;
    }
    jvec += bs;
  }
  double t1 = now_ms();
  std::fill(force, force + m.NV, 0.0);
  const int NV1 = m.NV1;
  for (long e = 0; e < m.ncell; ++e) {
    long cx = e % m.nc, cy = (e / m.nc) % m.nc, cz = e / ((long)m.nc * m.nc);
    if (childQ) {
      const double jxw = std::pow(0.25 * m.h, 3);  // child Gauss-2 weights are 1
      for (int iz = 0; iz < 2; ++iz) for (int iy = 0; iy < 2; ++iy) for (int ix = 0; ix < 2; ++ix) {
        double out[8] = {0};
        for (int kz = 0; kz < 2; ++kz) for (int ky = 0; ky < 2; ++ky) for (int kx = 0; kx < 2; ++kx) {
          int qx = 2 * ix + kx, qy = 2 * iy + ky, qz = 2 * iz + kz;
          double sv = Ssum[(size_t)e * 64 + (qz * 4 + qy) * 4 + qx] * jxw;
          for (int dz = 0; dz < 2; ++dz) for (int dy = 0; dy < 2; ++dy) for (int dx = 0; dx < 2; ++dx)
            out[(dz * 2 + dy) * 2 + dx] += m.cw[qx][dx] * m.cw[qy][dy] * m.cw[qz][dz] * sv;
        }
        for (int dz = 0; dz < 2; ++dz) for (int dy = 0; dy < 2; ++dy) for (int dx = 0; dx < 2; ++dx)
          force[((2 * cz + iz + dz) * NV1 + (2 * cy + iy + dy)) * (long)NV1 + (2 * cx + ix + dx)] +=
              out[(dz * 2 + dy) * 2 + dx];
      }
    } else {
      const int q1 = m.q1;
      for (int qz = 0; qz < q1; ++qz) for (int qy = 0; qy < q1; ++qy) for (int qx = 0; qx < q1; ++qx) {
        int q = (qz * q1 + qy) * q1 + qx;
        double sv = Ssum[(size_t)e * m.nq + q] * m.JxW[q];
        long ix = 2 * cx + m.qcc[qx], iy = 2 * cy + m.qcc[qy], iz = 2 * cz + m.qcc[qz];
        double wx[2] = {m.qlw0[qx], m.qlw1[qx]}, wy[2] = {m.qlw0[qy], m.qlw1[qy]},
               wz[2] = {m.qlw0[qz], m.qlw1[qz]};
        for (int dz = 0; dz < 2; ++dz) for (int dy = 0; dy < 2; ++dy) for (int dx = 0; dx < 2; ++dx)
          force[((iz + dz) * NV1 + (iy + dy)) * (long)NV1 + (ix + dx)] += wx[dx] * wy[dy] * wz[dz] * sv;
      }
    }
  }
  double t2 = now_ms();
  c->st.t_adjoint_ms = t1 - t0; c->st.t_force_ms = t2 - t1;
  c->lossPrev = c->loss;
  c->counter++;
  c->st.n_evaluations = c->counter;
  c->haveForward = false;
  return IDFT_OK;
}

}  // namespace

// ======================================================================= ABI
extern "C" {

const char* idft_backend(void) { return "cpu-oracle"; }

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
  d->minres_graph = 1;  // HIP launch option; no effect here
  d->relax_ratio = 1e3;  // mixed-precision option; no effect here
  d->minres_fused = 1;   // HIP launch option; no effect here
}

static bool desc_ok(const idft_desc* d) {
  return d && d->n_cells >= 1 && d->degree >= 1 && d->degree <= 8 && d->half_width > 0 &&
         d->n_eig >= 1;
}

int idft_geometry_sizes(const idft_desc* d, idft_sizes* s) {
  if (!desc_ok(d) || !s) return IDFT_EINVAL;
  long N1 = (long)d->n_cells * d->degree + 1, q = d->degree + 3, nv = 2L * d->n_cells + 1;
  s->n_nodes = N1 * N1 * N1; s->n_cells = (long)d->n_cells * d->n_cells * d->n_cells;
  s->nodes_per_cell = (long)(d->degree + 1) * (d->degree + 1) * (d->degree + 1);
  s->quad_per_cell = q * q * q; s->n_vxc = nv * nv * nv; s->child_quad_per_cell = 64;
  return IDFT_OK;
}

int idft_basis_tables(int32_t p, double* gll_x, double* gll_w, double* gauss_x, double* gauss_w) {
  if (p < 1 || p > 8) return IDFT_EINVAL;
  vector<double> a, b, c2, d2;
  gll_rule(p, a, b); gauss_rule(p + 3, c2, d2);
  std::copy(a.begin(), a.end(), gll_x); std::copy(b.begin(), b.end(), gll_w);
  std::copy(c2.begin(), c2.end(), gauss_x); std::copy(d2.begin(), d2.end(), gauss_w);
  return IDFT_OK;
}

int idft_geometry_quad_points(const idft_desc* d, double* xyz) {
  if (!desc_ok(d) || !xyz) return IDFT_EINVAL;
  Mesh m; m.init(*d);
  for (long e = 0; e < m.ncell; ++e) {
    long cx = e % m.nc, cy = (e / m.nc) % m.nc, cz = e / ((long)m.nc * m.nc);
    for (int qz = 0; qz < m.q1; ++qz) for (int qy = 0; qy < m.q1; ++qy) for (int qx = 0; qx < m.q1; ++qx) {
      size_t i = (size_t)e * m.nq + (qz * m.q1 + qy) * m.q1 + qx;
      xyz[3 * i + 0] = -m.L + cx * m.h + 0.5 * m.h * (m.qx[qx] + 1.0);
      xyz[3 * i + 1] = -m.L + cy * m.h + 0.5 * m.h * (m.qx[qy] + 1.0);
      xyz[3 * i + 2] = -m.L + cz * m.h + 0.5 * m.h * (m.qx[qz] + 1.0);
    }
  }
  return IDFT_OK;
}

int idft_geometry_node_points(const idft_desc* d, double* xyz) {
  if (!desc_ok(d) || !xyz) return IDFT_EINVAL;
  Mesh m; m.init(*d);
  vector<double> x1(m.N1);
  for (int c = 0; c < m.nc; ++c)
    for (int a = 0; a < m.p; ++a) x1[c * m.p + a] = -m.L + c * m.h + 0.5 * m.h * (m.gx[a] + 1.0);
  x1[m.N1 - 1] = m.L;
  for (long n = 0; n < m.N; ++n) {
    xyz[3 * n] = x1[n % m.N1]; xyz[3 * n + 1] = x1[(n / m.N1) % m.N1]; xyz[3 * n + 2] = x1[n / ((long)m.N1 * m.N1)];
  }
  return IDFT_OK;
}

int idft_geometry_vxc_points(const idft_desc* d, double* xyz) {
  if (!desc_ok(d) || !xyz) return IDFT_EINVAL;
  Mesh m; m.init(*d);
  for (long n = 0; n < m.NV; ++n) {
    long i = n % m.NV1, j = (n / m.NV1) % m.NV1, k = n / ((long)m.NV1 * m.NV1);
    xyz[3 * n] = -m.L + 0.5 * m.h * i; xyz[3 * n + 1] = -m.L + 0.5 * m.h * j; xyz[3 * n + 2] = -m.L + 0.5 * m.h * k;
  }
  return IDFT_OK;
}

int idft_create(const idft_desc* d, const double* pot_base_q, const double* rho_target_q,
                const double* weight_q, const double* rho_target_nodal, idft_ctx** out) {
  if (!out) return IDFT_EINVAL;
  *out = nullptr;
  if (!desc_ok(d) || !pot_base_q || !rho_target_q || !weight_q || !rho_target_nodal) return IDFT_EINVAL;
  if (d->precision != IDFT_FP64 && d->precision != IDFT_MIXED) return IDFT_EINVAL;
  idft_ctx* c = new idft_ctx();
  c->d = *d;
  c->m.init(*d);
  const Mesh& m = c->m;
  const long nQ = m.ncell * (long)m.nq;
  c->potBase.assign(pot_base_q, pot_base_q + nQ);
  c->rhoT.assign(rho_target_q, rho_target_q + nQ);
  c->weight.assign(weight_q, weight_q + nQ);
  c->rhoTnodal.assign(rho_target_nodal, rho_target_nodal + m.N);
  c->Mdiag.assign(m.N, 0.0); c->Kdiag.assign(m.N, 0.0);
  const double hh = 0.5 * m.h, J = hh * hh * hh;
  for (long e = 0; e < m.ncell; ++e)
    for (int cc = 0; cc < m.n1; ++cc) for (int b = 0; b < m.n1; ++b) for (int a = 0; a < m.n1; ++a) {
      int l = (cc * m.n1 + b) * m.n1 + a;
      long n = m.cellNodes[(size_t)e * m.ne + l];
      c->Mdiag[n] += J * m.gw[a] * m.gw[b] * m.gw[cc];
      c->Kdiag[n] += m.Ke[(size_t)l * m.ne + l];
    }
  c->bnd.assign(m.N, 0); c->adjMask.assign(m.N, 0);
  c->sqrtM.resize(m.N); c->invSqrtM.resize(m.N); c->S.resize(m.N); c->Djac.resize(m.N); c->DjacSqrt.resize(m.N);
  for (long n = 0; n < m.N; ++n) {
    c->bnd[n] = m.boundary(n);
    c->adjMask[n] = c->rhoTnodal[n] < d->rho_tol_constraints;
    c->sqrtM[n] = std::sqrt(c->Mdiag[n]);
    c->invSqrtM[n] = 1.0 / c->sqrtM[n];
    c->S[n] = c->bnd[n] ? 0.0 : c->invSqrtM[n];
    c->Djac[n] = 2.0 * c->Mdiag[n] / c->Kdiag[n];
    c->DjacSqrt[n] = std::sqrt(c->Djac[n]);
  }
  c->B = d->n_eig;
  c->adjointTol = d->adjoint_initial_tol;
  c->tolCheb = 1e-8;
  c->pot0.assign(m.NV, 0.0);
  c->potSol.assign(m.NV, 0.0);
  *out = c;
  return IDFT_OK;
}

void idft_destroy(idft_ctx* c) { delete c; }
const char* idft_last_error(const idft_ctx* c) { return c ? c->err.c_str() : "null context"; }

int idft_get_sizes(const idft_ctx* c, idft_sizes* s) {
  if (!c || !s) return IDFT_EINVAL;
  return idft_geometry_sizes(&c->d, s);
}
int idft_get_stats(const idft_ctx* c, idft_stats* s) {
  if (!c || !s) return IDFT_EINVAL;
  *s = c->st;
  return IDFT_OK;
}

int idft_set_potential(idft_ctx* c, const double* pot) {
  if (!c || !pot) return IDFT_EINVAL;
  vector<double> vq;
  vxc_to_quad(c, pot, vq);
  const long nQ = c->m.ncell * (long)c->m.nq;
  c->veff.resize(nQ);
  for (long q = 0; q < nQ; ++q) c->veff[q] = c->potBase[q] + vq[q];
  build_hamiltonian(c);
  return IDFT_OK;
}

int idft_set_veff_quad(idft_ctx* c, const double* vq) {
  if (!c || !vq) return IDFT_EINVAL;
  c->veff.assign(vq, vq + c->m.ncell * (long)c->m.nq);
  build_hamiltonian(c);
  return IDFT_OK;
}

int idft_get_eigenvalues(const idft_ctx* c, double* ev) {
  if (!c || !ev) return IDFT_EINVAL;
  if (!c->haveRitz || c->eig.empty()) return IDFT_ESTATE;
  std::copy(c->eig.begin(), c->eig.end(), ev);
  return IDFT_OK;
}

int idft_get_veff_quad(const idft_ctx* c, double* vq) {
  if (!c || !vq) return IDFT_EINVAL;
  if (!c->haveH) return IDFT_ESTATE;
  std::copy(c->veff.begin(), c->veff.end(), vq);
  return IDFT_OK;
}

int idft_eigensolve(idft_ctx* c, double tol, int32_t maxp, double* ev, double* mr, int32_t* passes) {
  if (!c) return IDFT_EINVAL;
  int p = 0;
  int rc = solve_passes(c, tol, maxp, &p, mr);
  if (rc) return rc;
  if (passes) *passes = p;
  if (ev) std::copy(c->eig.begin(), c->eig.end(), ev);
  return IDFT_OK;
}

int idft_get_eigenvectors(const idft_ctx* c, double* psi) {
  if (!c || !psi) return IDFT_EINVAL;
  if (!c->haveX) return IDFT_ESTATE;
  vector<double> p;
  physical_psi(c, p);
  std::copy(p.begin(), p.end(), psi);
  return IDFT_OK;
}

int idft_set_eigenvectors(idft_ctx* c, const double* psi) {
  if (!c || !psi) return IDFT_EINVAL;
  c->X.assign((size_t)c->m.N * c->B, 0.0);
  for (long n = 0; n < c->m.N; ++n)
    for (int j = 0; j < c->B; ++j)
      c->X[(size_t)n * c->B + j] = c->bnd[n] ? 0.0 : c->sqrtM[n] * psi[(size_t)n * c->B + j];
  c->haveX = true;
  c->haveRitz = false;
  c->needRR = true;
  return IDFT_OK;
}

int idft_get_occupancies(const idft_ctx* c, double* occ, double* fermi) {
  if (!c) return IDFT_EINVAL;
  if (c->occ.empty()) return IDFT_ESTATE;
  if (occ) std::copy(c->occ.begin(), c->occ.end(), occ);
  if (fermi) *fermi = c->fermi;
  return IDFT_OK;
}

int idft_get_density(idft_ctx* c, double* rho) {
  if (!c || !rho) return IDFT_EINVAL;
  if (!c->haveX || c->occ.empty()) return fail(c, IDFT_ESTATE, "no eigenvectors/occupancies");
  vector<double> r;
  density(c, r);
  std::copy(r.begin(), r.end(), rho);
  return IDFT_OK;
}

int idft_apply_hamiltonian(idft_ctx* c, const double* x, double* y, int32_t ncols, int32_t prec) {
  if (!c || !x || !y || ncols < 1) return IDFT_EINVAL;
  if (!c->haveH) return fail(c, IDFT_ESTATE, "no Hamiltonian");
  (void)prec;
  apply_cells(c, c->He, x, y, ncols, true);
  return IDFT_OK;
}

int idft_chebyshev_filter(idft_ctx* c, double* x, int32_t ncols, int32_t deg, double a, double b,
                          double a0, int32_t prec) {
  if (!c || !x || ncols < 1 || deg < 1) return IDFT_EINVAL;
  if (!c->haveH) return fail(c, IDFT_ESTATE, "no Hamiltonian");
  (void)prec;
  chebyshev_filter(c, x, ncols, deg, a, b, a0);
  return IDFT_OK;
}

int idft_mv_create(idft_ctx* c, int32_t ncols, idft_mv** out) {
  if (!c || !out || ncols < 1) return IDFT_EINVAL;
  idft_mv* v = new idft_mv();
  v->ncols = ncols;
  v->v.assign((size_t)c->m.N * ncols, 0.0);
  *out = v;
  return IDFT_OK;
}
void idft_mv_destroy(idft_mv* v) { delete v; }
int idft_mv_set(idft_ctx* c, idft_mv* v, const double* h) {
  if (!c || !v || !h) return IDFT_EINVAL;
  std::copy(h, h + (size_t)c->m.N * v->ncols, v->v.begin());
  return IDFT_OK;
}
int idft_mv_get(idft_ctx* c, const idft_mv* v, double* h) {
  if (!c || !v || !h) return IDFT_EINVAL;
  std::copy(v->v.begin(), v->v.end(), h);
  return IDFT_OK;
}

int idft_adj_update_input_psi(idft_ctx* c, idft_mv* psi, const double* occ, const double* u_q,
                              const int32_t* pairs, int32_t np, double fermi, double* eigv,
                              int32_t block) {
  if (!c || !psi || !occ || !u_q || (!pairs && np) || !eigv) return IDFT_EINVAL;
  return update_input_psi(c, psi, occ, u_q, pairs, np, fermi, eigv, block);
}
int idft_adj_compute_rhs(idft_ctx* c, idft_mv* ndbc, idft_mv* x, idft_mv* rhs) {
  if (!c || !x || !rhs) return IDFT_EINVAL;
  (void)ndbc;
  return adj_compute_rhs(c, x, rhs);
}
int idft_adj_vmult(idft_ctx* c, idft_mv* Ax, idft_mv* x, int32_t block) {
  if (!c || !Ax || !x) return IDFT_EINVAL;
  if (block != c->adj.B) return fail(c, IDFT_EINVAL, "block size mismatch");
  return adj_vmult(c, Ax, x);
}
int idft_adj_precondition_jacobi(idft_ctx* c, idft_mv* dst, const idft_mv* src, double) {
  if (!c || !dst || !src || dst->ncols != src->ncols) return IDFT_EINVAL;
  precondition(c, c->Djac, dst, src);
  return IDFT_OK;
}
int idft_adj_precondition_jacobi_sqrt(idft_ctx* c, idft_mv* dst, const idft_mv* src, double) {
  if (!c || !dst || !src || dst->ncols != src->ncols) return IDFT_EINVAL;
  precondition(c, c->DjacSqrt, dst, src);
  return IDFT_OK;
}
int idft_adj_distribute_x(idft_ctx* c) {
  if (!c) return IDFT_EINVAL;
  return adj_distribute_x(c);
}
int idft_adj_dot_quad_wise(idft_ctx* c, idft_mv* v1, idft_mv* v2, double* out) {
  if (!c || !v1 || !v2 || !out) return IDFT_EINVAL;
  if (!c->adj.ready) return fail(c, IDFT_ESTATE, "no adjoint problem");
  vector<double> o;
  dot_quad_wise(c, v1->v.data(), v2->v.data(), o);
  std::copy(o.begin(), o.end(), out);
  return IDFT_OK;
}
int idft_minres_solve(idft_ctx* c, idft_mv* x, idft_mv* ndbc, double tol, int32_t maxit,
                      int32_t distribute, int32_t* iters) {
  if (!c || !x) return IDFT_EINVAL;
  if (!c->adj.ready) return fail(c, IDFT_ESTATE, "no adjoint problem");
  c->st.minres_iters_max = 0; c->st.minres_iters_total = 0;
  return minres(c, x, ndbc, tol, maxit, distribute, iters);
}

int idft_forward(idft_ctx* c, const double* pot, double* loss) {
  if (!c || !pot) return IDFT_EINVAL;
  return forward(c, pot, loss);
}
int idft_backward(idft_ctx* c, double* force) {
  if (!c || !force) return IDFT_EINVAL;
  return backward(c, force);
}

int idft_get_force(idft_ctx* c, double* pot, double* force, double* loss) {
  if (!c || !pot || !force || !loss) return IDFT_EINVAL;
  double t0 = now_ms();
  int rc = forward(c, pot, loss);
  if (rc) return rc;
  rc = backward(c, force);
  c->st.t_total_ms = now_ms() - t0;
  return rc;
}

int idft_dot(idft_ctx* c, const double* v1, const double* v2, int64_t n, int32_t block, double* out) {
  if (!c || !v1 || !v2 || !out || block < 1) return IDFT_EINVAL;
  for (int i = 0; i < block; ++i) out[i] = 0.0;
  for (int64_t i = 0; i < n; ++i) out[i % block] += v1[i] * v2[i];
  return IDFT_OK;
}

int idft_set_initial_guess(idft_ctx* c, const double* pot) {
  if (!c || !pot) return IDFT_EINVAL;
  c->pot0.assign(pot, pot + c->m.NV);
  return IDFT_OK;
}
int idft_get_initial_guess(const idft_ctx* c, double* pot) {
  if (!c || !pot) return IDFT_EINVAL;
  std::copy(c->pot0.begin(), c->pot0.end(), pot);
  return IDFT_OK;
}
int idft_set_solution(idft_ctx* c, const double* pot) {
  if (!c || !pot) return IDFT_EINVAL;
  c->potSol.assign(pot, pot + c->m.NV);
  return IDFT_OK;
}
int idft_get_solution(const idft_ctx* c, double* pot) {
  if (!c || !pot) return IDFT_EINVAL;
  std::copy(c->potSol.begin(), c->potSol.end(), pot);
  return IDFT_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ L-BFGS
// BFGSInverseDFTSolver (BFGSInverseDFTSolver.cpp:66-627), one component.
namespace {

struct Bfgs {
  idft_ctx* c;
  long n;
  int historySize;
  vector<double> x, g, p;
  std::vector<vector<double>> Y, S;
  std::vector<double> rho;
  int k = 0;
  int err = IDFT_OK;

  double dot(const vector<double>& a, const vector<double>& b) {
    double o = 0.0;
    idft_dot(c, a.data(), b.data(), n, 1, &o);
    return o;
  }
  // two-loop recursion (.cpp:66-118)
  void inv_jac_times(const vector<double>& gin, vector<double>& z) {
    vector<double> q = gin;
    double gamma = 1.0;
    const int N = k;
    if (!Y.empty()) gamma = dot(S.back(), Y.back()) / dot(Y.back(), Y.back());
    vector<double> alpha(N, 0.0), beta(N, 0.0);
    for (int i = N - 1; i >= 0; --i) {
      alpha[i] = rho[i] * dot(S[i], q);
      for (long t = 0; t < n; ++t) q[t] += -alpha[i] * Y[i][t];
    }
    z = q;
    for (long t = 0; t < n; ++t) z[t] *= gamma;
    for (int i = 0; i < N; ++i) {
      beta[i] = rho[i] * dot(Y[i], z);
      for (long t = 0; t < n; ++t) z[t] += (alpha[i] - beta[i]) * S[i][t];
    }
  }
  bool eval(const vector<double>& xin, vector<double>& gout, double& L) {
    vector<double> xx = xin;
    gout.assign(n, 0.0);
    int rc = idft_get_force(c, xx.data(), gout.data(), &L);
    if (rc) { err = rc; return false; }
    return true;
  }
  bool fnorm(int type, double lam, double& f) {
    vector<double> xn(n), gn;
    for (long t = 0; t < n; ++t) xn[t] = x[t] + lam * p[t];
    double L;
    if (!eval(xn, gn, L)) return false;
    if (type == IDFT_LS_SECANT_LOSS) f = L;
    else if (type == IDFT_LS_SECANT_FORCE_NORM) f = std::sqrt(dot(gn, gn));
    else f = dot(gn, p);
    return true;
  }
  // solveLineSearchCP (.cpp:215-303)
  bool ls_cp(vector<double>& lam, vector<double>& f, int maxIter, double tol) {
    for (int i = 0; i < maxIter; ++i) {
      int N = (int)lam.size(), M = (int)f.size();
      double l0 = lam[N - 2], l1 = lam[N - 1], f0, f1;
      if (M > 0) f0 = f[M - 1];
      else { if (!fnorm(IDFT_LS_CP, l0, f0)) return false; f.push_back(f0); }
      if (std::sqrt(std::fabs(f0)) < tol) { lam.pop_back(); break; }
      if (!fnorm(IDFT_LS_CP, l1, f1)) return false;
      f.push_back(f1);
      if (std::sqrt(std::fabs(f1)) < tol) break;
      if (std::fabs((f1 - f0) / f1) < 1e-15) { c->err = "secant line search failed: f'(x) = 0"; err = IDFT_ENOCONV; return false; }
      double sl = (f1 - f0) / (l1 - l0);
      if (sl > 0.0) sl = -sl;
      double ln = l1 - f1 / sl;
      if (ln < 0.0) ln = l1 + f1 / sl;
      lam.push_back(ln);
    }
    return true;
  }
  // solveLineSearchSecantLoss / SecantForceNorm (.cpp:305-463)
  bool ls_secant(int type, vector<double>& lam, vector<double>& f, int maxIter) {
    double lOld = lam[0], lNew = lam[1], lMid = 0.5 * (lNew + lOld), fOld = f[0], fMid, fNew;
    for (int i = 0; i < maxIter; ++i) {
      if (!fnorm(type, lMid, fMid)) return false;
      if (!fnorm(type, lNew, fNew)) return false;
      lNew = 0.5 * (lNew + lOld);
      lMid = 0.5 * (lNew + lOld);
      double dl = lNew - lOld;
      double dF = (3. * fNew - 4. * fMid + fOld) / dl;
      double dFOld = (-3. * fOld + 4. * fMid - fNew) / dl;
      double d2F = (dF - dFOld) / dl;
      double lu;
      if (d2F > 0.) lu = lNew - dF / d2F;
      else if (d2F < 0.) lu = lNew + dF / d2F;
      else break;
      if (lu < 0.0) lu = 0.5 * (lNew + lOld);
      lOld = lNew; lNew = lu; fOld = fNew;
      lMid = 0.5 * (lNew + lOld);
      lam.push_back(lNew); f.push_back(fNew);
    }
    return true;
  }
};

}  // namespace

extern "C" int idft_bfgs_solve(idft_ctx* c, int32_t ls_type, int32_t max_iter, double tol,
                               double ls_tol, int32_t history, int32_t n_ls, double* loss_hist,
                               double* gnorm_hist, double* time_hist, int32_t* n_done) {
  const double tStart = now_ms();
  if (!c || max_iter < 0 || history < 1) return IDFT_EINVAL;
  if (ls_type < IDFT_LS_SECANT_LOSS || ls_type > IDFT_LS_CP) return fail(c, IDFT_EINVAL, "invalid line search type");
  Bfgs b;
  b.c = c; b.n = c->m.NV; b.historySize = history;
  b.x = c->pot0;
  b.g.assign(b.n, 0.0); b.p.assign(b.n, 0.0);
  vector<double> y(b.n, 0.0), s(b.n, 0.0);
  double L = 0.0;
  int iter = 0;
  for (; iter < max_iter; ++iter) {
    if (iter == 0 && !b.eval(b.x, b.g, L)) return b.err;
    double gn = std::sqrt(b.dot(b.g, b.g));
    if (time_hist) time_hist[iter] = now_ms() - tStart;
    if (loss_hist) loss_hist[iter] = L;
    if (gnorm_hist) gnorm_hist[iter] = gn;
    bool conv = gn < tol;
    b.inv_jac_times(b.g, b.p);
    for (long t = 0; t < b.n; ++t) b.p[t] = -b.p[t];
    double f0 = ls_type == IDFT_LS_SECANT_LOSS ? L : ls_type == IDFT_LS_SECANT_FORCE_NORM ? gn : b.dot(b.g, b.p);
    vector<double> lam = {0.0, 1.0}, fs = {f0};
    if (conv) break;
    bool ok = ls_type == IDFT_LS_CP ? b.ls_cp(lam, fs, n_ls, ls_tol) : b.ls_secant(ls_type, lam, fs, n_ls);
    if (!ok) return b.err;
    double lamOpt = lam.back() > 0.0 ? lam.back() : 1.0;
    for (long t = 0; t < b.n; ++t) { s[t] = lamOpt * b.p[t]; b.x[t] += s[t]; y[t] = -b.g[t]; }
    if (!b.eval(b.x, b.g, L)) return b.err;
    for (long t = 0; t < b.n; ++t) y[t] += b.g[t];
    double curv = b.dot(y, s);
    if (curv > 0.0) { b.Y.push_back(y); b.S.push_back(s); b.rho.push_back(1.0 / curv); b.k++; }
    if (b.k > history) { b.k = history; b.Y.erase(b.Y.begin()); b.S.erase(b.S.begin()); b.rho.erase(b.rho.begin()); }
  }
  if (iter == max_iter && loss_hist) {
    loss_hist[iter] = L;
    if (gnorm_hist) gnorm_hist[iter] = std::sqrt(b.dot(b.g, b.g));
  }
  if (time_hist) time_hist[iter] = now_ms() - tStart;
  c->potSol = b.x;
  if (n_done) *n_done = iter;
  return IDFT_OK;
}

// ============================================================ set-up stage
// include/idft_setup.h on the CPU (test infrastructure, like the rest of this
// file): the atomic-orbital target density as gaussianFunctionManager.cpp
// evaluates it (spherical coordinates, associated Legendre functions), the
// loss weights, the Poisson solve and the initial v_xc guess with assembled
// cell matrices and plain CG.
#include "../include/idft_setup.h"

#include <fstream>
#include <sstream>

struct idft_ao_basis {
  struct Fn {
    double c[3];
    int l = 0, m = 0;
    vector<double> alpha, cn;  // exponents, coefficient * primitive norm
    double norm = 1.0;         // basisNormConst
  };
  vector<Fn> f;
};

namespace {

const double kPi = 3.14159265358979323846;
// DFT-FE's Angstrom -> Bohr conversion (C_AngToBohr).
const double kAngToBohr = 1.0 / 0.52917721067;

double dfact(int n) {  // doubleFactorial (gaussianFunctionManager.cpp:133-137)
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}
// getNormConsts (:251-262)
double prim_norm(double a, int l) {
  const double ov = dfact(2 * l + 1) * std::sqrt(kPi) / (std::pow(2.0, 2 * l + 3.5) * std::pow(a, l + 1.5));
  return 1.0 / std::sqrt(ov);
}
// Blm / Clm (:271-287) and Dm (:264-269)
double blm(int l, int m) { return m == 0 ? std::sqrt((2.0 * l + 1) / 2.0) : blm(l, m - 1) / std::sqrt((l - m + 1.0) * (l + m)); }
double dm(int m) { return m == 0 ? 1.0 / std::sqrt(2 * kPi) : 1.0 / std::sqrt(kPi); }
// Associated Legendre P_l^m(x), m >= 0, without the Condon-Shortley phase
// (Plm, :310-320), by the standard upward recurrence in l.
double plm(int l, int m, double x) {
  if (m > l) return 0.0;
  const double s = std::sqrt(std::max(0.0, 1.0 - x * x));
  double pmm = dfact(2 * m - 1) * std::pow(s, m);
  if (l == m) return pmm;
  double pm1 = x * (2 * m + 1) * pmm;
  if (l == m + 1) return pm1;
  double pl = 0.0;
  for (int k = m + 2; k <= l; ++k) {
    pl = (x * (2 * k - 1) * pm1 - (k + m - 1) * pmm) / (k - m);
    pmm = pm1;
    pm1 = pl;
  }
  return pl;
}
double fact(int n) { double r = 1.0; for (int k = 2; k <= n; ++k) r *= k; return r; }
// dPlmDTheta (:322-350)
double dplm_dtheta(int l, int m, double theta) {
  const double ct = std::cos(theta);
  if (std::abs(m) > l || l == 0) return 0.0;
  if (m < 0) {
    const int am = -m;
    return std::pow(-1.0, m) * fact(l - am) / fact(l + am) * dplm_dtheta(l, am, theta);
  }
  if (m == 0) return -plm(l, 1, ct);
  if (m == l) return l * plm(l, l - 1, ct);
  return 0.5 * ((l + m) * (l - m + 1) * plm(l, m - 1, ct) - plm(l, m + 1, ct));
}
// Qm, dQm/dphi (:289-308)
double qm(int m, double phi) { return m > 0 ? std::cos(m * phi) : m == 0 ? 1.0 : std::sin(-m * phi); }
double dqm(int m, double phi) { return m > 0 ? -m * std::sin(m * phi) : m == 0 ? 0.0 : -m * std::cos(-m * phi); }
// convertCartesianToSpherical (:147-172)
void to_spherical(const double* d, double& r, double& th, double& ph) {
  r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  th = 0.0; ph = 0.0;
  if (r == 0) return;
  th = std::acos(d[2] / r);
  if (std::fabs(th) >= 1e-12 && std::fabs(th - kPi) >= 1e-12) ph = std::atan2(d[1], d[0]);
}
// gfRadialPartDerivative (:385-394)
double radial_der(double r, double a, int l, int k) {
  if (k == 0) return l >= 0 ? std::pow(r, l) * std::exp(-a * r * r) : 0.0;
  return l * radial_der(r, a, l - 1, k - 1) - 2 * a * radial_der(r, a, l + 1, k - 1);
}
// evaluateBasisValue (:416-446)
double ao_value(const idft_ao_basis::Fn& b, const double* x) {
  const double d[3] = {x[0] - b.c[0], x[1] - b.c[1], x[2] - b.c[2]};
  double r, th, ph;
  to_spherical(d, r, th, ph);
  double rad = 0.0;
  for (size_t k = 0; k < b.alpha.size(); ++k) rad += b.cn[k] * std::pow(r, b.l) * std::exp(-b.alpha[k] * r * r);
  const int am = std::abs(b.m);
  return rad * blm(b.l, am) * dm(b.m) * plm(b.l, am, std::cos(th)) * qm(b.m, ph) * b.norm;
}
// getSphericalGaussianGradient (:449-576) summed over the contraction
// (getBasisFunctionGradient, :1633-1663).
void ao_gradient(const idft_ao_basis::Fn& b, const double* x, double* g) {
  const double d[3] = {x[0] - b.c[0], x[1] - b.c[1], x[2] - b.c[2]};
  double r, th, ph;
  to_spherical(d, r, th, ph);
  const int l = b.l, m = b.m, am = std::abs(m);
  const double C = blm(l, am) * dm(m);
  g[0] = g[1] = g[2] = 0.0;
  for (size_t k = 0; k < b.alpha.size(); ++k) {
    const double a = b.alpha[k];
    double v[3] = {0, 0, 0};
    if (r < 1e-15) {
      if (l == 1) v[m == 1 ? 0 : m == -1 ? 1 : 2] = C;
    } else if (std::fabs(th) < 1e-12 || std::fabs(th - kPi) < 1e-12) {
      const double R = radial_der(r, a, l, 0), dR = radial_der(r, a, l, 1);
      const double ct = std::cos(th), P = plm(l, am, ct);
      const double sgn = std::fabs(th) < 1e-12 ? 1.0 : std::pow(-1.0, l + 1);
      if (m == 0) v[2] = C * dR * P * ct;
      else if (m == 1) v[0] = C * (R / r) * l * (l + 1) / 2.0 * sgn;
      else if (m == -1) v[1] = C * (R / r) * l * (l + 1) / 2.0 * sgn;
    } else {
      const double R = radial_der(r, a, l, 0), dR = radial_der(r, a, l, 1);
      const double ct = std::cos(th), st = std::sin(th);
      const double P = plm(l, am, ct), dP = dplm_dtheta(l, am, th), Q = qm(m, ph), dQ = dqm(m, ph);
      const double J[3][3] = {{st * std::cos(ph), ct * std::cos(ph) / r, -std::sin(ph) / (r * st)},
                              {st * std::sin(ph), ct * std::sin(ph) / r, std::cos(ph) / (r * st)},
                              {ct, -st / r, 0.0}};
      const double pd[3] = {dR * P * Q, R * dP * Q, R * P * dQ};
      for (int i = 0; i < 3; ++i) v[i] = C * (J[i][0] * pd[0] + J[i][1] * pd[1] + J[i][2] * pd[2]);
    }
    for (int i = 0; i < 3; ++i) g[i] += b.cn[k] * v[i];
  }
  for (int i = 0; i < 3; ++i) g[i] *= b.norm;
}

bool is_number(const std::string& w, double& v) {
  char* end = nullptr;
  v = std::strtod(w.c_str(), &end);
  return end && *end == '\0' && !w.empty();
}

std::string join_dir(const char* dir, const std::string& f) {
  if (!dir || !*dir || (!f.empty() && f[0] == '/')) return f;
  return std::string(dir) + "/" + f;
}

// readGaussianFiles (:720-825): shells of one basis file.
struct Shell { int l; vector<double> a, c; };
bool read_shells(const std::string& path, vector<Shell>& out) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  std::getline(in, line);  // first line ignored
  const std::string lch = "SPDFGHI";
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string w;
    if (!(ls >> w)) continue;
    for (auto& ch : w) ch = (char)std::toupper(ch);
    if (w.find_first_not_of(lch) != std::string::npos) return false;
    std::string ns;
    if (!(ls >> ns)) return false;
    char* end = nullptr;
    const long n = std::strtol(ns.c_str(), &end, 10);
    if (!end || *end != '\0' || n < 1) return false;
    vector<double> a(n);
    vector<vector<double>> c(n, vector<double>(w.size()));
    for (long i = 0; i < n; ++i) {
      if (!std::getline(in, line) || line.empty()) return false;
      std::istringstream lc(line);
      std::string t;
      if (!(lc >> t) || !is_number(t, a[i])) return false;
      for (size_t j = 0; j < w.size(); ++j)
        if (!(lc >> t) || !is_number(t, c[i][j])) return false;
    }
    for (size_t j = 0; j < w.size(); ++j) {
      Shell s;
      s.l = (int)lch.find(w[j]);
      s.a = a;
      for (long i = 0; i < n; ++i) s.c.push_back(c[i][j]);
      out.push_back(s);
    }
  }
  return true;
}

// Parent quadrature point coordinates of cell e, point q.
void quad_xyz(const Mesh& m, long e, int q, double* x) {
  const long cx = e % m.nc, cy = (e / m.nc) % m.nc, cz = e / ((long)m.nc * m.nc);
  const int qx = q % m.q1, qy = (q / m.q1) % m.q1, qz = q / (m.q1 * m.q1);
  x[0] = -m.L + cx * m.h + 0.5 * m.h * (m.qx[qx] + 1.0);
  x[1] = -m.L + cy * m.h + 0.5 * m.h * (m.qx[qy] + 1.0);
  x[2] = -m.L + cz * m.h + 0.5 * m.h * (m.qx[qz] + 1.0);
}
void node_xyz(const Mesh& m, long n, double* x) {
  const long i = n % m.N1, j = (n / m.N1) % m.N1, k = n / ((long)m.N1 * m.N1);
  const int ai = (int)(i % m.p), aj = (int)(j % m.p), ak = (int)(k % m.p);
  const long ci = i / m.p, cj = j / m.p, ck = k / m.p;
  x[0] = -m.L + ci * m.h + 0.5 * m.h * (m.gx[ai] + 1.0);
  x[1] = -m.L + cj * m.h + 0.5 * m.h * (m.gx[aj] + 1.0);
  x[2] = -m.L + ck * m.h + 0.5 * m.h * (m.gx[ak] + 1.0);
}

// y = sum_e P_e^T A_e P_e x over all nodes (A_e: [ne][ne]) with rows and
// columns of boundary nodes dropped (zero) when `interior`.
void cell_apply(const Mesh& m, const vector<double>& Ae, const vector<double>& x, vector<double>& y,
                bool interior) {
  const int ne = m.ne;
  y.assign(m.N, 0.0);
  for (int col = 0; col < 8; ++col) {
#pragma omp parallel for schedule(static)
    for (long e = 0; e < m.ncell; ++e) {
      if (m.colour(e) != col) continue;
      const long* nodes = &m.cellNodes[(size_t)e * ne];
      for (int a = 0; a < ne; ++a) {
        if (interior && m.boundary(nodes[a])) continue;
        double s = 0.0;
        for (int b = 0; b < ne; ++b)
          if (!(interior && m.boundary(nodes[b]))) s += Ae[(size_t)a * ne + b] * x[nodes[b]];
        y[nodes[a]] += s;
      }
    }
  }
}
// f_a = sum_e sum_q N_a(q) w[e][q]  (integration against the parent basis)
void integrate_basis(const Mesh& m, const vector<double>& w, vector<double>& f) {
  const int ne = m.ne, nq = m.nq;
  f.assign(m.N, 0.0);
  for (int col = 0; col < 8; ++col) {
#pragma omp parallel for schedule(static)
    for (long e = 0; e < m.ncell; ++e) {
      if (m.colour(e) != col) continue;
      const long* nodes = &m.cellNodes[(size_t)e * ne];
      for (int a = 0; a < ne; ++a) {
        double s = 0.0;
        for (int q = 0; q < nq; ++q) s += m.Nq[(size_t)q * ne + a] * w[(size_t)e * nq + q];
        f[nodes[a]] += s;
      }
    }
  }
}
// Preconditioned CG on the interior nodes: A x = b, D the diagonal
// preconditioner, x0 = 0; stops at ||r||_D <= tol ||b||_D.
template <typename Apply>
int pcg(const Mesh& m, Apply&& A, const vector<double>& b, const vector<double>& D, double tol,
        int maxit, vector<double>& x, double* relres) {
  const long N = m.N;
  x.assign(N, 0.0);
  vector<double> r = b, z(N), p(N), q;
  auto dot = [&](const vector<double>& u, const vector<double>& v) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s)
    for (long i = 0; i < N; ++i) s += u[i] * v[i];
    return s;
  };
  for (long i = 0; i < N; ++i) z[i] = D[i] * r[i];
  p = z;
  double rz = dot(r, z);
  const double bn = std::sqrt(std::max(rz, 0.0));
  int it = 0;
  double rn = bn;
  while (it < maxit && rn > tol * bn && bn > 0.0) {
    A(p, q);
    const double alpha = rz / dot(p, q);
    for (long i = 0; i < N; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    for (long i = 0; i < N; ++i) z[i] = D[i] * r[i];
    const double rz2 = dot(r, z);
    rn = std::sqrt(std::max(rz2, 0.0));
    const double beta = rz2 / rz;
    rz = rz2;
    for (long i = 0; i < N; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  if (relres) *relres = bn > 0.0 ? rn / bn : 0.0;
  return it;
}
// Cell mass matrix M_e = sum_q N_a N_b JxW (consistent mass at the parent quadrature).
vector<double> cell_mass(const Mesh& m) {
  const int ne = m.ne, nq = m.nq;
  vector<double> Me((size_t)ne * ne, 0.0);
  for (int q = 0; q < nq; ++q)
    for (int a = 0; a < ne; ++a)
      for (int b = 0; b < ne; ++b) Me[(size_t)a * ne + b] += m.Nq[(size_t)q * ne + a] * m.Nq[(size_t)q * ne + b] * m.JxW[q];
  return Me;
}
// out[e][q] = sum_a N_a(q) x[node_a]
void nodal_to_quad(const Mesh& m, const vector<double>& x, double* out) {
  const int ne = m.ne, nq = m.nq;
#pragma omp parallel for schedule(static)
  for (long e = 0; e < m.ncell; ++e) {
    const long* nodes = &m.cellNodes[(size_t)e * ne];
    for (int q = 0; q < nq; ++q) {
      double s = 0.0;
      for (int a = 0; a < ne; ++a) s += m.Nq[(size_t)q * ne + a] * x[nodes[a]];
      out[(size_t)e * nq + q] = s;
    }
  }
}
vector<double> diag_of(const Mesh& m, const vector<double>& Ae) {
  vector<double> d(m.N, 0.0);
  for (long e = 0; e < m.ncell; ++e)
    for (int a = 0; a < m.ne; ++a) d[m.cellNodes[(size_t)e * m.ne + a]] += Ae[(size_t)a * m.ne + a];
  return d;
}

// ---- exchange-correlation potentials (libxc is not in the reference; the
// published forms are restated).
// Slater exchange of the total density n: v_x = -(3 n / pi)^{1/3}.
double slater_vx(double n) { return n > 1e-30 ? -std::cbrt(3.0 * n / kPi) : 0.0; }
// van Leeuwen-Baerends 1994 exchange potential (beta = 0.05) of the total
// density n with sigma = |grad n|^2, per spin channel n/2.
double lb94_vx(double n, double sigma) {
  if (n <= 1e-30) return 0.0;
  const double beta = 0.05, ns = 0.5 * n;
  const double x = 0.5 * std::sqrt(std::max(sigma, 0.0)) / std::pow(ns, 4.0 / 3.0);
  return slater_vx(n) - beta * std::cbrt(ns) * x * x / (1.0 + 3.0 * beta * x * std::asinh(x));
}
// Perdew-Wang 1992 correlation (spin-unpolarised, Table I parameters of the
// paper): eps_c(r_s) = -2A(1 + a1 r_s) ln(1 + 1/(2A(b1 rs^1/2 + b2 rs + b3 rs^3/2 + b4 rs^2))),
// v_c = eps_c - (r_s/3) d eps_c / d r_s.
double pw92_vc(double n) {
  if (n <= 1e-30) return 0.0;
  const double A = 0.031091, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
  const double rs = std::cbrt(3.0 / (4.0 * kPi * n)), srs = std::sqrt(rs);
  const double Q0 = -2.0 * A * (1.0 + a1 * rs);
  const double Q1 = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  const double Q1p = A * (b1 / srs + 2.0 * b2 + 3.0 * b3 * srs + 4.0 * b4 * rs);
  const double lg = std::log(1.0 + 1.0 / Q1);
  const double eps = Q0 * lg;
  const double deps = -2.0 * A * a1 * lg - Q0 * Q1p / (Q1 * Q1 + Q1);
  return eps - rs / 3.0 * deps;
}

}  // namespace

extern "C" {

int idft_ao_basis_read(const char* coords_file, const char* basis_dir, int32_t angstrom,
                       idft_ao_basis** out) {
  if (!coords_file || !out) return IDFT_EINVAL;
  *out = nullptr;
  std::ifstream in(coords_file);
  if (!in) return IDFT_EINVAL;
  auto* b = new idft_ao_basis();
  std::map<std::string, vector<Shell>> cache;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string sym, xs, ys, zs, file;
    if (!(ls >> sym)) continue;
    double c[3];
    if (!(ls >> xs >> ys >> zs >> file) || !is_number(xs, c[0]) || !is_number(ys, c[1]) ||
        !is_number(zs, c[2])) { delete b; return IDFT_EINVAL; }
    if (angstrom) for (double& v : c) v *= kAngToBohr;
    const std::string path = join_dir(basis_dir, file);
    if (!cache.count(path)) {
      vector<Shell> sh;
      if (!read_shells(path, sh)) { delete b; return IDFT_EINVAL; }
      cache[path] = sh;
    }
    for (const Shell& s : cache[path]) {
      vector<int> ms;
      if (s.l == 1) ms = {1, -1, 0};
      else for (int mm = -s.l; mm <= s.l; ++mm) ms.push_back(mm);
      for (int mm : ms) {
        idft_ao_basis::Fn f;
        std::copy(c, c + 3, f.c);
        f.l = s.l; f.m = mm; f.alpha = s.a;
        for (size_t k = 0; k < s.a.size(); ++k) f.cn.push_back(s.c[k] * prim_norm(s.a[k], s.l));
        b->f.push_back(f);
      }
    }
  }
  if (b->f.empty()) { delete b; return IDFT_EINVAL; }
  *out = b;
  return IDFT_OK;
}

void idft_ao_basis_destroy(idft_ao_basis* b) { delete b; }

int idft_ao_basis_size(const idft_ao_basis* b, int32_t* n) {
  if (!b || !n) return IDFT_EINVAL;
  *n = (int32_t)b->f.size();
  return IDFT_OK;
}

int idft_ao_read_matrix(const char* path, int32_t n, double* out) {
  if (!path || !out || n < 1) return IDFT_EINVAL;
  std::ifstream in(path);
  if (!in) return IDFT_EINVAL;
  std::string line, w;
  int row = 0;
  while (std::getline(in, line)) {
    if (line.empty()) return IDFT_EINVAL;
    std::istringstream ls(line);
    int col = 0;
    while (ls >> w) {
      double v;
      if (!is_number(w, v) || row >= n || col >= n) return IDFT_EINVAL;
      out[(size_t)row * n + col++] = v;
    }
    if (col != n) return IDFT_EINVAL;
    ++row;
  }
  return row == n ? IDFT_OK : IDFT_EINVAL;
}

int idft_ao_values(const idft_ao_basis* b, int64_t np, const double* xyz, double* vals) {
  if (!b || !xyz || !vals || np < 0) return IDFT_EINVAL;
  const size_t nb = b->f.size();
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)nb; ++i)
    for (int64_t p = 0; p < np; ++p) vals[(size_t)i * np + p] = ao_value(b->f[i], xyz + 3 * p);
  return IDFT_OK;
}

int idft_ao_density_quad(const idft_desc* d, idft_ao_basis* b, int32_t normalize, int32_t n_spin,
                         const double* P, double* rho_q, double* grad_q, double* s_diag) {
  if (!desc_ok(d) || !b || !P || !rho_q || n_spin < 1 || n_spin > 2) return IDFT_EINVAL;
  Mesh m; m.init(*d);
  const long nQ = m.ncell * m.nq;
  const int nb = (int)b->f.size();
  for (auto& f : b->f) f.norm = 1.0;
  // self-overlaps on this quadrature (evaluateForQuad, :1252-1305)
  vector<double> S(nb, 0.0);
#pragma omp parallel
  {
    vector<double> loc(nb, 0.0);
#pragma omp for schedule(static)
    for (long i = 0; i < nQ; ++i) {
      double x[3];
      quad_xyz(m, i / m.nq, (int)(i % m.nq), x);
      const double w = m.JxW[i % m.nq];
      for (int k = 0; k < nb; ++k) { const double v = ao_value(b->f[k], x); loc[k] += v * v * w; }
    }
#pragma omp critical
    for (int k = 0; k < nb; ++k) S[k] += loc[k];
  }
  if (s_diag) std::copy(S.begin(), S.end(), s_diag);
  if (normalize)
    for (int k = 0; k < nb; ++k) b->f[k].norm = 1.0 / std::sqrt(S[k]);
  // getRhoValue (:1387-1430) and getRhoGradient (:1481-1516)
#pragma omp parallel
  {
    vector<double> phi(nb), gphi(3 * (size_t)nb);
#pragma omp for schedule(static)
    for (long i = 0; i < nQ; ++i) {
      double x[3];
      quad_xyz(m, i / m.nq, (int)(i % m.nq), x);
      for (int k = 0; k < nb; ++k) {
        phi[k] = ao_value(b->f[k], x);
        if (grad_q) ao_gradient(b->f[k], x, &gphi[3 * k]);
      }
      for (int s = 0; s < n_spin; ++s) {
        const double* Ps = P + (size_t)s * nb * nb;
        double val = 0.0, g[3] = {0, 0, 0};
        for (int a = 0; a < nb; ++a) {
          val += Ps[(size_t)a * nb + a] * phi[a] * phi[a];
          for (int c = a + 1; c < nb; ++c) val += 2.0 * Ps[(size_t)a * nb + c] * phi[a] * phi[c];
          if (grad_q)
            for (int c = 0; c < nb; ++c) {
              const double pac = Ps[(size_t)std::min(a, c) * nb + std::max(a, c)];
              for (int k = 0; k < 3; ++k) g[k] += pac * (phi[a] * gphi[3 * c + k] + phi[c] * gphi[3 * a + k]);
            }
        }
        rho_q[(size_t)s * nQ + i] = val;
        if (grad_q)
          for (int k = 0; k < 3; ++k) grad_q[((size_t)s * nQ + i) * 3 + k] = g[k];
      }
    }
  }
  return IDFT_OK;
}

int idft_loss_weights(int64_t n, const double* rho, double sf, double a1, double a2, double tau,
                      double* w) {
  if (n < 0 || !rho || !w) return IDFT_EINVAL;
  for (int64_t i = 0; i < n; ++i) {
    const double r = sf * rho[i];
    w[i] = 1.0 / (std::pow(r, a1) + tau) + std::pow(r, a2);
  }
  return IDFT_OK;
}

int idft_poisson_solve(const idft_desc* d, const double* cq, const double* bc, double tol,
                       int32_t maxit, double* phi_q, double* phi_nodal, idft_poisson_stats* st) {
  if (!desc_ok(d) || !cq || !(tol > 0) || maxit < 1) return IDFT_EINVAL;
  Mesh m; m.init(*d);
  const long nQ = m.ncell * m.nq;
  // charge moments about the origin
  double Q = 0.0, dip[3] = {0, 0, 0};
  for (long i = 0; i < nQ; ++i) {
    double x[3];
    quad_xyz(m, i / m.nq, (int)(i % m.nq), x);
    const double w = cq[i] * m.JxW[i % m.nq];
    Q += w;
    for (int k = 0; k < 3; ++k) dip[k] += w * x[k];
  }
  vector<double> g(m.N, 0.0);
  for (long n = 0; n < m.N; ++n) {
    if (!m.boundary(n)) continue;
    if (bc) { g[n] = bc[n]; continue; }
    double x[3];
    node_xyz(m, n, x);
    const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    g[n] = Q / r + (dip[0] * x[0] + dip[1] * x[1] + dip[2] * x[2]) / (r * r * r);
  }
  // K phi_I = 4 pi int N c - K g on the interior nodes
  vector<double> w4((size_t)nQ), f, Kg;
  for (long i = 0; i < nQ; ++i) w4[i] = 4.0 * kPi * cq[i] * m.JxW[i % m.nq];
  integrate_basis(m, w4, f);
  cell_apply(m, m.Ke, g, Kg, false);
  vector<double> b(m.N, 0.0), Dinv(m.N, 0.0);
  const vector<double> Kd = diag_of(m, m.Ke);
  for (long n = 0; n < m.N; ++n)
    if (!m.boundary(n)) { b[n] = f[n] - Kg[n]; Dinv[n] = 1.0 / Kd[n]; }
  vector<double> x;
  double rr = 0.0;
  const int it = pcg(m, [&](const vector<double>& u, vector<double>& v) { cell_apply(m, m.Ke, u, v, true); },
                     b, Dinv, tol, maxit, x, &rr);
  for (long n = 0; n < m.N; ++n) x[n] = m.boundary(n) ? g[n] : x[n];
  if (phi_nodal) std::copy(x.begin(), x.end(), phi_nodal);
  if (phi_q) nodal_to_quad(m, x, phi_q);
  if (st) {
    st->iterations = it; st->rel_residual = rr; st->charge = Q;
    for (int k = 0; k < 3; ++k) st->dipole[k] = dip[k];
  }
  return it >= maxit && rr > tol ? IDFT_ENOCONV : IDFT_OK;
}

}  // extern "C"

// ---- initial v_xc guess (setInitialPotL2Proj, InverseDFTEngine.cpp:1044-1431)
extern "C" {

void idft_guess_params_default(idft_guess_params* gp) {
  if (!gp) return;
  gp->lb94 = 1; gp->tau_vx = 1e-2; gp->fermi_amaldi = 1; gp->tau_fa = 1e-2;
  gp->factor_fa = 1.0; gp->factor_lda = 0.0; gp->spin_factor = 2.0; gp->n_electrons = 0.0;
}

int idft_xc_potentials(int64_t n, const double* r, const double* sigma, int32_t lb94, double* vx,
                       double* vc) {
  if (n < 0 || !r || !vx || !vc || (lb94 && !sigma)) return IDFT_EINVAL;
  for (int64_t i = 0; i < n; ++i) {
    vx[i] = lb94 ? lb94_vx(r[i], sigma[i]) : slater_vx(r[i]);
    vc[i] = pw92_vc(r[i]);
  }
  return IDFT_OK;
}

int idft_l2_project_parent(const idft_desc* d, const double* v_q, double* x_nodal) {
  if (!desc_ok(d) || !v_q || !x_nodal) return IDFT_EINVAL;
  Mesh m; m.init(*d);
  const long nQ = m.ncell * m.nq;
  const vector<double> Me = cell_mass(m);
  vector<double> w((size_t)nQ), f;
  for (long i = 0; i < nQ; ++i) w[i] = v_q[i] * m.JxW[i % m.nq];
  integrate_basis(m, w, f);
  const vector<double> Md = diag_of(m, Me);
  vector<double> Dinv(m.N, 0.0);
  for (long n = 0; n < m.N; ++n)
    if (!m.boundary(n)) Dinv[n] = 1.0 / Md[n]; else f[n] = 0.0;
  vector<double> xp;
  double rr = 0.0;
  pcg(m, [&](const vector<double>& u, vector<double>& y) { cell_apply(m, Me, u, y, true); }, f, Dinv,
      1e-14, 2000, xp, &rr);
  std::copy(xp.begin(), xp.end(), x_nodal);
  return IDFT_OK;
}

// Values of the parent field at the 4 child Gauss points per parent axis,
// times scale, integrated against the child Q1 basis, then the child
// consistent mass (assembled element by element, Gauss 2) solved by CG.
int idft_parent_to_child(const idft_desc* d, const double* x_nodal, double scale, double* out) {
  if (!desc_ok(d) || !x_nodal || !out) return IDFT_EINVAL;
  Mesh m; m.init(*d);
  const long NV1 = m.NV1, NV = m.NV;
  const double jc = std::pow(0.25 * m.h, 3);
  vector<double> rhs(NV, 0.0);
  for (long e = 0; e < m.ncell; ++e) {
    const long cx = e % m.nc, cy = (e / m.nc) % m.nc, cz = e / ((long)m.nc * m.nc);
    const long* nodes = &m.cellNodes[(size_t)e * m.ne];
    for (int z = 0; z < 4; ++z) for (int y = 0; y < 4; ++y) for (int x = 0; x < 4; ++x) {
      double val = 0.0;
      for (int c = 0; c < m.n1; ++c) for (int b = 0; b < m.n1; ++b) for (int a = 0; a < m.n1; ++a)
        val += m.Lc[x * m.n1 + a] * m.Lc[y * m.n1 + b] * m.Lc[z * m.n1 + c] * x_nodal[nodes[(c * m.n1 + b) * m.n1 + a]];
      val *= scale;
      for (int kc = 0; kc < 2; ++kc) for (int kb = 0; kb < 2; ++kb) for (int ka = 0; ka < 2; ++ka) {
        const long i = 2 * cx + m.ccell[x] + ka, j = 2 * cy + m.ccell[y] + kb, k = 2 * cz + m.ccell[z] + kc;
        rhs[(k * NV1 + j) * NV1 + i] += m.cw[x][ka] * m.cw[y][kb] * m.cw[z][kc] * val * jc;
      }
    }
  }
  // child element mass from the child basis at its own Gauss points
  const double g2 = 1.0 / std::sqrt(3.0), hc = 0.5 * m.h;
  double me1[2][2] = {{0, 0}, {0, 0}};
  for (int k = 0; k < 2; ++k) {
    const double t = k ? g2 : -g2, n0 = 0.5 * (1 - t), n1v = 0.5 * (1 + t);
    me1[0][0] += 0.5 * hc * n0 * n0; me1[0][1] += 0.5 * hc * n0 * n1v;
    me1[1][0] += 0.5 * hc * n1v * n0; me1[1][1] += 0.5 * hc * n1v * n1v;
  }
  auto childMass = [&](const vector<double>& u, vector<double>& y) {
    y.assign(NV, 0.0);
    for (long ck = 0; ck < NV1 - 1; ++ck) for (long cj = 0; cj < NV1 - 1; ++cj) for (long ci = 0; ci < NV1 - 1; ++ci)
      for (int kc = 0; kc < 2; ++kc) for (int kb = 0; kb < 2; ++kb) for (int ka = 0; ka < 2; ++ka) {
        double s = 0.0;
        for (int lc = 0; lc < 2; ++lc) for (int lb = 0; lb < 2; ++lb) for (int la = 0; la < 2; ++la)
          s += me1[ka][la] * me1[kb][lb] * me1[kc][lc] * u[((ck + lc) * NV1 + cj + lb) * NV1 + ci + la];
        y[((ck + kc) * NV1 + cj + kb) * NV1 + ci + ka] += s;
      }
  };
  vector<double> xc(NV, 0.0), r = rhs, p, q, z(NV), dg(NV);
  {
    vector<double> e1(NV, 0.0);
    // diagonal: product over axes of the 1D diagonal (end nodes have one element)
    for (long n = 0; n < NV; ++n) {
      const long i = n % NV1, j = (n / NV1) % NV1, k = n / (NV1 * NV1);
      auto f1 = [&](long t) { return (t == 0 || t == NV1 - 1) ? me1[0][0] : 2.0 * me1[0][0]; };
      dg[n] = f1(i) * f1(j) * f1(k);
    }
  }
  for (long n = 0; n < NV; ++n) z[n] = r[n] / dg[n];
  p = z;
  double rz = 0.0;
  for (long n = 0; n < NV; ++n) rz += r[n] * z[n];
  const double bn = std::sqrt(std::max(rz, 0.0));
  for (int it = 0; it < 2000 && std::sqrt(std::max(rz, 0.0)) > 1e-14 * bn && bn > 0; ++it) {
    childMass(p, q);
    double pq = 0.0;
    for (long n = 0; n < NV; ++n) pq += p[n] * q[n];
    const double al = rz / pq;
    for (long n = 0; n < NV; ++n) { xc[n] += al * p[n]; r[n] -= al * q[n]; z[n] = r[n] / dg[n]; }
    double rz2 = 0.0;
    for (long n = 0; n < NV; ++n) rz2 += r[n] * z[n];
    const double be = rz2 / rz;
    rz = rz2;
    for (long n = 0; n < NV; ++n) p[n] = z[n] + be * p[n];
  }
  std::copy(xc.begin(), xc.end(), out);
  return IDFT_OK;
}

int idft_initial_vxc(const idft_desc* d, const idft_guess_params* gp, const double* rho_spin_q,
                     const double* sigma_q, const double* vH_q, double* v_lda_q, double* vxc_child) {
  if (!desc_ok(d) || !gp || !rho_spin_q || (gp->lb94 && !sigma_q) ||
      (gp->fermi_amaldi && (!vH_q || !(gp->n_electrons > 0))))
    return IDFT_EINVAL;
  Mesh m; m.init(*d);
  const long nQ = m.ncell * m.nq;
  vector<double> v((size_t)nQ);
  for (long i = 0; i < nQ; ++i) {
    const double r = gp->spin_factor * rho_spin_q[i];
    const double vx = gp->lb94 ? lb94_vx(r, sigma_q[i]) : slater_vx(r);
    const double pre = r / (r + gp->tau_vx);
    double val = (1.0 - pre) * vx + pre * (vx + pw92_vc(r));
    if (gp->fermi_amaldi) {
      const double pb = r / (r + gp->tau_fa);
      val = pb * val + (1.0 - pb) * (-1.0 / gp->n_electrons) * gp->factor_fa * vH_q[i];
    }
    v[i] = val;
  }
  if (v_lda_q) std::copy(v.begin(), v.end(), v_lda_q);
  if (!vxc_child) return IDFT_OK;
  vector<double> xp(m.N);
  int rc = idft_l2_project_parent(d, v.data(), xp.data());
  if (rc) return rc;
  return idft_parent_to_child(d, xp.data(), 1.0 - gp->factor_lda, vxc_child);
}

}  // extern "C"
