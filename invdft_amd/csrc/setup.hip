// Device kernels of the set-up stage (include/idft_setup.h): atomic-orbital
// basis values and the target density, the exchange-correlation guess, the
// loss weights, nodal -> quadrature interpolation and the tridiagonal solves
// of the child-mesh L2 projection.  All are once-per-run, HBM-bound kernels
// except the density's n_basis^2 contraction, which is a library GEMM
// (rocBLAS) between the basis-value kernel and the row-reduction kernel.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.h"
#include "setup.h"

namespace idft {

namespace {

constexpr int kT = 256;

// Parent quadrature point of flat index i = e * Q^3 + q.
__device__ inline void quad_point(const QuadGeo& qg, int64_t i, double* x) {
  const int nq = qg.q1 * qg.q1 * qg.q1;
  const int64_t e = i / nq;
  const int q = (int)(i - e * nq);
  const int64_t cx = e % qg.nc, cy = (e / qg.nc) % qg.nc, cz = e / ((int64_t)qg.nc * qg.nc);
  const int qx = q % qg.q1, qy = (q / qg.q1) % qg.q1, qz = q / (qg.q1 * qg.q1);
  x[0] = -qg.L + cx * qg.h + 0.5 * qg.h * (qg.gx[qx] + 1.0);
  x[1] = -qg.L + cy * qg.h + 0.5 * qg.h * (qg.gx[qy] + 1.0);
  x[2] = -qg.L + cz * qg.h + 0.5 * qg.h * (qg.gx[qz] + 1.0);
}

// Basis values (and gradients) at `rows` quadrature points from row0:
// phi[r][i], gphi[r][i][3].  One thread per (point, function); a function is
// norm * Y(d) * R(r^2) with Y its real solid harmonic as a polynomial in
// d = x - centre (monomial list) and R = sum_k cn_k exp(-a_k r^2).
// With pts, the points are pts[3 (row0 + r) ..] instead of the quadrature.
__global__ void k_ao_values(QuadGeo qg, AoDev ao, int64_t row0, int64_t rows,
                            const double* __restrict__ pts, double* __restrict__ phi,
                            double* __restrict__ gphi) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = ao.nb;
  if (idx >= rows * nb) return;
  const int64_t r = idx / nb;
  const int i = (int)(idx - r * nb);
  double x[3];
  if (pts) {
    x[0] = pts[3 * (row0 + r)]; x[1] = pts[3 * (row0 + r) + 1]; x[2] = pts[3 * (row0 + r) + 2];
  } else {
    quad_point(qg, row0 + r, x);
  }
  const double d0 = x[0] - ao.center[3 * i], d1 = x[1] - ao.center[3 * i + 1],
               d2 = x[2] - ao.center[3 * i + 2];
  const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
  double R = 0.0, dR = 0.0;  // dR: d R / d(r^2) * 2 = sum -2 a cn e^{-a r^2}
  for (int k = ao.prim_start[i]; k < ao.prim_start[i + 1]; ++k) {
    const double ex = ao.cn[k] * exp(-ao.alpha[k] * r2);
    R += ex;
    dR -= 2.0 * ao.alpha[k] * ex;
  }
  // powers d^0..d^6 per axis
  double px[7], py[7], pz[7];
  px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
  for (int t = 1; t < 7; ++t) { px[t] = px[t - 1] * d0; py[t] = py[t - 1] * d1; pz[t] = pz[t - 1] * d2; }
  double Y = 0.0, Yx = 0.0, Yy = 0.0, Yz = 0.0;
  for (int t = ao.mono_start[i]; t < ao.mono_start[i + 1]; ++t) {
    const int a = ao.mono_pow[3 * t], b = ao.mono_pow[3 * t + 1], c = ao.mono_pow[3 * t + 2];
    const double cf = ao.mono_coef[t];
    Y += cf * px[a] * py[b] * pz[c];
    if (gphi) {
      if (a) Yx += cf * a * px[a - 1] * py[b] * pz[c];
      if (b) Yy += cf * b * px[a] * py[b - 1] * pz[c];
      if (c) Yz += cf * c * px[a] * py[b] * pz[c - 1];
    }
  }
  const double nrm = ao.norm[i];
  phi[r * nb + i] = nrm * Y * R;
  if (gphi) {
    double* gp = gphi + (r * nb + i) * 3;
    gp[0] = nrm * (Yx * R + Y * dR * d0);
    gp[1] = nrm * (Yy * R + Y * dR * d1);
    gp[2] = nrm * (Yz * R + Y * dR * d2);
  }
}

// partial[blk][i] = sum over this block's rows of phi[r][i]^2 JxW(row0 + r)
__global__ void k_ao_selfoverlap(QuadGeo qg, int nb, int64_t row0, int64_t rows,
                                 const double* __restrict__ phi, const double* __restrict__ jxw,
                                 int rows_per_block, double* __restrict__ partial) {
  const int nq = qg.q1 * qg.q1 * qg.q1;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = min(rows, r0 + rows_per_block);
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    double s = 0.0;
    for (int64_t r = r0; r < r1; ++r) {
      const double v = phi[r * nb + i];
      s += v * v * jxw[(row0 + r) % nq];
    }
    partial[(int64_t)blockIdx.x * nb + i] = s;
  }
}

// rho[r] = sum_i T[r][i] phi[r][i]; grad[r][k] = 2 sum_i T[r][i] gphi[r][i][k].
// One wave per row, fixed-order shuffle reduction.
__global__ void k_ao_rows(int nb, int64_t rows, const double* __restrict__ T,
                          const double* __restrict__ phi, const double* __restrict__ gphi,
                          double* __restrict__ rho, double* __restrict__ grad) {
  const int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (r >= rows) return;
  double s = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
  for (int i = lane; i < nb; i += 64) {
    const double t = T[r * nb + i];
    s += t * phi[r * nb + i];
    if (gphi) {
      const double* gp = gphi + (r * nb + i) * 3;
      g0 += t * gp[0]; g1 += t * gp[1]; g2 += t * gp[2];
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off, 64);
    if (gphi) {
      g0 += __shfl_down(g0, off, 64);
      g1 += __shfl_down(g1, off, 64);
      g2 += __shfl_down(g2, off, 64);
    }
  }
  if (lane == 0) {
    rho[r] = s;
    if (gphi) { grad[3 * r] = 2.0 * g0; grad[3 * r + 1] = 2.0 * g1; grad[3 * r + 2] = 2.0 * g2; }
  }
}

__global__ void k_loss_weights(int64_t n, const double* __restrict__ rho, double sf, double a1,
                               double a2, double tau, double* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = sf * rho[i];
  w[i] = 1.0 / (pow(r, a1) + tau) + pow(r, a2);
}

// Exchange-correlation potentials of the total density n (published forms):
// Slater exchange, van Leeuwen-Baerends 1994 exchange (beta = 0.05, per
// spin channel n/2), Perdew-Wang 1992 correlation (unpolarised).
__device__ inline double dev_slater(double n) { return n > 1e-30 ? -cbrt(3.0 * n / M_PI) : 0.0; }
__device__ inline double dev_lb94(double n, double sigma) {
  if (n <= 1e-30) return 0.0;
  const double beta = 0.05, ns = 0.5 * n;
  const double x = 0.5 * sqrt(fmax(sigma, 0.0)) / pow(ns, 4.0 / 3.0);
  return dev_slater(n) - beta * cbrt(ns) * x * x / (1.0 + 3.0 * beta * x * asinh(x));
}
__device__ inline double dev_pw92(double n) {
  if (n <= 1e-30) return 0.0;
  const double A = 0.031091, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
  const double rs = cbrt(3.0 / (4.0 * M_PI * n)), srs = sqrt(rs);
  const double Q0 = -2.0 * A * (1.0 + a1 * rs);
  const double Q1 = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  const double Q1p = A * (b1 / srs + 2.0 * b2 + 3.0 * b3 * srs + 4.0 * b4 * rs);
  const double lg = log(1.0 + 1.0 / Q1);
  const double eps = Q0 * lg;
  const double deps = -2.0 * A * a1 * lg - Q0 * Q1p / (Q1 * Q1 + Q1);
  return eps - rs / 3.0 * deps;
}

__global__ void k_xc(int64_t n, const double* __restrict__ r, const double* __restrict__ sigma,
                     int lb94, double* __restrict__ vx, double* __restrict__ vc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vx[i] = lb94 ? dev_lb94(r[i], sigma[i]) : dev_slater(r[i]);
  vc[i] = dev_pw92(r[i]);
}

// The guess at a quadrature point (setInitialPotL2Proj, InverseDFTEngine.cpp:1196-1268).
__global__ void k_guess(int64_t n, GuessArgs ga, const double* __restrict__ rho_spin,
                        const double* __restrict__ sigma, const double* __restrict__ vH,
                        double* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = ga.spin_factor * rho_spin[i];
  const double vx = ga.lb94 ? dev_lb94(r, sigma[i]) : dev_slater(r);
  const double pre = r / (r + ga.tau_vx);
  double val = (1.0 - pre) * vx + pre * (vx + dev_pw92(r));
  if (ga.fermi_amaldi) {
    const double pb = r / (r + ga.tau_fa);
    val = pb * val + (1.0 - pb) * (-1.0 / ga.n_electrons) * ga.factor_fa * vH[i];
  }
  v[i] = val;
}

// out[e][q] = sum_abc Lq[qx][a] Lq[qy][b] Lq[qz][c] x[node(e, a, b, c)]
__global__ void k_nodal_to_quad(Geo g, const double* __restrict__ Lq, const double* __restrict__ x,
                                double* __restrict__ out) {
  const int nq = g.q1 * g.q1 * g.q1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.ncell * nq) return;
  const int64_t e = i / nq;
  const int q = (int)(i - e * nq);
  const int64_t cx = e % g.nc, cy = (e / g.nc) % g.nc, cz = e / ((int64_t)g.nc * g.nc);
  const int qx = q % g.q1, qy = (q / g.q1) % g.q1, qz = q / (g.q1 * g.q1);
  double s = 0.0;
  for (int c = 0; c < g.n1; ++c) {
    const double lc = Lq[qz * g.n1 + c];
    for (int b = 0; b < g.n1; ++b) {
      const double lbc = lc * Lq[qy * g.n1 + b];
      const int64_t row = ((cz * g.p + c) * g.N1 + (cy * g.p + b)) * (int64_t)g.N1 + cx * g.p;
      for (int a = 0; a < g.n1; ++a) s += lbc * Lq[qx * g.n1 + a] * x[row + a];
    }
  }
  out[i] = s;
}

// In-place solve along one axis of the child grid (NV1^3 nodes) with the 1D
// Q1 mass matrix of element size hc, (hc/6) tridiag(1, 4, 1) with hc/3 at
// both ends, by the Thomas algorithm, one thread per grid line.  The matrix
// is the same for every line, so its elimination factors come from the host:
// cp[i] (modified super-diagonal) and inv[i] (1 / modified diagonal); sub is
// the constant sub-diagonal hc/6.
__global__ void k_tridiag_axis(int NV1, int axis, double sub, const double* __restrict__ cp,
                               const double* __restrict__ inv, double* __restrict__ v) {
  const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= (int64_t)NV1 * NV1) return;
  const int64_t u = line % NV1, w = line / NV1;
  int64_t base, stride;
  if (axis == 0) { base = (w * NV1 + u) * NV1; stride = 1; }
  else if (axis == 1) { base = w * NV1 * NV1 + u; stride = NV1; }
  else { base = w * NV1 + u; stride = (int64_t)NV1 * NV1; }
  double dprev = 0.0;
  for (int i = 0; i < NV1; ++i) {
    const double dd = (v[base + i * stride] - (i ? sub * dprev : 0.0)) * inv[i];
    v[base + i * stride] = dd;
    dprev = dd;
  }
  for (int i = NV1 - 2; i >= 0; --i) v[base + i * stride] -= cp[i] * v[base + (i + 1) * stride];
}

}  // namespace

static unsigned grid_of(int64_t n, int t = kT) { return (unsigned)((n + t - 1) / t); }

void launch_ao_values(const QuadGeo& qg, const AoDev& ao, int64_t row0, int64_t rows, double* phi,
                      double* gphi, hipStream_t s) {
  if (rows <= 0) return;
  k_ao_values<<<grid_of(rows * ao.nb), kT, 0, s>>>(qg, ao, row0, rows, nullptr, phi, gphi);
  IDFT_HIP(hipGetLastError());
}
void launch_ao_values_at(const AoDev& ao, int64_t np, const double* pts, double* phi,
                         hipStream_t s) {
  if (np <= 0) return;
  QuadGeo qg{};
  k_ao_values<<<grid_of(np * ao.nb), kT, 0, s>>>(qg, ao, 0, np, pts, phi, nullptr);
  IDFT_HIP(hipGetLastError());
}
int launch_ao_selfoverlap(const QuadGeo& qg, int nb, int64_t row0, int64_t rows, const double* phi,
                          const double* jxw, double* partial, int max_blocks, hipStream_t s) {
  const int rpb = (int)std::max<int64_t>(1, (rows + max_blocks - 1) / max_blocks);
  const int nblk = (int)((rows + rpb - 1) / rpb);
  k_ao_selfoverlap<<<nblk, kT, 0, s>>>(qg, nb, row0, rows, phi, jxw, rpb, partial);
  IDFT_HIP(hipGetLastError());
  return nblk;
}
void launch_ao_rows(int nb, int64_t rows, const double* T, const double* phi, const double* gphi,
                    double* rho, double* grad, hipStream_t s) {
  if (rows <= 0) return;
  k_ao_rows<<<grid_of(rows, 4), kT, 0, s>>>(nb, rows, T, phi, gphi, rho, grad);
  IDFT_HIP(hipGetLastError());
}
void launch_loss_weights(int64_t n, const double* rho, double sf, double a1, double a2, double tau,
                         double* w, hipStream_t s) {
  if (n <= 0) return;
  k_loss_weights<<<grid_of(n), kT, 0, s>>>(n, rho, sf, a1, a2, tau, w);
  IDFT_HIP(hipGetLastError());
}
void launch_xc(int64_t n, const double* r, const double* sigma, int lb94, double* vx, double* vc,
               hipStream_t s) {
  if (n <= 0) return;
  k_xc<<<grid_of(n), kT, 0, s>>>(n, r, sigma, lb94, vx, vc);
  IDFT_HIP(hipGetLastError());
}
void launch_guess(int64_t n, const GuessArgs& ga, const double* rho_spin, const double* sigma,
                  const double* vH, double* v, hipStream_t s) {
  if (n <= 0) return;
  k_guess<<<grid_of(n), kT, 0, s>>>(n, ga, rho_spin, sigma, vH, v);
  IDFT_HIP(hipGetLastError());
}
void launch_nodal_to_quad(const Geo& g, const double* Lq, const double* x, double* out,
                          hipStream_t s) {
  const int64_t n = g.ncell * (int64_t)g.q1 * g.q1 * g.q1;
  k_nodal_to_quad<<<grid_of(n), kT, 0, s>>>(g, Lq, x, out);
  IDFT_HIP(hipGetLastError());
}
void launch_child_mass_solve(int NV1, double hc, const double* cp, const double* inv, double* v,
                             hipStream_t s) {
  const int64_t lines = (int64_t)NV1 * NV1;
  for (int axis = 0; axis < 3; ++axis) {
    k_tridiag_axis<<<grid_of(lines, 64), 64, 0, s>>>(NV1, axis, hc / 6.0, cp, inv, v);
    IDFT_HIP(hipGetLastError());
  }
}

// Elimination factors of the 1D child mass matrix (see k_tridiag_axis).
void child_mass_factors(int NV1, double hc, std::vector<double>& cp, std::vector<double>& inv) {
  cp.assign(NV1, 0.0);
  inv.assign(NV1, 0.0);
  const double off = hc / 6.0;
  double cprev = 0.0;
  for (int i = 0; i < NV1; ++i) {
    const double diag = (i == 0 || i == NV1 - 1) ? hc / 3.0 : 2.0 * hc / 3.0;
    const double den = diag - (i ? off * cprev : 0.0);
    inv[i] = 1.0 / den;
    cp[i] = (i == NV1 - 1 ? 0.0 : off) / den;
    cprev = cp[i];
  }
}

}  // namespace idft
