// Internal declarations of the set-up stage (include/idft_setup.h) of the
// HIP library: device views of the atomic-orbital basis and the launchers of
// setup.hip.
#pragma once

#include <vector>

#include "common.h"

namespace idft {

// Parent-quadrature geometry for kernels that build points on the fly.
struct QuadGeo {
  int nc, q1;
  double L, h;
  double gx[16];  // Gauss abscissae on [-1, 1] (q1 <= 16)
};

// Device view of an atomic-orbital basis: function i has primitives
// [prim_start[i], prim_start[i+1]) (exponent alpha, coefficient x norm cn)
// and solid-harmonic monomials [mono_start[i], mono_start[i+1])
// (powers mono_pow[3t..3t+2] of x, y, z about center[3i..], coefficient
// mono_coef[t]), and the normalisation norm[i].
struct AoDev {
  int nb;
  const double* center;
  const int* prim_start;
  const double* alpha;
  const double* cn;
  const int* mono_start;
  const int* mono_pow;
  const double* mono_coef;
  const double* norm;
};

struct GuessArgs {
  int lb94, fermi_amaldi;
  double tau_vx, tau_fa, factor_fa, spin_factor, n_electrons;
};

void launch_ao_values(const QuadGeo& qg, const AoDev& ao, int64_t row0, int64_t rows, double* phi,
                      double* gphi, hipStream_t s);
// the same at np caller points pts [np][3] (no gradients)
void launch_ao_values_at(const AoDev& ao, int64_t np, const double* pts, double* phi,
                         hipStream_t s);
// returns the number of partial rows (each nb wide)
int launch_ao_selfoverlap(const QuadGeo& qg, int nb, int64_t row0, int64_t rows, const double* phi,
                          const double* jxw, double* partial, int max_blocks, hipStream_t s);
void launch_ao_rows(int nb, int64_t rows, const double* T, const double* phi, const double* gphi,
                    double* rho, double* grad, hipStream_t s);
void launch_loss_weights(int64_t n, const double* rho, double sf, double a1, double a2, double tau,
                         double* w, hipStream_t s);
void launch_xc(int64_t n, const double* r, const double* sigma, int lb94, double* vx, double* vc,
               hipStream_t s);
void launch_guess(int64_t n, const GuessArgs& ga, const double* rho_spin, const double* sigma,
                  const double* vH, double* v, hipStream_t s);
void launch_nodal_to_quad(const Geo& g, const double* Lq, const double* x, double* out,
                          hipStream_t s);
void child_mass_factors(int NV1, double hc, std::vector<double>& cp, std::vector<double>& inv);
void launch_child_mass_solve(int NV1, double hc, const double* cp, const double* inv, double* v,
                             hipStream_t s);

}  // namespace idft
