// Drives both reference contracts through include/idft_shim.hpp, the way the
// reference's callers do (BFGSInverseDFTSolver -> InverseDFTSolverFunction;
// the multivector MINRES solver -> MultiVectorAdjointLinearSolverProblem),
// against whichever library it is linked with, and prints "key value" lines
// for tests/test_shim.py to compare between the HIP library and the oracle.
#include <cmath>
#include <cstdio>
#include <vector>

#include "idft_shim.hpp"

using idft_shim::InverseDFTSolverFunction;
using idft_shim::MultiVector;
using idft_shim::MultiVectorAdjointLinearSolverProblem;

static double sumsq(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return s;
}
static double wsum(const std::vector<double>& v) {
  double s = 0.0;
  for (size_t i = 0; i < v.size(); ++i) s += v[i] * std::sin(0.37 * (double)(i % 1013));
  return s;
}

int main() {
  idft_desc d;
  idft_desc_default(&d);
  d.n_cells = 3; d.degree = 4; d.half_width = 4.0; d.n_eig = 6; d.n_electrons = 4.0;
  d.cheb_tol = 1e-11; d.initial_tol_cheb = 1e-11;
  idft_sizes sz;
  idft_shim::check(idft_geometry_sizes(&d, &sz), nullptr, "sizes");
  const long nq = (long)(sz.n_cells * sz.quad_per_cell);
  std::vector<double> xyz(3 * nq), pot(nq), zero(nq, 0.0), one(nq, 1.0), nodal(sz.n_nodes, 1.0);
  idft_geometry_quad_points(&d, xyz.data());
  for (long q = 0; q < nq; ++q) {
    const double* x = &xyz[3 * q];
    const double r2a = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    const double r2b = x[0] * x[0] + x[1] * x[1] + (x[2] - 1.5) * (x[2] - 1.5);
    pot[q] = -2.0 / std::sqrt(r2a + 0.5) - 1.0 / std::sqrt(r2b + 0.5);
  }
  // target density: the ground state of the wells alone (spin density)
  idft_ctx* t = nullptr;
  idft_shim::check(idft_create(&d, pot.data(), zero.data(), one.data(), nodal.data(), &t), t, "create");
  idft_shim::check(idft_set_veff_quad(t, pot.data()), t, "veff");
  std::vector<double> ev(d.n_eig), rho(nq);
  double mr = 0.0;
  int32_t passes = 0;
  idft_shim::check(idft_eigensolve(t, 1e-11, 100, ev.data(), &mr, &passes), t, "eigensolve");
  idft_shim::check(idft_get_density(t, rho.data()), t, "density");
  for (double& v : rho) v *= 0.5;
  idft_destroy(t);

  idft_ctx* c = nullptr;
  idft_shim::check(idft_create(&d, pot.data(), rho.data(), one.data(), nodal.data(), &c), c, "create");
  // ---- contract 1, as BFGSInverseDFTSolver::solve calls it
  InverseDFTSolverFunction f(c);
  std::vector<std::vector<double>> v = f.getInitialGuess(), force;
  std::vector<double> vxy(3 * sz.n_vxc);
  idft_geometry_vxc_points(&d, vxy.data());
  for (long n = 0; n < (long)sz.n_vxc; ++n) {
    const double* x = &vxy[3 * n];
    v[0][n] = 0.02 * std::exp(-(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) / 2.0);
  }
  std::vector<double> loss(1);
  f.getForceVector(v, force, loss);
  std::vector<double> dots;
  f.dotProduct(force[0], force[0], 1, dots);
  std::printf("loss %.17e\nforce_sumsq %.17e\nforce_wsum %.17e\ndot %.17e\n", loss[0],
              sumsq(force[0]), wsum(force[0]), dots[0]);

  // ---- contract 2, as the multivector MINRES solver calls it
  idft_shim::check(idft_set_potential(c, v[0].data()), c, "potential");
  idft_shim::check(idft_eigensolve(c, 1e-11, 100, ev.data(), &mr, &passes), c, "eigensolve");
  std::vector<double> occ(d.n_eig), rq(nq), psi((size_t)sz.n_nodes * d.n_eig);
  double fermi = 0.0;
  idft_shim::check(idft_get_occupancies(c, occ.data(), &fermi), c, "occ");
  idft_shim::check(idft_get_density(c, rq.data()), c, "density");
  idft_shim::check(idft_get_eigenvectors(c, psi.data()), c, "eigenvectors");
  std::vector<double> u(nq);
  for (long q = 0; q < nq; ++q) u[q] = rho[q] - 0.5 * rq[q];
  const unsigned B = (unsigned)d.n_eig;
  std::vector<std::vector<unsigned>> deg(B);
  for (unsigned i = 0; i < B; ++i)
    for (unsigned j = 0; j < B; ++j)
      if (std::fabs(ev[i] - ev[j]) < d.degeneracy_tol) deg[i].push_back(j);
  MultiVector mpsi(c, (int)B), x(c, (int)B), ax(c, (int)B), pj(c, (int)B), pjs(c, (int)B),
      ndbc(c, (int)B), sol(c, (int)B);
  mpsi.set(psi);
  std::vector<double> eneg = ev;
  MultiVectorAdjointLinearSolverProblem prob(c);
  prob.updateInputPsi(mpsi, occ, u, deg, fermi, eneg, B);
  std::printf("negated_eigenvalue0 %.17e\neigenvalue0 %.17e\n", eneg[0], ev[0]);
  MultiVector& rhs = prob.computeRhs(ndbc, x, B);
  std::vector<double> xin((size_t)sz.n_nodes * B);
  for (size_t i = 0; i < xin.size(); ++i) xin[i] = std::cos(0.11 * (double)(i % 997));
  x.set(xin);
  prob.vmult(ax, x, B);
  prob.precondition_Jacobi(pj, x, 1.0);
  prob.precondition_JacobiSqrt(pjs, x, 1.0);
  std::printf("rhs_sumsq %.17e\nvmult_wsum %.17e\njacobi_wsum %.17e\njacobi_sqrt_wsum %.17e\n",
              sumsq(rhs.get()), wsum(ax.get()), wsum(pj.get()), wsum(pjs.get()));
  // MultiVectorMinResSolver::solve stand-in: computeRhs, MINRES, distributeX
  std::vector<int32_t> its(B);
  idft_shim::check(idft_minres_solve(c, sol.handle(), ndbc.handle(), 1e-13, 3000, 1, its.data()), c,
                   "minres");
  std::printf("solution_sumsq %.17e\nsolution_wsum %.17e\n", sumsq(sol.get()), wsum(sol.get()));
  idft_destroy(c);
  return 0;
}
