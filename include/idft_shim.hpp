// idft_shim.hpp -- the reference's two C++ interfaces over the C ABI of
// idft.h, header-only.  A caller written against
//   invDFT::InverseDFTSolverFunction              (InverseDFTSolverFunction.h:51-183)
//   invDFT::MultiVectorAdjointLinearSolverProblem (MultiVectorAdjointLinearSolverProblem.h:29-213)
// keeps its call sequence and argument meaning and links libidft.so (HIP) or
// the CPU oracle instead.  deal.II / DFT-FE types are replaced by plain
// containers: a distributed vector is a std::vector<double> over the child
// v_xc nodes, a MultiVector is idft_shim::MultiVector (device memory of the
// library, node-major [node][col]).  Errors the reference reports by
// AssertThrow come back as idft_shim::Error (what() = idft_last_error).
#ifndef IDFT_SHIM_HPP
#define IDFT_SHIM_HPP

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "idft.h"

namespace idft_shim {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const idft_ctx* c, const char* what) {
  if (rc != IDFT_OK)
    throw Error(rc, std::string(what) + ": " + (c ? idft_last_error(c) : "invalid argument"));
}

// dftfe::linearAlgebra::MultiVector stand-in: a context-owned block of
// `cols` vectors over the parent nodes (idft_mv).
class MultiVector {
 public:
  MultiVector(idft_ctx* ctx, int cols) : ctx_(ctx), cols_(cols) {
    check(idft_mv_create(ctx, cols, &mv_), ctx, "idft_mv_create");
  }
  MultiVector(const MultiVector&) = delete;
  MultiVector& operator=(const MultiVector&) = delete;
  MultiVector(MultiVector&& o) noexcept : ctx_(o.ctx_), mv_(o.mv_), cols_(o.cols_) { o.mv_ = nullptr; }
  ~MultiVector() {
    if (mv_) idft_mv_destroy(mv_);
  }
  int numVectors() const { return cols_; }
  idft_mv* handle() const { return mv_; }
  void set(const std::vector<double>& host) {
    check(idft_mv_set(ctx_, mv_, host.data()), ctx_, "idft_mv_set");
  }
  std::vector<double> get() const {
    idft_sizes s;
    check(idft_get_sizes(ctx_, &s), ctx_, "idft_get_sizes");
    std::vector<double> out((size_t)s.n_nodes * cols_);
    check(idft_mv_get(ctx_, mv_, out.data()), ctx_, "idft_mv_get");
    return out;
  }

 private:
  idft_ctx* ctx_;
  idft_mv* mv_ = nullptr;
  int cols_;
};

// InverseDFTSolverFunction (Contract 1): what BFGSInverseDFTSolver calls.
// One spin channel (n_spin = 1): the vectors of vectors have one entry.
class InverseDFTSolverFunction {
 public:
  explicit InverseDFTSolverFunction(idft_ctx* ctx) : ctx_(ctx) {
    check(idft_get_sizes(ctx, &sz_), ctx, "idft_get_sizes");
  }
  idft_ctx* context() const { return ctx_; }
  // getInitialGuess (InverseDFTSolverFunction.cpp:1676-1683)
  std::vector<std::vector<double>> getInitialGuess() const {
    std::vector<std::vector<double>> v(1, std::vector<double>((size_t)sz_.n_vxc));
    check(idft_get_initial_guess(ctx_, v[0].data()), ctx_, "getInitialGuess");
    return v;
  }
  void setInitialGuess(const std::vector<std::vector<double>>& pot) {
    check(idft_set_initial_guess(ctx_, pot.at(0).data()), ctx_, "setInitialGuess");
  }
  // getForceVector (.cpp:657-1311): pot is borrowed and mutated (constrained
  // entries zeroed); force is resized here; loss must hold n_spin entries.
  void getForceVector(std::vector<std::vector<double>>& pot,
                      std::vector<std::vector<double>>& force, std::vector<double>& loss) {
    if (pot.size() != 1 || loss.size() != 1) throw Error(IDFT_EINVAL, "one spin channel");
    force.assign(1, std::vector<double>((size_t)sz_.n_vxc));
    check(idft_get_force(ctx_, pot[0].data(), force[0].data(), loss.data()), ctx_,
          "getForceVector");
  }
  // dotProduct (.cpp:1691-1705): blockSize interleaved components.
  void dotProduct(const std::vector<double>& v1, const std::vector<double>& v2,
                  unsigned int blockSize, std::vector<double>& out) {
    out.assign(blockSize, 0.0);
    check(idft_dot(ctx_, v1.data(), v2.data(), (int64_t)v1.size(), (int32_t)blockSize, out.data()),
          ctx_, "dotProduct");
  }
  // setSolution (.cpp:1685-1689)
  void setSolution(const std::vector<std::vector<double>>& pot) {
    check(idft_set_solution(ctx_, pot.at(0).data()), ctx_, "setSolution");
  }

 private:
  idft_ctx* ctx_;
  idft_sizes sz_;
};

// MultiVectorAdjointLinearSolverProblem (Contract 2): what the multivector
// MINRES solver calls, plus updateInputPsi from getForceVector.
class MultiVectorAdjointLinearSolverProblem {
 public:
  explicit MultiVectorAdjointLinearSolverProblem(idft_ctx* ctx) : ctx_(ctx) {}
  // updateInputPsi (.cpp:817-910): psi borrowed; eigenValues negated in place
  // as the reference does (.cpp:898-903).
  void updateInputPsi(MultiVector& psi, const std::vector<double>& effectiveOrbitalOccupancy,
                      const std::vector<double>& uQuad,
                      const std::vector<std::vector<unsigned int>>& degeneracy,
                      double fermiEnergy, std::vector<double>& eigenValues,
                      unsigned int blockSize) {
    std::vector<int32_t> pairs;
    for (unsigned int i = 0; i < degeneracy.size(); ++i)
      for (unsigned int j : degeneracy[i]) {
        pairs.push_back((int32_t)i);
        pairs.push_back((int32_t)j);
      }
    check(idft_adj_update_input_psi(ctx_, psi.handle(), effectiveOrbitalOccupancy.data(),
                                    uQuad.data(), pairs.data(), (int32_t)(pairs.size() / 2),
                                    fermiEnergy, eigenValues.data(), (int32_t)blockSize),
          ctx_, "updateInputPsi");
    rhs_.reset(new MultiVector(ctx_, (int)blockSize));
  }
  // computeRhs (.cpp:614-812): the problem-owned right-hand side; x is kept
  // for distributeX.
  MultiVector& computeRhs(MultiVector& NDBCVec, MultiVector& outputVec, unsigned int blockSize) {
    (void)blockSize;
    check(idft_adj_compute_rhs(ctx_, NDBCVec.handle(), outputVec.handle(), rhs_->handle()), ctx_,
          "computeRhs");
    return *rhs_;
  }
  // vmult (.cpp:1345-1367)
  void vmult(MultiVector& Ax, MultiVector& x, unsigned int blockSize) {
    check(idft_adj_vmult(ctx_, Ax.handle(), x.handle(), (int32_t)blockSize), ctx_, "vmult");
  }
  // precondition_Jacobi / precondition_JacobiSqrt (.cpp:587-612); omega unused
  void precondition_Jacobi(MultiVector& dst, const MultiVector& src, double omega) const {
    check(idft_adj_precondition_jacobi(ctx_, dst.handle(), src.handle(), omega), ctx_,
          "precondition_Jacobi");
  }
  void precondition_JacobiSqrt(MultiVector& dst, const MultiVector& src, double omega) const {
    check(idft_adj_precondition_jacobi_sqrt(ctx_, dst.handle(), src.handle(), omega), ctx_,
          "precondition_JacobiSqrt");
  }
  // distributeX (.cpp:1002-1141)
  void distributeX() { check(idft_adj_distribute_x(ctx_), ctx_, "distributeX"); }

 private:
  idft_ctx* ctx_;
  std::unique_ptr<MultiVector> rhs_;
};

}  // namespace idft_shim

#endif
