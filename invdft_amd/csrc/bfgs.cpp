// L-BFGS driver over contract 1 (getForceVector / dotProduct), with the
// semantics of BFGSInverseDFTSolver (BFGSInverseDFTSolver.cpp:66-627): two-loop
// inverse-Hessian product, CP / secant-loss / secant-force-norm line searches,
// curvature test y.s > 0, history cap.  Vectors are host arrays of n_vxc
// doubles (the reference keeps them as host deal.II vectors too).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "common.h"

namespace {

using Vec = std::vector<double>;

const char kMagic[8] = {'I', 'D', 'F', 'T', 'L', 'B', 'F', '1'};

// A failing dotProduct inside the iteration, carried to idft_bfgs_solve's status.
struct DotFailure {
  int rc;
};

class Lbfgs {
 public:
  Lbfgs(idft_ctx* c, int64_t n, int history) : c_(c), n_(n), history_(history) {}

  int run(int lsType, int maxIter, double tol, double lsTol, int nLs, double* lossHist,
          double* gnormHist, double* timeHist, int32_t* nDone) {
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&] {
      return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    const idft::BfgsCheckpoint& ck = idft::bfgs_checkpoint_of(c_);
    x_.assign(n_, 0.0);
    g_.assign(n_, 0.0);
    p_.assign(n_, 0.0);
    Vec y(n_, 0.0), s(n_, 0.0);
    double L = 0.0;
    bool resumed = false;
    int rc = IDFT_OK;
    if (ck.resume && !ck.path.empty()) {
      rc = load(ck.path, L);
      if (rc == IDFT_OK) resumed = true;
      else if (rc != kNoFile) return rc;
    }
    if (!resumed && (rc = idft_get_initial_guess(c_, x_.data()))) return rc;
    int iter = 0;
    for (; iter < maxIter; ++iter) {
      if (iter == 0 && !resumed && (rc = evaluate(x_, g_, L))) return rc;
      const double gn = std::sqrt(dot(g_, g_));
      if (timeHist) timeHist[iter] = elapsed();
      if (lossHist) lossHist[iter] = L;
      if (gnormHist) gnormHist[iter] = gn;
      const bool converged = gn < tol;
      twoLoop(g_, p_);
      for (double& v : p_) v = -v;
      double f0 = lsType == IDFT_LS_SECANT_LOSS ? L
                  : lsType == IDFT_LS_SECANT_FORCE_NORM ? gn
                                                        : dot(g_, p_);
      std::vector<double> lam = {0.0, 1.0}, fs = {f0};
      if (converged) break;
      rc = lsType == IDFT_LS_CP ? lineSearchCP(lam, fs, nLs, lsTol)
                                : lineSearchSecant(lsType, lam, fs, nLs);
      if (rc) return rc;
      const double step = lam.back() > 0.0 ? lam.back() : 1.0;
      for (int64_t t = 0; t < n_; ++t) {
        s[t] = step * p_[t];
        x_[t] += s[t];
        y[t] = -g_[t];
      }
      if ((rc = evaluate(x_, g_, L))) return rc;
      for (int64_t t = 0; t < n_; ++t) y[t] += g_[t];
      const double curvature = dot(y, s);
      if (curvature > 0.0) {
        Y_.push_back(y);
        S_.push_back(s);
        rho_.push_back(1.0 / curvature);
      }
      if ((int)Y_.size() > history_) {
        Y_.pop_front();
        S_.pop_front();
        rho_.pop_front();
      }
      if (!ck.path.empty() && (iter + 1) % ck.every == 0 && (rc = save(ck.path, L))) return rc;
    }
    if (iter == maxIter) {
      if (lossHist) lossHist[iter] = L;
      if (gnormHist) gnormHist[iter] = std::sqrt(dot(g_, g_));
    }
    if (timeHist) timeHist[iter] = elapsed();
    idft_set_solution(c_, x_.data());
    if (nDone) *nDone = iter;
    return IDFT_OK;
  }

 private:
  static constexpr int kNoFile = 1;

  // Checkpoint: "IDFTLBF1", int64 n, int32 m, double L, x[n], g[n], then m
  // records (s[n], y[n], rho), oldest first.  Written to path + ".tmp" and
  // renamed, so a reader never sees a partial file.
  int save(const std::string& path, double L) const {
    const std::string tmp = path + ".tmp";
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return IDFT_EINVAL;
    const int64_t n = n_;
    const int32_t m = (int32_t)Y_.size();
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(&n, sizeof n, 1, f) == 1 &&
              std::fwrite(&m, sizeof m, 1, f) == 1 && std::fwrite(&L, sizeof L, 1, f) == 1 &&
              std::fwrite(x_.data(), sizeof(double), n_, f) == (size_t)n_ &&
              std::fwrite(g_.data(), sizeof(double), n_, f) == (size_t)n_;
    for (int32_t k = 0; ok && k < m; ++k)
      ok = std::fwrite(S_[k].data(), sizeof(double), n_, f) == (size_t)n_ &&
           std::fwrite(Y_[k].data(), sizeof(double), n_, f) == (size_t)n_ &&
           std::fwrite(&rho_[k], sizeof(double), 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) return IDFT_EINVAL;
    return IDFT_OK;
  }

  int load(const std::string& path, double& L) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return kNoFile;
    char magic[8];
    int64_t n = 0;
    int32_t m = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
              std::fread(&n, sizeof n, 1, f) == 1 && n == n_ &&
              std::fread(&m, sizeof m, 1, f) == 1 && m >= 0 && m <= history_ + 1 &&
              std::fread(&L, sizeof L, 1, f) == 1 &&
              std::fread(x_.data(), sizeof(double), n_, f) == (size_t)n_ &&
              std::fread(g_.data(), sizeof(double), n_, f) == (size_t)n_;
    Y_.clear();
    S_.clear();
    rho_.clear();
    for (int32_t k = 0; ok && k < m; ++k) {
      Vec sv(n_), yv(n_);
      double r = 0.0;
      ok = std::fread(sv.data(), sizeof(double), n_, f) == (size_t)n_ &&
           std::fread(yv.data(), sizeof(double), n_, f) == (size_t)n_ &&
           std::fread(&r, sizeof(double), 1, f) == 1;
      if (ok) {
        S_.push_back(std::move(sv));
        Y_.push_back(std::move(yv));
        rho_.push_back(r);
      }
    }
    std::fclose(f);
    while ((int)Y_.size() > history_) {
      Y_.pop_front();
      S_.pop_front();
      rho_.pop_front();
    }
    return ok ? IDFT_OK : IDFT_EINVAL;
  }

  double dot(const Vec& a, const Vec& b) {
    double o = 0.0;
    const int rc = idft_dot(c_, a.data(), b.data(), n_, 1, &o);
    if (rc != IDFT_OK) throw DotFailure{rc};
    return o;
  }

  int evaluate(const Vec& x, Vec& g, double& L) {
    Vec xx = x;  // getForceVector may modify the potential (constraints)
    g.assign(n_, 0.0);
    return idft_get_force(c_, xx.data(), g.data(), &L);
  }

  // z = H_k g (two-loop recursion, BFGSInverseDFTSolver.cpp:66-118)
  void twoLoop(const Vec& g, Vec& z) {
    Vec q = g;
    const int m = (int)Y_.size();
    double gamma = 1.0;
    if (m > 0) gamma = dot(S_.back(), Y_.back()) / dot(Y_.back(), Y_.back());
    Vec alpha(m, 0.0);
    for (int i = m - 1; i >= 0; --i) {
      alpha[i] = rho_[i] * dot(S_[i], q);
      for (int64_t t = 0; t < n_; ++t) q[t] -= alpha[i] * Y_[i][t];
    }
    z = q;
    for (double& v : z) v *= gamma;
    for (int i = 0; i < m; ++i) {
      const double beta = rho_[i] * dot(Y_[i], z);
      for (int64_t t = 0; t < n_; ++t) z[t] += (alpha[i] - beta) * S_[i][t];
    }
  }

  int lineValue(int type, double lam, double& f) {
    Vec xn(n_), gn;
    for (int64_t t = 0; t < n_; ++t) xn[t] = x_[t] + lam * p_[t];
    double L = 0.0;
    int rc = evaluate(xn, gn, L);
    if (rc) return rc;
    if (type == IDFT_LS_SECANT_LOSS) f = L;
    else if (type == IDFT_LS_SECANT_FORCE_NORM) f = std::sqrt(dot(gn, gn));
    else f = dot(gn, p_);
    return IDFT_OK;
  }

  // solveLineSearchCP (BFGSInverseDFTSolver.cpp:215-303)
  int lineSearchCP(std::vector<double>& lam, std::vector<double>& f, int maxIter, double tol) {
    for (int i = 0; i < maxIter; ++i) {
      const double l0 = lam[lam.size() - 2], l1 = lam.back();
      double f0, f1;
      int rc;
      if (!f.empty()) f0 = f.back();
      else {
        if ((rc = lineValue(IDFT_LS_CP, l0, f0))) return rc;
        f.push_back(f0);
      }
      if (std::sqrt(std::fabs(f0)) < tol) { lam.pop_back(); break; }
      if ((rc = lineValue(IDFT_LS_CP, l1, f1))) return rc;
      f.push_back(f1);
      if (std::sqrt(std::fabs(f1)) < tol) break;
      if (std::fabs((f1 - f0) / f1) < 1e-15) return IDFT_ENOCONV;
      double slope = (f1 - f0) / (l1 - l0);
      if (slope > 0.0) slope = -slope;
      double next = l1 - f1 / slope;
      if (next < 0.0) next = l1 + f1 / slope;
      lam.push_back(next);
    }
    return IDFT_OK;
  }

  // solveLineSearchSecantLoss / SecantForceNorm (BFGSInverseDFTSolver.cpp:305-463)
  int lineSearchSecant(int type, std::vector<double>& lam, std::vector<double>& f, int maxIter) {
    double lOld = lam[0], lNew = lam[1], lMid = 0.5 * (lNew + lOld), fOld = f[0], fMid, fNew;
    for (int i = 0; i < maxIter; ++i) {
      int rc;
      if ((rc = lineValue(type, lMid, fMid))) return rc;
      if ((rc = lineValue(type, lNew, fNew))) return rc;
      lNew = 0.5 * (lNew + lOld);
      lMid = 0.5 * (lNew + lOld);
      const double dl = lNew - lOld;
      const double dF = (3. * fNew - 4. * fMid + fOld) / dl;
      const double dFOld = (-3. * fOld + 4. * fMid - fNew) / dl;
      const double d2F = (dF - dFOld) / dl;
      double upd;
      if (d2F > 0.) upd = lNew - dF / d2F;
      else if (d2F < 0.) upd = lNew + dF / d2F;
      else break;
      if (upd < 0.0) upd = 0.5 * (lNew + lOld);
      lOld = lNew;
      lNew = upd;
      fOld = fNew;
      lMid = 0.5 * (lNew + lOld);
      lam.push_back(lNew);
      f.push_back(fNew);
    }
    return IDFT_OK;
  }

  idft_ctx* c_;
  int64_t n_;
  int history_;
  Vec x_, g_, p_;
  std::deque<Vec> Y_, S_;
  std::deque<double> rho_;
};

}  // namespace

extern "C" int idft_bfgs_solve(idft_ctx* c, int32_t ls_type, int32_t max_iter, double tol,
                               double ls_tol, int32_t history, int32_t n_line_search,
                               double* loss_hist, double* gnorm_hist, double* time_hist,
                               int32_t* n_done) {
  if (!c || max_iter < 0 || history < 1 || n_line_search < 0) return IDFT_EINVAL;
  if (ls_type < IDFT_LS_SECANT_LOSS || ls_type > IDFT_LS_CP) return IDFT_EINVAL;
  idft_sizes sz;
  int rc = idft_get_sizes(c, &sz);
  if (rc) return rc;
  try {
    Lbfgs solver(c, sz.n_vxc, history);
    return solver.run(ls_type, max_iter, tol, ls_tol, n_line_search, loss_hist, gnorm_hist,
                      time_hist, n_done);
  } catch (const DotFailure& f) {
    return f.rc;
  } catch (const std::exception&) {
    return IDFT_ENOMEM;
  }
}
