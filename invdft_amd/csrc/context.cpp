// Host orchestration of the MI355X inverse-DFT hot path: context set-up, the
// ChFSI eigensolve, density/loss, the adjoint problem contract, MINRES and
// the force.  All N-sized work runs in HIP kernels (kernels.hip) on the
// context's stream; the host handles only B x B dense algebra and scalars.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <type_traits>

#include "../../include/idft_setup.h"
#include <rocblas/rocblas.h>

#include "common.h"
#include "setup.h"
#include "sf_tables.h"

using namespace idft;

struct idft_mv {
  idft_ctx* ctx = nullptr;
  int ncols = 0;
  DevBuf<double> own;     // storage of idft_mv_create
  double* ext = nullptr;  // caller's device storage (idft_mv_wrap), not freed here
  double* p() const { return ext ? ext : own.p; }
};

namespace {

double wall_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct AdjointState {
  bool ready = false;
  int B = 0;
  idft_mv* psi = nullptr;
  idft_mv* x = nullptr;
  std::vector<double> occ, eig;
  std::vector<int> pairs;
  double fermi = 0.0;
  DevBuf<int> dpairs;
  DevBuf<double> docc4, dnegEig, uJxW;
  std::vector<double> Mu;
  // sum_q u JxW psi_i psi_j of the pairs, for the current psi and u (set by
  // computeRhs, reused by distributeX)
  std::vector<double> pairQ;
  bool havePairQ = false;
};

// Work buffers of one MINRES precision.
template <typename T>
struct MinresWork {
  DevBuf<T> r1, r2, y, w1, w, rhs;
  void ensure(size_t n) {
    r1.resize(n); r2.resize(n); y.resize(n); w1.resize(n); w.resize(n); rhs.resize(n);
  }
};

}  // namespace

struct idft_ctx {
  idft_desc d{};
  Tables t;
  Geo g{};
  DevTables dt;
  hipStream_t s = nullptr;
  std::string err;
  int B = 0;
  int64_t nQ = 0;
  bool mixed = false;
  bool verbose = false;  // IDFT_VERBOSE set: solver progress on stderr
  // static per-node / per-quadrature data
  DevBuf<double> S, sqrtM, invSqrtM, Djac, DjacSqrt, Mdiag, jxw;
  DevBuf<float> S32;
  DevBuf<unsigned char> bnd, adjMask;
  DevBuf<double> potBase, rhoT, weight, veff, rho, u;
  std::vector<double> hMdiag;
  // Hamiltonian
  DevBuf<double> H64, Yc64, W64;
  DevBuf<float> H32, Yc32, W32;
  GemmShape shape64, shape32;  // cell-product launch shapes (autotuned)
  bool haveH = false;          // operator of the current potential set (W64/W32)
  bool storedH = false;        // H64/H32 built for it (only the stored-H kernels read them)
  // eigen state (Loewdin vectors) and work
  DevBuf<double> X, X2, HX, Xtmp, T2buf, rref, xdelta, psiAll, force;
  DevBuf<float> F0, F1, F2, Rf;
  DevBuf<double> ydev;
  DevBuf<float> Xck;    // FP32 checkpoint of X between passes (psi_checkpoint)
  bool checkpointed = false;  // X released, Xck holds it
  bool haveHX = false;  // HX == H~ X for the current X and Hamiltonian
  DevBuf<double> partial, small, small2, occD;
  std::vector<double> eig, occ, resid;
  double fermi = 0.0, bup = 0.0, blow = 0.0;
  bool haveX = false, haveRitz = false, needRR = false;
  // inverse-problem state
  int64_t counter = 0;
  double tolCheb = 1e-8, adjointTol = 1e-11, lossPrev = 0.0, loss = 0.0;
  bool haveForward = false;
  std::vector<double> pot0, potSol;
  // v_xc output of getForceVector and L-BFGS checkpoint settings (idft_io.h)
  std::string vxcFolder, vxcPostfix;
  int vxcFreq = 0;
  std::vector<double> potEval;  // input potential of the current evaluation
  BfgsCheckpoint ckpt;
  idft_stats st{};
  AdjointState adj;
  MinresWork<double> mw64;
  DevBuf<double> mst, mx, mtmp, mnegc, mcell;
  DevBuf<int> mflags, msel;
  idft_mv mvPsi, mvX;

  // rocBLAS handle on s (created on first use), for the Gram and X Q products
  // of blocks of kBlasCols columns or more: there they are GEMMs bound by
  // arithmetic (N B^2 multiply-adds), which the library's DGEMM runs on the
  // matrix cores.  Atomics off: the split-K reductions sum in a fixed order.
  static constexpr int kBlasCols = 24;
  rocblas_handle blas = nullptr;
  rocblas_handle blas_handle() {
    if (!blas) {
      if (rocblas_create_handle(&blas) != rocblas_status_success) {
        blas = nullptr;
        throw Error(IDFT_EHIP, "rocblas_create_handle failed");
      }
      if (rocblas_set_stream(blas, s) != rocblas_status_success ||
          rocblas_set_atomics_mode(blas, rocblas_atomics_not_allowed) != rocblas_status_success)
        throw Error(IDFT_EHIP, "rocblas handle set-up failed");
    }
    return blas;
  }

  void sync() { IDFT_HIP(hipStreamSynchronize(s)); }
  size_t NB(int b) const { return (size_t)t.N * b; }

  // ---------------------------------------------------------------- apply
  // Y += alpha S H S X (Loewdin Hamiltonian), FP64 or FP32 cell matrices.
  // One application of the Loewdin cell operator with a fused epilogue
  // (see ApplyEpilogue); returns the number of dot-partial rows.
  template <typename T>
  int apply(const T* Xin, int b, const ApplyEpilogue<T>& ep) {
    if constexpr (std::is_same<T, double>::value) {
      Yc64.resize(cell_scratch_elems(g, b));
      st.applies_fp64++;
      st.apply_cols_fp64 += b;
      return launch_cell_apply<double, double>(g, op64(), S.p, S.p, Xin, Yc64.p, b, ep, shape64,
                                               s);
    } else {
      Yc32.resize(cell_scratch_elems(g, b));
      st.applies_fp32++;
      st.apply_cols_fp32 += b;
      return launch_cell_apply<float, float>(g, op32(), S32.p, S32.p, Xin, Yc32.p, b, ep, shape32,
                                             s);
    }
  }
  // FP32 cell operator (FP32 H_e or weights, arithmetic and cell scratch) on
  // FP64 vectors: input scaled and rounded once, gathered sums and epilogue
  // in FP64.
  int apply_f32op(const double* Xin, int b, const ApplyEpilogue<double>& ep) {
    Yc32.resize(cell_scratch_elems(g, b));
    st.applies_fp32++;
    st.apply_cols_fp32 += b;
    return launch_cell_apply<float, double>(g, op32(), S32.p, S.p, Xin, Yc32.p, b, ep, shape32, s);
  }
  // The cell operator at each precision for the current shapes; the stored
  // cell matrices are built on first use by a stored-H kernel.
  CellOp<double> op64() {
    if (shape64.kernel != CK_SUM_FACTOR) ensure_stored_h();
    CellOp<double> o;
    o.H = H64.p;
    o.W = W64.p;
    o.kin_scale = 0.25 * t.h;
    return o;
  }
  CellOp<float> op32() {
    if (!mixed) throw Error(IDFT_ESTATE, "FP32 Hamiltonian not built (precision is FP64)");
    if (shape32.kernel != CK_SUM_FACTOR) ensure_stored_h();
    CellOp<float> o;
    o.H = H32.p;
    o.W = W32.p;
    o.kin_scale = 0.25 * t.h;
    return o;
  }
  // Y = alpha S H S X + beta Y
  template <typename T>
  void happly(const T* Xv, T* Yv, int b, double alpha, double beta = 0.0) {
    ApplyEpilogue<T> ep;
    ep.Y = Yv;
    ep.alpha = alpha;
    ep.beta = beta;
    apply<T>(Xv, b, ep);
  }
  void happly64(const double* Xv, double* Yv, int b, double alpha, double beta = 0.0) {
    happly<double>(Xv, Yv, b, alpha, beta);
  }
  void happly32(const float* Xv, float* Yv, int b, double alpha, double beta = 0.0) {
    happly<float>(Xv, Yv, b, alpha, beta);
  }

  // ------------------------------------------------------- reductions
  std::vector<double> coldot(const double* a, const double* b2, int b, const double* w = nullptr) {
    partial.resize((size_t)1024 * std::max(b, 1));
    small.resize(std::max(b, 16));
    const int nb = launch_coldot_partial<double>(t.N, b, a, b2, w, partial.p, s);
    launch_sum_partials(nb, b, partial.p, small.p, s);
    std::vector<double> out(b);
    small.download(out.data(), b, s);
    sync();
    return out;
  }
  // G[i][j] = sum_n a[n][i] b2[n][j] (row-major B x B)
  std::vector<double> gram(const double* a, const double* b2, int b) {
    small.resize((size_t)b * b);
    if (b >= kBlasCols) {
      // column-major: a and b2 are B x N (leading dimension B), and
      // G^T = b2 a^T is G read row-major.  The node sum is split into nk
      // chunks of klen (8192) nodes (one batched DGEMM, one B x B partial
      // each, plus one for the remainder), then the partials are summed in
      // order: a single product would leave the long reduction to few
      // work-groups.
      const int64_t klen = std::min<int64_t>(8192, t.N);
      const int64_t nk = t.N / klen, rem = t.N - nk * klen;
      const int64_t bb = (int64_t)b * b;
      partial.resize((size_t)(nk + 1) * bb);
      const double one = 1.0, zero = 0.0;
      if (rocblas_dgemm_strided_batched(blas_handle(), rocblas_operation_none,
                                        rocblas_operation_transpose, b, b, (rocblas_int)klen, &one,
                                        b2, b, (rocblas_stride)(klen * b), a, b,
                                        (rocblas_stride)(klen * b), &zero, partial.p, b,
                                        (rocblas_stride)bb, (rocblas_int)nk) !=
          rocblas_status_success)
        throw Error(IDFT_EHIP, "rocblas_dgemm_strided_batched failed (Gram)");
      int np = (int)nk;
      if (rem > 0) {
        const size_t off = (size_t)(nk * klen) * b;
        if (rocblas_dgemm(blas_handle(), rocblas_operation_none, rocblas_operation_transpose, b, b,
                          (rocblas_int)rem, &one, b2 + off, b, a + off, b, &zero,
                          partial.p + (size_t)nk * bb, b) != rocblas_status_success)
          throw Error(IDFT_EHIP, "rocblas_dgemm failed (Gram remainder)");
        ++np;
      }
      launch_sum_partials(np, (int)bb, partial.p, small.p, s);
    } else {
      partial.resize((size_t)gram_partial_blocks(t.N, b) * b * b);
      launch_gram(t.N, b, a, b2, partial.p, small.p, s);
    }
    std::vector<double> G((size_t)b * b);
    small.download(G.data(), G.size(), s);
    sync();
    return G;
  }
  // out = beta out + diag(ns) a Q  (Q row-major B x B)
  void right_mult(const double* a, const std::vector<double>& Q, double* out, int b,
                  const double* ns = nullptr, double beta = 0.0) {
    small2.resize(Q.size());
    small2.upload(Q.data(), Q.size(), s);
    if (b >= kBlasCols && !ns) {
      // column-major: out (B x N) = Q^T a, and Q row-major is Q^T
      // column-major
      const double one = 1.0;
      if (rocblas_dgemm(blas_handle(), rocblas_operation_none, rocblas_operation_none, b,
                        (rocblas_int)t.N, b, &one, small2.p, b, a, b, &beta, out, b) !=
          rocblas_status_success)
        throw Error(IDFT_EHIP, "rocblas_dgemm failed (X Q)");
      return;
    }
    launch_right_mult(t.N, b, a, small2.p, out, ns, beta, s);
  }
  void zero_boundary(double* v, int b) { launch_zero_rows(t.N, b, v, bnd.p, s); }

  // ------------------------------------------------------ Hamiltonian
  // Operator of the current v_eff: the quadrature weights w = v_eff JxW
  // (FP64, and FP32 in MIXED); the stored H_e waits for a kernel that reads it.
  void build_h() {
    W64.resize(nQ);
    launch_mul_q(nQ, t.nq, veff.p, jxw.p, W64.p, s);
    if (mixed) {
      W32.resize(nQ);
      launch_convert_d2f(nQ, W32.p, W64.p, s);
    }
    haveH = true;
    storedH = false;
    haveHX = false;
  }
  void ensure_stored_h() {
    if (storedH) return;
    if (!haveH) throw Error(IDFT_ESTATE, "no Hamiltonian: call set_potential first");
    const double hh = 0.5 * t.h;
    H64.resize((size_t)t.ncell * hamiltonian_tile_elems(t.p));
    if (mixed) H32.resize((size_t)t.ncell * hamiltonian_tile_elems(t.p));
    launch_build_hamiltonian(g, dt, hh, veff.p, jxw.p, H64.p, mixed ? H32.p : nullptr, s);
    storedH = true;
  }
  // Frees the stored cell matrices when neither precision's kernel reads them.
  void release_unused_h() {
    if (shape64.kernel != CK_SUM_FACTOR || (mixed && shape32.kernel != CK_SUM_FACTOR)) return;
    H64.release();
    H32.release();
    storedH = false;
  }
  // Device ms per operator application on ncols random columns (HIP events
  // on the context stream, 2 warm-up + reps applications), FP32 or FP64
  // operator.  scaled: the input carries a node and a column scale, as in the
  // MINRES steps (FP32 operator on FP64 vectors there); ms_gemm (optional):
  // the cell-product kernel alone, unscaled.
  // Random input blocks of the profiling calls, made once per width and kept
  // while the autotuner times its shapes: kProfBlocks different blocks taken
  // in turn, so that an application does not find its input in the caches
  // (as in the solvers, where other vector kernels run between applications).
  DevBuf<double> profX, profY, profOnes;
  DevBuf<float> profXf, profYf;
  int profCols = 0;
  void release_profile_block() {
    profX.release();
    profY.release();
    profOnes.release();
    profXf.release();
    profYf.release();
    profCols = 0;
  }
  double profile_ms(int ncols, bool f32, int reps, bool scaled, double* ms_gemm) {
    constexpr int kProfBlocks = 3;
    const size_t n = NB(ncols);
    if (profCols != ncols) {
      std::vector<double> h(n * kProfBlocks);
      uint64_t sd = 12345;
      for (auto& v : h) v = rand_pm1(sd);
      profX.upload(h.data(), h.size(), s);
      profY.resize(n);
      const std::vector<double> one(ncols, 1.0);
      profOnes.upload(one.data(), one.size(), s);
      profXf.release();
      profYf.release();
      profCols = ncols;
    }
    if (f32 && !scaled && !profXf.p) {
      profXf.resize(n * kProfBlocks);
      profYf.resize(n);
      launch_convert_d2f((int64_t)(n * kProfBlocks), profXf.p, profX.p, s);
    }
    int blk = 0;  // input block of the next application
    auto xdp = [&] { return profX.p + n * (blk++ % kProfBlocks); };
    auto xfp = [&] { return profXf.p + n * (blk++ % kProfBlocks); };
    DevBuf<double>& yd = profY;
    DevBuf<double>& ones = profOnes;
    DevBuf<float>& yf = profYf;
    auto once = [&] {
      if (scaled) {
        ApplyEpilogue<double> ep;
        ep.Y = yd.p;
        ep.xn = Djac.p;
        ep.xcol = ones.p;
        if (f32) apply_f32op(xdp(), ncols, ep);
        else apply<double>(xdp(), ncols, ep);
      } else if (f32) {
        happly32(xfp(), yf.p, ncols, 1.0);
      } else {
        happly64(xdp(), yd.p, ncols, 1.0);
      }
    };
    auto gemm = [&] {
      if (f32) launch_cell_gemm<float>(g, op32(), S32.p, xfp(), nullptr, Yc32.p, ncols, shape32, s);
      else launch_cell_gemm<double>(g, op64(), S.p, xdp(), nullptr, Yc64.p, ncols, shape64, s);
    };
    hipEvent_t e0, e1;
    IDFT_HIP(hipEventCreate(&e0));
    IDFT_HIP(hipEventCreate(&e1));
    auto timed = [&](auto&& fn) {
      fn();
      fn();
      IDFT_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < reps; ++r) fn();
      IDFT_HIP(hipEventRecord(e1, s));
      IDFT_HIP(hipEventSynchronize(e1));
      float el = 0.0f;
      IDFT_HIP(hipEventElapsedTime(&el, e0, e1));
      return (double)el / reps;
    };
    const double ms = timed(once);
    if (ms_gemm && !scaled) *ms_gemm = timed(gemm);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
  }
  // The sum-factorised kernel's compile-time 1D tables against this
  // context's FE tables (fe.cpp): they must agree to rounding.
  void check_sf_tables() const {
    double d = 0.0;
    auto cmp = [&](const double* L, const double* M, const double* D) {
      for (int i = 0; i < t.q1 * t.n1; ++i) d = std::max(d, std::fabs(L[i] - t.Lq[i]));
      for (int i = 0; i < t.n1 * t.n1; ++i) {
        d = std::max(d, std::fabs(M[i] - t.M1[i]));
        d = std::max(d, std::fabs(D[i] - t.D1[i]) / 16.0);
      }
    };
    switch (t.p) {
      case 2: cmp(SfC<2>::v.L, SfC<2>::v.M, SfC<2>::v.D); break;
      case 3: cmp(SfC<3>::v.L, SfC<3>::v.M, SfC<3>::v.D); break;
      case 4: cmp(SfC<4>::v.L, SfC<4>::v.M, SfC<4>::v.D); break;
      case 5: cmp(SfC<5>::v.L, SfC<5>::v.M, SfC<5>::v.D); break;
      default: break;
    }
    if (!(d < 1e-13)) throw Error(IDFT_EINVAL, "sum-factorisation tables disagree with the FE tables");
  }

  void set_potential(const double* pot_host) {
    DevBuf<double> pot;
    pot.upload(pot_host, (size_t)t.NV, s);
    veff.resize(nQ);
    launch_vxc_to_quad(g, dt, pot.p, potBase.p, veff.p, s);
    build_h();
    sync();
  }

  // ------------------------------------------------------------ ChFSI
  void init_subspace() {
    std::vector<double> h(NB(B));
    uint64_t sd = d.seed;
    std::vector<unsigned char> hb((size_t)t.N);
    bnd.download(hb.data(), hb.size(), s);
    sync();
    for (int64_t n = 0; n < t.N; ++n)
      for (int j = 0; j < B; ++j) {
        const double r = rand_pm1(sd);
        h[(size_t)n * B + j] = hb[n] ? 0.0 : r;
      }
    X.upload(h.data(), h.size(), s);
    haveX = true;
    haveRitz = false;
    haveHX = false;
  }

  template <typename T>
  void lanczos_t(double& upper, double& lowest) {
    const int k = std::max(2, d.lanczos_steps);
    std::vector<double> hv((size_t)t.N), hb;
    std::vector<unsigned char> mask((size_t)t.N);
    bnd.download(mask.data(), mask.size(), s);
    sync();
    uint64_t sd = d.seed ^ 0x5bd1e995ULL;
    double nrm = 0.0;
    for (int64_t n = 0; n < t.N; ++n) {
      const double r = rand_pm1(sd);
      hv[n] = mask[n] ? 0.0 : r;
      nrm += hv[n] * hv[n];
    }
    nrm = std::sqrt(nrm);
    for (auto& v : hv) v /= nrm;
    DevBuf<double> v64(t.N);
    v64.upload(hv.data(), hv.size(), s);
    DevBuf<T> v(t.N), vp(t.N), w(t.N);
    convert_to(v64.p, v.p, t.N);
    launch_lincomb<T>(t.N, 1, vp.p, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    std::vector<double> al, be;
    double beta = 0.0;
    for (int j = 0; j < k; ++j) {
      happly<T>(v.p, w.p, 1, 1.0);
      const double a = coldot_t<T>(v.p, w.p, 1)[0];
      launch_lincomb<T>(t.N, 1, w.p, 1.0, v.p, -a, nullptr, vp.p, -beta, nullptr, nullptr, s);
      const double b2 = std::sqrt(std::max(coldot_t<T>(w.p, w.p, 1)[0], 0.0));
      al.push_back(a);
      be.push_back(b2);
      beta = b2;
      if (b2 == 0.0) break;
      std::swap(vp.p, v.p);
      launch_lincomb<T>(t.N, 1, v.p, 0.0, w.p, 1.0 / b2, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    }
    const int kk = (int)al.size();
    std::vector<double> T3((size_t)kk * kk, 0.0), ev, V;
    for (int i = 0; i < kk; ++i) {
      T3[(size_t)i * kk + i] = al[i];
      if (i + 1 < kk) T3[(size_t)i * kk + i + 1] = T3[(size_t)(i + 1) * kk + i] = be[i];
    }
    sym_eig(kk, T3, ev, V);
    upper = ev[kk - 1] + std::fabs(be[kk - 1]);
    lowest = ev[0];
  }

  template <typename T>
  std::vector<double> coldot_t(const T* a, const T* b2, int b) {
    partial.resize((size_t)1024 * std::max(b, 1));
    small.resize(std::max(b, 16));
    const int nb = launch_coldot_partial<T>(t.N, b, a, b2, nullptr, partial.p, s);
    launch_sum_partials(nb, b, partial.p, small.p, s);
    std::vector<double> out(b);
    small.download(out.data(), b, s);
    sync();
    return out;
  }
  void convert_to(const double* a, double* o, int64_t n) {
    IDFT_HIP(hipMemcpyAsync(o, a, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  void convert_to(const double* a, float* o, int64_t n) { launch_convert_d2f(n, o, a, s); }

  // Scaled Chebyshev filter (Zhou-Saad) of X in place, FP64 or FP32.
  template <typename T>
  void filter_t(T* x0, T* y0, T* z0, int b, int deg, double a, double bb, double a0, T** result) {
    const double e = 0.5 * (bb - a), c = 0.5 * (bb + a);
    double sigma = e / (a0 - c);
    const double sigma1 = sigma, tau = 2.0 / sigma1;
    T *xk = x0, *yk = y0, *zk = z0;
    {
      ApplyEpilogue<T> ep;
      ep.Y = yk; ep.alpha = sigma1 / e;
      ep.X = xk; ep.a = -c * sigma1 / e;
      apply<T>(xk, b, ep);
    }
    for (int k = 2; k <= deg; ++k) {
      const double sigma2 = 1.0 / (tau - sigma);
      const double f = 2.0 * sigma2 / e;
      ApplyEpilogue<T> ep;
      ep.Y = zk; ep.alpha = f;
      ep.X = yk; ep.a = -f * c;
      ep.Z = xk; ep.b = -sigma * sigma2;
      apply<T>(yk, b, ep);
      T* tmp = xk; xk = yk; yk = zk; zk = tmp;
      sigma = sigma2;
    }
    *result = yk;
  }

  void chebyshev_filter_X(double* Xd, int b, int deg, double a, double bb, double a0, bool fp32) {
    if (fp32) {
      F0.resize(NB(b)); F1.resize(NB(b)); F2.resize(NB(b));
      launch_convert_d2f((int64_t)NB(b), F0.p, Xd, s);
      float* res = nullptr;
      filter_t<float>(F0.p, F1.p, F2.p, b, deg, a, bb, a0, &res);
      launch_convert_f2d((int64_t)NB(b), Xd, res, s);
    } else {
      X2.resize(NB(b)); HX.resize(NB(b));
      double* res = nullptr;
      Xtmp.resize(NB(b));
      IDFT_HIP(hipMemcpyAsync(Xtmp.p, Xd, NB(b) * sizeof(double), hipMemcpyDeviceToDevice, s));
      filter_t<double>(Xtmp.p, X2.p, HX.p, b, deg, a, bb, a0, &res);
      IDFT_HIP(hipMemcpyAsync(Xd, res, NB(b) * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
  }

  // Mixed-precision filter in defect form.  With Ritz pairs (X, L) of the
  // current Hamiltonian and R = H~X - X L (FP64), D_k = Y_k - X y_k(L) obeys
  //   D_{k+1} = f (H~ - c) D_k - g D_{k-1} + f R diag(y_k(L)),  D_0 = 0,
  //   D_1 = (sigma1/e) R,
  // where y_k are the scalar filter polynomials.  D is carried in FP32 on the
  // FP32 cell operator; Y_m = X y_m(L) + D_m is formed in FP64, so the FP32
  // rounding error is proportional to the residual and vanishes at convergence.
  void chebyshev_defect_filter(int deg, double a, double bb, double a0) {
    const int b = B;
    const size_t n = NB(b);
    F0.resize(n); F1.resize(n); F2.resize(n); Rf.resize(n); X2.resize(n);
    // R = HX - X diag(eig)
    small2.resize(b);
    small2.upload(eig.data(), b, s);
    launch_lincomb_d2f(t.N, b, Rf.p, HX.p, 1.0, nullptr, X.p, -1.0, small2.p, s);
    const double e = 0.5 * (bb - a), c = 0.5 * (bb + a);
    double sigma = e / (a0 - c);
    const double sigma1 = sigma, tau = 2.0 / sigma1;
    // scalar sequences y_k(lambda_j), k = 0..deg
    std::vector<double> ys((size_t)(deg + 1) * b);
    for (int j = 0; j < b; ++j) { ys[j] = 1.0; ys[b + j] = (sigma1 / e) * (eig[j] - c); }
    {
      double sg = sigma;
      for (int k = 1; k < deg; ++k) {
        const double s2 = 1.0 / (tau - sg), f = 2.0 * s2 / e, gk = sg * s2;
        for (int j = 0; j < b; ++j)
          ys[(size_t)(k + 1) * b + j] = f * (eig[j] - c) * ys[(size_t)k * b + j] - gk * ys[(size_t)(k - 1) * b + j];
        sg = s2;
      }
    }
    ydev.resize(ys.size());
    ydev.upload(ys.data(), ys.size(), s);
    float *dm = F0.p, *dk = F1.p, *dn = F2.p;
    launch_lincomb<float>(t.N, b, dk, 0.0, Rf.p, sigma1 / e, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    for (int k = 1; k < deg; ++k) {
      const double s2 = 1.0 / (tau - sigma), f = 2.0 * s2 / e, gk = sigma * s2;
      ApplyEpilogue<float> ep;
      ep.Y = dn; ep.alpha = f;
      ep.X = dk; ep.a = -f * c;
      // D_0 = 0: the first step has no -g D_{k-1} term (dm holds nothing yet)
      ep.Z = k > 1 ? dm : nullptr; ep.b = -gk;
      ep.W = Rf.p; ep.c = f; ep.cc = ydev.p + (size_t)k * b;
      apply<float>(dk, b, ep);
      float* tmp = dm; dm = dk; dk = dn; dn = tmp;
      sigma = s2;
    }
    launch_lincomb_df2d(t.N, b, X.p, X.p, 1.0, ydev.p + (size_t)deg * b, dk, 1.0, nullptr, s);
  }

  // Two sweeps of Cholesky-QR (SVQB when the Gram matrix is not positive definite).
  void orthonormalize() {
    X2.resize(NB(B));
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<double> G = gram(X.p, X.p, B), L = G, Q((size_t)B * B, 0.0);
      if (cholesky_lower(B, L)) {
        for (int j = 0; j < B; ++j) {
          std::vector<double> yv(B, 0.0);
          for (int i = 0; i < B; ++i) {
            double sv = (i == j) ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) sv -= L[(size_t)i * B + k] * yv[k];
            yv[i] = sv / L[(size_t)i * B + i];
          }
          for (int i = 0; i < B; ++i) Q[(size_t)j * B + i] = yv[i];
        }
      } else {
        std::vector<double> lam, V;
        sym_eig(B, G, lam, V);
        const double tau = 1e-14 * std::max(lam[B - 1], 1e-300);
        for (int i = 0; i < B; ++i)
          for (int j = 0; j < B; ++j)
            Q[(size_t)i * B + j] = V[(size_t)i * B + j] / std::sqrt(std::max(lam[j], tau));
      }
      right_mult(X.p, Q, X2.p, B);
      std::swap(X.p, X2.p);
    }
  }

  void rayleigh_ritz() {
    HX.resize(NB(B)); X2.resize(NB(B));
    happly64(X.p, HX.p, B, 1.0);
    std::vector<double> Hp = gram(X.p, HX.p, B), Q;
    for (int i = 0; i < B; ++i)
      for (int j = i + 1; j < B; ++j) {
        const double v = 0.5 * (Hp[(size_t)i * B + j] + Hp[(size_t)j * B + i]);
        Hp[(size_t)i * B + j] = Hp[(size_t)j * B + i] = v;
      }
    sym_eig(B, Hp, eig, Q);
    right_mult(X.p, Q, X2.p, B);
    std::swap(X.p, X2.p);
    right_mult(HX.p, Q, X2.p, B);
    std::swap(HX.p, X2.p);
    small2.resize(B);
    small2.upload(eig.data(), B, s);
    partial.resize((size_t)1024 * B);
    small.resize(std::max(B, 16));
    const int nb = launch_residual_partial(t.N, B, HX.p, X.p, small2.p, partial.p, s);
    launch_sum_partials(nb, B, partial.p, small.p, s);
    resid.assign(B, 0.0);
    small.download(resid.data(), B, s);
    sync();
    for (auto& r : resid) r = std::sqrt(std::max(r, 0.0));
    haveRitz = true;
    haveHX = true;
  }

  void update_occupancy() {
    fermi = fermi_level(eig, d.n_electrons, d.temperature);
    occ.assign(B, 0.0);
    for (int i = 0; i < B; ++i) occ[i] = fd_occ(eig[i], fermi, d.temperature);
  }

  void chfsi_pass() {
    double a0, a;
    if (haveRitz) { a0 = eig[0]; a = eig[B - 1]; }
    else { a0 = blow; a = a0 + 0.1 * (bup - a0); }
    if (mixed && haveRitz) {
      if (!haveHX) rayleigh_ritz();  // Ritz pairs of the current Hamiltonian
      a0 = eig[0]; a = eig[B - 1];
      chebyshev_defect_filter(d.cheb_degree, a, bup, a0);
    } else {
      chebyshev_filter_X(X.p, B, d.cheb_degree, a, bup, a0, mixed);
    }
    haveHX = false;
    orthonormalize();
    rayleigh_ritz();
    for (double r : resid)
      if (!std::isfinite(r)) throw Error(IDFT_ENAN, "non-finite Ritz residual");
  }

  void solve_passes(double tol, int maxPasses, int* passesOut, double* maxResOut) {
    if (!haveH) throw Error(IDFT_ESTATE, "no Hamiltonian: call set_potential first");
    if (!haveX) init_subspace();
    if (mixed) lanczos_t<float>(bup, blow);
    else lanczos_t<double>(bup, blow);
    st.spectrum_upper_bound = bup;
    if (needRR) {
      orthonormalize();
      rayleigh_ritz();
      needRR = false;
    }
    int pass = 0;
    double maxRes = 0.0;
    do {
      chfsi_pass();
      update_occupancy();
      maxRes = 0.0;
      int homo = 0;
      for (int i = 0; i < B; ++i)
        if (eig[i] < fermi + 1e-3 && homo < i) homo = i;
      for (int i = 0; i < B; ++i)
        if (occ[i] > d.frac_occ_tol || i <= homo + d.additional_states)
          maxRes = std::max(maxRes, resid[i]);
      pass++;
    } while (maxRes > tol && pass < maxPasses);
    st.cheb_passes = pass;
    st.max_residual = maxRes;
    st.fermi_energy = fermi;
    if (passesOut) *passesOut = pass;
    if (maxResOut) *maxResOut = maxRes;
  }

  void physical_psi() {
    psiAll.resize(NB(B));
    launch_lincomb<double>(t.N, B, psiAll.p, 0.0, X.p, 1.0, nullptr, nullptr, 0.0, nullptr, S.p, s);
  }

  void density_into(double* out) {
    occD.resize(B);
    occD.upload(occ.data(), B, s);
    launch_density(g, dt, X.p, S.p, B, occD.p, out, s);
  }

  // ----------------------------------------------------- adjoint problem
  std::vector<double> pair_quad_sum(const double* A, const double* Bv, int b, const double* wq,
                                    int64_t wstride) {
    const int np = (int)adj.pairs.size() / 2;
    if (np == 0) return {};
    partial.resize((size_t)t.ncell * np);
    small.resize(std::max(np, 16));
    launch_pair_quad(g, dt, A, Bv, b, adj.dpairs.p, np, wq, wstride, partial.p, s);
    launch_sum_partials((int)t.ncell, np, partial.p, small.p, s);
    std::vector<double> out(np);
    small.download(out.data(), np, s);
    sync();
    return out;
  }

  void compute_mu(const std::vector<double>& occv) {
    const int Bb = adj.B;
    if (!adj.havePairQ) {
      adj.pairQ = pair_quad_sum(adj.psi->p(), adj.psi->p(), Bb, adj.uJxW.p, t.nq);
      adj.havePairQ = true;
    }
    const std::vector<double>& v = adj.pairQ;
    adj.Mu.assign((size_t)Bb * Bb, 0.0);
    for (size_t k = 0; k < v.size(); ++k) {
      const int i = adj.pairs[2 * k], j = adj.pairs[2 * k + 1];
      adj.Mu[(size_t)i * Bb + j] += 2.0 * occv[i] * v[k];
    }
  }

  void update_input_psi(idft_mv* psi, const double* occv, const double* u_dev_or_null,
                        const double* u_host, const int32_t* pairs, int np, double fermiE,
                        double* eigv, int b) {
    if (!psi || psi->ncols != b || b <= 0) throw Error(IDFT_EINVAL, "psi block size mismatch");
    AdjointState& A = adj;
    A.B = b; A.psi = psi; A.fermi = fermiE; A.x = nullptr;
    A.havePairQ = false;
    zero_boundary(psi->p(), b);
    A.occ.assign(occv, occv + b);
    std::vector<double> occ4(b);
    for (int i = 0; i < b; ++i) occ4[i] = 4.0 * occv[i];
    A.pairs.assign(pairs, pairs + 2 * np);
    for (int v : A.pairs)
      if (v < 0 || v >= b) throw Error(IDFT_EINVAL, "pair index out of range");
    A.eig.assign(eigv, eigv + b);
    for (int i = 0; i < b; ++i) eigv[i] = -eigv[i];
    A.docc4.upload(occ4.data(), b, s);
    A.dnegEig.upload(eigv, b, s);
    if (np) A.dpairs.upload(A.pairs.data(), A.pairs.size(), s);
    A.uJxW.resize(nQ);
    if (u_dev_or_null) {
      launch_mul_q(nQ, t.nq, u_dev_or_null, jxw.p, A.uJxW.p, s);
    } else {
      DevBuf<double> uq;
      uq.upload(u_host, nQ, s);
      launch_mul_q(nQ, t.nq, uq.p, jxw.p, A.uJxW.p, s);
      sync();
    }
    A.ready = true;
  }

  void compute_rhs(idft_mv* x, double* rhs) {
    AdjointState& A = adj;
    if (!A.ready) throw Error(IDFT_ESTATE, "computeRhs before updateInputPsi");
    const int b = A.B;
    if (x && x->ncols != b) throw Error(IDFT_EINVAL, "block size mismatch");
    A.x = x;
    const double* psi = A.psi->p();
    // T2 = R psi, R = N^T diag(u JxW) N assembled over the cells
    T2buf.resize(NB(b));
    double* T2 = T2buf.p;
    launch_lincomb<double>(t.N, b, T2, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    for (int colour = 0; colour < 8; ++colour)
      launch_quad_operator(g, dt, psi, b, A.uJxW.p, T2, colour, s);
    // the pairs' quadrature sums sum_q u JxW psi_i psi_j = psi_i . (R psi)_j
    // (R assembled from the same quadrature), before T2's boundary rows are
    // cleared; more pairs than the node-dot kernel takes (256) are summed at
    // the quadrature points by compute_mu instead
    const int np = (int)A.pairs.size() / 2;
    A.havePairQ = false;
    if (np > 0 && np <= 256) {
      partial.resize((size_t)1024 * np);
      small.resize(std::max(np, 16));
      const int nb = launch_pairdot_partial(t.N, b, psi, T2, nullptr, A.dpairs.p, np, partial.p, s);
      launch_sum_partials(nb, np, partial.p, small.p, s);
      A.pairQ.resize(np);
      small.download(A.pairQ.data(), np, s);
      sync();
      A.havePairQ = true;
    } else if (np == 0) {
      A.pairQ.clear();
      A.havePairQ = true;
    }
    compute_mu(A.occ);
    // rhs = -2 sqrt(M) psi Mu   (column i: -2 sum_k Mu[k][i] M^{1/2} psi_k)
    std::vector<double> C((size_t)b * b);
    for (size_t k = 0; k < C.size(); ++k) C[k] = -2.0 * A.Mu[k];
    right_mult(psi, C, rhs, b, sqrtM.p, 0.0);
    // rhs += 4 f M^{-1/2} R psi
    zero_boundary(T2, b);
    launch_lincomb<double>(t.N, b, rhs, 1.0, T2, 1.0, A.docc4.p, nullptr, 0.0, nullptr, invSqrtM.p, s);
    zero_boundary(rhs, b);
    sync();
  }

  void vmult(double* Ax, double* x, int b) {
    if (!adj.ready) throw Error(IDFT_ESTATE, "vmult before updateInputPsi");
    zero_boundary(x, b);
    ApplyEpilogue<double> ep;
    ep.Y = Ax; ep.alpha = 1.0;
    ep.X = x; ep.a = 1.0; ep.ca = adj.dnegEig.p;
    apply<double>(x, b, ep);
  }

  // vmult with the cell-product scratch in the library's fixed device array
  // (noalloc.hip) instead of Yc64: no device allocation.  The operator must
  // need nothing built on first use (the sum-factorised kernel, or stored
  // cell matrices already built).  Calls that overlap on one GPU share the
  // array and race.
  void vmult_noalloc(double* Ax, double* x, int b) {
    if (!adj.ready) throw Error(IDFT_ESTATE, "vmult before updateInputPsi");
    if (!haveH) throw Error(IDFT_ESTATE, "no Hamiltonian: call set_potential first");
    if (shape64.kernel != CK_SUM_FACTOR && !storedH)
      throw Error(IDFT_ESTATE,
                  "the selected cell-product kernel reads stored cell matrices, which are not "
                  "built (building them allocates)");
    const int64_t need = cell_scratch_elems(g, b);
    if (need > kNoallocScratchElems)
      throw Error(IDFT_EINVAL, "cells*(p+1)^3*block = " + std::to_string(need) +
                                   " exceeds the fixed scratch of " +
                                   std::to_string(kNoallocScratchElems) + " doubles");
    double* scr = noalloc_scratch();
    zero_boundary(x, b);
    ApplyEpilogue<double> ep;
    ep.Y = Ax; ep.alpha = 1.0;
    ep.X = x; ep.a = 1.0; ep.ca = adj.dnegEig.p;
    CellOp<double> o;
    o.H = H64.p;
    o.W = W64.p;
    o.kin_scale = 0.25 * t.h;
    st.applies_fp64++;
    st.apply_cols_fp64 += b;
    launch_cell_apply<double, double>(g, o, S.p, S.p, x, scr, b, ep, shape64, s);
  }

  std::vector<double> dot_quad_wise(const double* v1, const double* v2, int b) {
    return pair_quad_sum(v1, v2, b, jxw.p, 0);
  }

  void distribute_x() {
    AdjointState& A = adj;
    if (!A.ready || !A.x) throw Error(IDFT_ESTATE, "distributeX before computeRhs");
    const int b = A.B;
    double* x = A.x->p();
    const double* psi = A.psi->p();
    launch_lincomb<double>(t.N, b, x, 0.0, x, 1.0, nullptr, nullptr, 0.0, nullptr, invSqrtM.p, s);
    zero_boundary(x, b);
    const int np = (int)A.pairs.size() / 2;
    std::vector<double> dot;
    if (d.gradient_mode == IDFT_GRADIENT_EXACT) {
      partial.resize((size_t)1024 * std::max(np, 1));
      small.resize(std::max(np, 16));
      const int nb = launch_pairdot_partial(t.N, b, x, psi, Mdiag.p, A.dpairs.p, np, partial.p, s);
      launch_sum_partials(nb, np, partial.p, small.p, s);
      dot.resize(np);
      small.download(dot.data(), np, s);
      sync();
    } else {
      dot = dot_quad_wise(x, psi, b);
    }
    compute_mu(std::vector<double>(b, 1.0));
    double sumD = 0.0, etaR = 0.0;
    for (int i = 0; i < b; ++i) {
      const double dF = -fd_occ_der(A.eig[i], A.fermi, d.temperature);
      sumD += dF;
      etaR += dF * A.Mu[(size_t)i * b + i];
    }
    const double eta = std::fabs(etaR) < 1e-16 ? 0.0 : etaR / sumD;
    std::vector<double> C((size_t)b * b, 0.0);  // x[:, i] += sum_j C[j][i] psi[:, j]
    for (int k = 0; k < np; ++k) {
      const int i = A.pairs[2 * k], j = A.pairs[2 * k + 1];
      const double dFi = fd_occ_der(A.eig[i], A.fermi, d.temperature);
      double corr = -dFi * A.Mu[(size_t)i * b + j];
      if (i == j) corr += dFi * eta;
      C[(size_t)j * b + i] += -dot[k] + corr;
    }
    right_mult(psi, C, x, b, nullptr, 1.0);
    sync();
  }

  // ------------------------------------------------------------ MINRES
  // Preconditioned MINRES (FP64 recurrence) from x = 0 (x accumulates), per-
  // column tolerance.  relax > 0 (MIXED): the FP64 cell operator is applied
  // until every open column's residual estimate phibar is below relax times
  // its tolerance, the FP32 cell operator after that -- the matvec accuracy
  // a Krylov method needs falls as its residual falls (inexact-Krylov
  // relaxation), so the late steps tolerate FP32 products.
  //
  // Only open columns are streamed: the recurrence runs on a compact block of
  // the columns not yet converged (`orig` maps its columns to the caller's),
  // repacked at the host convergence checks once enough columns have
  // finished, and a column whose right-hand side is already below its
  // tolerance (an unoccupied state: f = 0 makes b~ = 0) never enters it.
  // Each column's recurrence is independent of the others', so the iterates
  // are those of the full block.
  std::vector<int> minres_core(const double* rhs, double* x, int b, const std::vector<double>& tol,
                               int maxit, double relax) {
    MinresWork<double>& W = mw64;
    W.ensure(NB(b));
    mx.resize(NB(b));
    mtmp.resize(NB(b));
    mst.resize((size_t)MS_COUNT * b);
    mflags.resize(b);
    mst.upload(tol.data(), b, s);  // slot 0 temporarily
    IDFT_HIP(hipMemcpyAsync(mst.p + (size_t)MS_TOL * b, mst.p, b * sizeof(double),
                            hipMemcpyDeviceToDevice, s));
    const int64_t N = t.N;
    double *r1 = W.r1.p, *r2 = W.r2.p, *w1 = W.w1.p, *w = W.w.p, *xc = mx.p;
    launch_lincomb<double>(N, b, r1, 0.0, rhs, 1.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    launch_lincomb<double>(N, b, r2, 0.0, rhs, 1.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    launch_lincomb<double>(N, b, w1, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    launch_lincomb<double>(N, b, w, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    launch_lincomb<double>(N, b, xc, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    const bool canRelax = relax > 0.0 && mixed;
    // stored cell matrices (if a stored-H kernel is selected) built before
    // any step is captured into a graph
    (void)op64();
    if (canRelax) (void)op32();
    // dot partials of every block width the compaction may reach
    int nblk = 1024;
    for (int bw = 1; bw <= b; ++bw) {
      nblk = std::max(nblk, gather_blocks(g, bw, shape64.gather_rows, 8));
      if (canRelax) nblk = std::max(nblk, gather_blocks(g, bw, shape32.gather_rows, 8));
    }
    // Fused steps (sum-factorised cell products at every precision the solve
    // may use): alpha comes from the cell kernel's per-cell dots, so the
    // Lanczos recurrence runs inside the gather and y = H~ v is never stored.
    const bool fused = d.minres_fused != 0 && shape64.kernel == CK_SUM_FACTOR &&
                       (!canRelax || shape32.kernel == CK_SUM_FACTOR);
    partial.resize((size_t)nblk * b * (fused ? 3 : 1));
    if (fused) {
      mcell.resize((size_t)t.ncell * b);
      const int nb = launch_minres_start_dots(N, b, r1, Djac.p, partial.p, s);
      launch_minres_start(b, nb, partial.p, adj.dnegEig.p, mst.p, s, 2);
    } else {
      const int nb = launch_coldot_partial<double>(N, b, r1, r1, Djac.p, partial.p, s);
      launch_minres_start(b, nb, partial.p, adj.dnegEig.p, mst.p, s);
    }
    // the compact block: bc columns, orig[k] the caller's column of column k
    int bc = b;
    std::vector<int> orig(b);
    for (int j = 0; j < b; ++j) orig[j] = j;
    std::vector<double> negc(b);
    for (int j = 0; j < b; ++j) negc[j] = -adj.eig[j];
    mnegc.upload(negc.data(), b, s);
    std::vector<int> its(b, 0);
    const int check = std::max(1, d.check_every);
    std::vector<int> flags(b);
    std::vector<double> phibar(b);
    bool f32op = false;
    // one MINRES step with a stored-H cell product: 5 launches, all arguments
    // device-resident (scalars in mst, done flags per column), so a run of
    // steps can be replayed as a graph.  The two buffer swaps return to the
    // start after two steps.
    // N*bc streams per step: 1 (cell product) + 3 (gather) + 3 (Lanczos) + 6
    // (update).  r2 = r_k (newest Lanczos residual), r1 = r_{k-1}; the step's
    // Lanczos vector v = cv D r2 is formed on the fly (not stored).
    // Fused step: 4 launches and N*bc streams 1 (cell product) + 3 (gather:
    // r2, r1 in, r1 out) + 5 (the previous step's update inside the gather:
    // w, w1 and x in, w and x out; r1 is read anyway).  The last step's
    // update of a run carries over into the next run's first gather; only
    // before a compaction and at the end does finish_run() apply it on its own
    // (6 streams), since those read x.
    bool pending = false;  // a step whose update has not run yet
    auto fused_step = [&] {
      const double* cv = mst.p + (size_t)MS_CV * bc;
      if (f32op) {
        st.applies_fp32++;
        st.apply_cols_fp32 += bc;
        launch_cell_dot<float, double>(g, op32(), S32.p, r2, bc, Djac.p, cv, Yc32.p, mcell.p,
                                       shape32, s);
      } else {
        st.applies_fp64++;
        st.apply_cols_fp64 += bc;
        launch_cell_dot<double, double>(g, op64(), S.p, r2, bc, Djac.p, cv, Yc64.p, mcell.p,
                                        shape64, s);
      }
      launch_minres_alpha_cell(bc, (int)t.ncell, mcell.p, mst.p, s);
      // the previous step's update (w <- its new direction, x) if pending:
      // after that step's swaps w holds its oldest direction and w1 the newer
      double* wu = pending ? w : nullptr;
      const double* w1u = pending ? w1 : nullptr;
      double* xu = pending ? xc : nullptr;
      const int nbk =
          f32op ? launch_minres_gather<float>(g, Yc32.p, S.p, Djac.p, r2, r1, bc, mst.p, partial.p,
                                              shape32.gather_rows, s, wu, w1u, xu)
                : launch_minres_gather<double>(g, Yc64.p, S.p, Djac.p, r2, r1, bc, mst.p,
                                               partial.p, shape64.gather_rows, s, wu, w1u, xu);
      launch_minres_givens(bc, nbk, partial.p, mnegc.p, mst.p, s, 3);
      pending = true;
      std::swap(r1, r2);
      std::swap(w1, w);
    };
    // the last step's update of a run of fused steps (r1 = its Lanczos
    // residual after the swap, w its pending direction, w1 the newer one)
    auto finish_run = [&] {
      if (!pending) return;
      launch_minres_update<double>(N, bc, r1, Djac.p, w, w1, xc, mst.p, s);
      pending = false;
    };
    auto step = [&] {
      if (fused) {
        fused_step();
        return;
      }
      // y <- H~ v - eps v + cr1 r1, v = cv D r2;  partial dots of v.y  (one apply)
      ApplyEpilogue<double> ep;
      ep.Y = W.y.p; ep.alpha = 1.0;
      ep.X = r2; ep.a = 1.0; ep.ca = mst.p + (size_t)MS_CVN * bc;
      ep.Z = r1; ep.b = 1.0; ep.cb = mst.p + (size_t)MS_CR1 * bc;
      ep.xn = Djac.p;
      ep.xcol = mst.p + (size_t)MS_CV * bc;
      ep.dot_partial = partial.p;
      int nbk = f32op ? apply_f32op(r2, bc, ep) : apply<double>(r2, bc, ep);
      launch_minres_alpha(bc, nbk, partial.p, mst.p, s);
      nbk = launch_minres_lanczos<double>(N, bc, W.y.p, r2, r1, Djac.p, mst.p, partial.p, s);
      launch_minres_givens(bc, nbk, partial.p, mnegc.p, mst.p, s);
      launch_minres_update<double>(N, bc, r2, Djac.p, w1, w, xc, mst.p, s);
      std::swap(r1, r2);
      std::swap(w1, w);
    };
    // check_every steps between host convergence checks, captured once per
    // compact block and operator precision as a HIP graph and replayed
    // (kernels only: no memset or copy nodes); the apply scratch is sized
    // before capture.
    const bool useGraph = d.minres_graph != 0 && check % 2 == 0;
    Yc64.resize(cell_scratch_elems(g, b));
    if (canRelax) Yc32.resize(cell_scratch_elems(g, b));
    // graphs per (operator precision, update pending at the run's start)
    hipGraphExec_t execs[4] = {nullptr, nullptr, nullptr, nullptr};
    struct ExecGuard {
      hipGraphExec_t* e;
      void drop() {
        for (int i = 0; i < 4; ++i)
          if (e[i]) { (void)hipGraphExecDestroy(e[i]); e[i] = nullptr; }
      }
      ~ExecGuard() { drop(); }
    } guard{execs};
    std::vector<double> hst;
    // the one buffer not holding a block vector (the buffers stay owned by
    // their DevBufs; the block vectors are views that move between them)
    double* spare = mtmp.p;
    // Repack the compact block to the columns `keep` (positions in it): the
    // caller's x receives every current column first, then the vectors, the
    // per-column state and the shifts are gathered.
    auto compact = [&](const std::vector<int>& keep) {
      const int bn = (int)keep.size();
      msel.upload(orig.data(), bc, s);
      launch_scatter_cols(N, bc, xc, b, msel.p, x, s);
      hst.resize((size_t)MS_COUNT * bc);
      IDFT_HIP(hipMemcpyAsync(hst.data(), mst.p, hst.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      sync();
      for (int k = 0; k < bc; ++k) its[orig[k]] = (int)hst[(size_t)MS_IT * bc + k];
      if (bn == 0) { bc = 0; return; }
      msel.upload(keep.data(), bn, s);
      for (double** buf : {&r1, &r2, &w1, &w, &xc}) {
        launch_gather_cols(N, bc, *buf, bn, msel.p, spare, s);
        std::swap(*buf, spare);
      }
      std::vector<double> nst((size_t)MS_COUNT * bn);
      std::vector<int> norig(bn);
      std::vector<double> nneg(bn);
      for (int i = 0; i < bn; ++i) {
        for (int k = 0; k < MS_COUNT; ++k) nst[(size_t)k * bn + i] = hst[(size_t)k * bc + keep[i]];
        norig[i] = orig[keep[i]];
        nneg[i] = negc[keep[i]];
      }
      mst.upload(nst.data(), nst.size(), s);
      mnegc.upload(nneg.data(), bn, s);
      orig = norig;
      negc = nneg;
      bc = bn;
      guard.drop();
      sync();
    };
    int k = 0;
    for (;;) {
      launch_copy_done(bc, mst.p, mflags.p, s);
      IDFT_HIP(hipMemcpyAsync(flags.data(), mflags.p, bc * sizeof(int), hipMemcpyDeviceToHost, s));
      if (canRelax && !f32op)
        IDFT_HIP(hipMemcpyAsync(phibar.data(), mst.p + (size_t)MS_PHIBAR * bc, bc * sizeof(double),
                                hipMemcpyDeviceToHost, s));
      sync();
      std::vector<int> open;
      for (int j = 0; j < bc; ++j)
        if (!flags[j]) open.push_back(j);
      if (open.empty() || k >= maxit) {
        finish_run();
        break;
      }
      // repack once enough columns have finished, keeping the width even
      // (one finished column rides along, inert: its coefficients are zero)
      // so that the block kernels move two columns per 16-byte access
      std::vector<int> keep = open;
      if (keep.size() % 2 == 1 && (int)keep.size() < bc) {
        for (int j = 0; j < bc; ++j)
          if (flags[j]) { keep.push_back(j); break; }
        std::sort(keep.begin(), keep.end());
      }
      if ((int)keep.size() < bc && bc - (int)open.size() >= std::max(1, bc / 16)) {
        std::vector<int> kflags(keep.size());
        std::vector<double> pb(keep.size());
        for (size_t i = 0; i < keep.size(); ++i) {
          kflags[i] = flags[keep[i]];
          pb[i] = phibar[keep[i]];
        }
        finish_run();  // x complete before it is copied out
        compact(keep);
        st.minres_compactions++;
        if (verbose)
          std::fprintf(stderr, "[idft] minres: step %d, block compacted to %d columns\n", k, bc);
        // the compacted flags and estimates are the kept columns' values
        phibar = pb;
        flags = kflags;
      }
      if (canRelax && !f32op) {
        bool ready = true;
        for (int j = 0; j < bc; ++j)
          if (!flags[j] && !(phibar[j] <= relax * tol[orig[j]])) ready = false;
        if (ready) {
          f32op = true;
          if (verbose)
            std::fprintf(stderr, "[idft] minres: FP32 cell operator from step %d\n", k);
        }
      }
      const int n = std::min(check, maxit - k);
      st.minres_column_steps += (int64_t)n * bc;
      // a run of fused steps that starts without a pending update (the
      // first after the start or a compaction) is launched directly; the
      // runs after it, which all start with one, replay a graph
      if (useGraph && n == check && (pending || !fused)) {
        hipGraphExec_t& exec = execs[(f32op ? 1 : 0) + (pending ? 2 : 0)];
        if (!exec) {
          hipGraph_t graph = nullptr;
          IDFT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          try {
            for (int i = 0; i < n; ++i) step();
          } catch (...) {
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
          }
          IDFT_HIP(hipStreamEndCapture(s, &graph));
          const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          IDFT_HIP(ie);
        } else {
          // the capture counted its n steps; replays count theirs here (and,
          // with fused steps, leave the last update pending as the capture
          // did)
          if (f32op) {
            st.applies_fp32 += n;
            st.apply_cols_fp32 += (int64_t)n * bc;
          } else {
            st.applies_fp64 += n;
            st.apply_cols_fp64 += (int64_t)n * bc;
          }
          if (fused) pending = true;
        }
        IDFT_HIP(hipGraphLaunch(exec, s));
      } else {
        for (int i = 0; i < n; ++i) step();
      }
      k += n;
    }
    // the caller's x and the iteration counts of the columns still in the block
    if (bc > 0) compact({});
    return its;
  }

  void minres(idft_mv* x, double tol, int maxit, int distribute, int32_t* iters) {
    AdjointState& A = adj;
    if (!A.ready) throw Error(IDFT_ESTATE, "no adjoint problem");
    const int b = A.B;
    if (x->ncols != b) throw Error(IDFT_EINVAL, "block size mismatch");
    DevBuf<double>& rhs = mw64.rhs;
    rhs.resize(NB(b));
    compute_rhs(x, rhs.p);
    {
      std::vector<double> nr = coldot(rhs.p, rhs.p, b);
      for (double v : nr)
        if (!std::isfinite(v)) throw Error(IDFT_ENAN, "non-finite adjoint rhs");
    }
    double* xp = x->p();
    launch_lincomb<double>(t.N, b, xp, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    std::vector<int> total(b, 0);
    if (!mixed) {
      total = minres_core(rhs.p, xp, b, std::vector<double>(b, tol), maxit, 0.0);
      if (verbose) {
        const int mx = *std::max_element(total.begin(), total.end());
        std::fprintf(stderr, "[idft] minres fp64: b=%d tol=%.3e max_its=%d\n", b, tol, mx);
      }
    } else {
      // FP64 recurrence with the relaxed (FP64, then FP32) cell operator.  The
      // true residual r = b - (H~ - eps) x, in FP64, then decides per column
      // whether a further sweep solves for the correction (x += delta).
      rref.resize(NB(b));
      xdelta.resize(NB(b));
      double* r = rref.p;
      int used = 0, sweeps = 0;
      std::vector<double> tl(b, tol);
      while (used < maxit) {
        const bool first = sweeps == 0;
        std::vector<int> it = minres_core(first ? rhs.p : r, first ? xp : xdelta.p, b, tl,
                                          maxit - used, d.relax_ratio);
        if (!first)
          launch_lincomb<double>(t.N, b, xp, 1.0, xdelta.p, 1.0, nullptr, nullptr, 0.0, nullptr,
                                 nullptr, s);
        int mxi = 0;
        for (int j = 0; j < b; ++j) { total[j] += it[j]; mxi = std::max(mxi, it[j]); }
        used += std::max(mxi, 1);
        sweeps++;
        ApplyEpilogue<double> ep;  // r = rhs - (H~ - eps) x
        ep.Y = r; ep.alpha = -1.0;
        ep.X = xp; ep.a = -1.0; ep.ca = adj.dnegEig.p;
        ep.Z = rhs.p; ep.b = 1.0;
        apply<double>(xp, b, ep);
        std::vector<double> rn = coldot(r, r, b, Djac.p);
        bool all = true;
        double rmax = 0.0;
        for (int j = 0; j < b; ++j) {
          const double nrm = std::sqrt(std::max(rn[j], 0.0));
          if (!std::isfinite(nrm)) throw Error(IDFT_ENAN, "non-finite refinement residual");
          rmax = std::max(rmax, nrm);
          if (nrm <= tol) tl[j] = INFINITY;  // converged column: starts as done
          else { tl[j] = tol; all = false; }
        }
        if (verbose)
          std::fprintf(stderr, "[idft] minres mixed: sweep %d steps %d true max |r|_D %.3e tol %.3e\n",
                       sweeps, used, rmax, tol);
        if (all) break;
        launch_lincomb<double>(t.N, b, xdelta.p, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr,
                               nullptr, s);
      }
      st.refinements += sweeps - 1;
    }
    int mx = 0;
    int64_t tot = 0;
    for (int j = 0; j < b; ++j) {
      mx = std::max(mx, total[j]);
      tot += total[j];
      if (iters) iters[j] = total[j];
    }
    st.minres_iters_max = std::max(st.minres_iters_max, mx);
    st.minres_iters_total += tot;
    if (distribute) distribute_x();
    sync();
  }

  // ------------------------------------- set-up stage solvers (idft_setup.h)
  // One column: y = Sout (A (Sin x)) with the sum-factorised cell operator o
  // (quadrature weights o.W, kinetic scale o.kin_scale).
  void cell_apply1(const CellOp<double>& o, const double* Sin, const double* Sout, const double* x,
                   double* y) {
    GemmShape sh;
    sh.kernel = CK_SUM_FACTOR;
    sh.vectors_per_tile = 2;
    sh.cells_per_wg = 1;
    Yc64.resize(cell_scratch_elems(g, 1));
    ApplyEpilogue<double> ep;
    ep.Y = y;
    ep.alpha = 1.0;
    launch_cell_apply<double, double>(g, o, Sin, Sout, x, Yc64.p, 1, ep, sh, s);
  }
  double dot1(const double* a, const double* b, const double* w = nullptr) {
    return coldot(a, b, 1, w)[0];
  }
  // Preconditioned CG from x = 0 on one column: A(in, out) applies the
  // operator, D is the diagonal preconditioner; stops at ||r||_D <= tol
  // ||b||_D.  Returns the iterations; relres gets the final ratio.
  template <typename Op>
  int pcg1(Op&& A, const double* b, const double* D, double tol, int maxit, double* x,
           double* relres) {
    const int64_t N = t.N;
    DevBuf<double> r(N), z(N), p(N), q(N);
    launch_lincomb<double>(N, 1, x, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    launch_lincomb<double>(N, 1, r.p, 0.0, b, 1.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    launch_lincomb<double>(N, 1, z.p, 0.0, r.p, 1.0, nullptr, nullptr, 0.0, nullptr, D, s);
    launch_lincomb<double>(N, 1, p.p, 0.0, z.p, 1.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    double rz = dot1(r.p, z.p);
    const double bn = std::sqrt(std::max(rz, 0.0));
    double rn = bn;
    int it = 0;
    while (it < maxit && rn > tol * bn && bn > 0.0) {
      A(p.p, q.p);
      const double alpha = rz / dot1(p.p, q.p);
      launch_lincomb<double>(N, 1, x, 1.0, p.p, alpha, nullptr, nullptr, 0.0, nullptr, nullptr, s);
      launch_lincomb<double>(N, 1, r.p, 1.0, q.p, -alpha, nullptr, nullptr, 0.0, nullptr, nullptr, s);
      launch_lincomb<double>(N, 1, z.p, 0.0, r.p, 1.0, nullptr, nullptr, 0.0, nullptr, D, s);
      const double rz2 = dot1(r.p, z.p);
      rn = std::sqrt(std::max(rz2, 0.0));
      const double beta = rz2 / rz;
      rz = rz2;
      launch_lincomb<double>(N, 1, p.p, beta, z.p, 1.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
      ++it;
    }
    if (relres) *relres = bn > 0.0 ? rn / bn : 0.0;
    return it;
  }
  // f[node] = sum over cells and q of N_a(q) w[e][q] (w already holds JxW):
  // the quadrature operator N^T diag(w) N applied to the all-ones field
  // (N 1 = 1 by the partition of unity)
  void integrate_parent(const double* w, double* f) {
    DevBuf<double> ones(t.N);
    fill_ones(ones.p, t.N);
    launch_lincomb<double>(t.N, 1, f, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    for (int colour = 0; colour < 8; ++colour) launch_quad_operator(g, dt, ones.p, 1, w, f, colour, s);
  }
  void fill_ones(double* v, int64_t n) {
    std::vector<double> h((size_t)n, 1.0);
    IDFT_HIP(hipMemcpyAsync(v, h.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    sync();
  }
  // Poisson: -lap phi = 4 pi c with phi = g on the boundary, as
  //   (1/2 K~) phi~ = S (f/2 - (1/2 K) g),  phi = S phi~ + g,
  // K~ = S K S with S = M^{-1/2} on free nodes (0 on the boundary), CG with
  // the Jacobi diagonal D = 2 M / K_ii.
  void poisson(const double* cq_host, const double* bc_host, double tol, int maxit,
               double* phi_q_host, double* phi_nodal_host, idft_poisson_stats* pst) {
    const int64_t N = t.N;
    // charge moments about the origin and the boundary values
    std::vector<double> gx, gq(nQ);
    double Q = 0.0, dip[3] = {0, 0, 0};
    {
      std::vector<double> xyz((size_t)nQ * 3);
      idft_geometry_quad_points(&d, xyz.data());
      std::vector<double> jw(t.nq);
      const double hh = 0.5 * t.h, J = hh * hh * hh;
      for (int q = 0; q < t.nq; ++q)
        jw[q] = J * t.gq_w[q % t.q1] * t.gq_w[(q / t.q1) % t.q1] * t.gq_w[q / (t.q1 * t.q1)];
      for (int64_t i = 0; i < nQ; ++i) {
        const double w = cq_host[i] * jw[i % t.nq];
        Q += w;
        for (int k = 0; k < 3; ++k) dip[k] += w * xyz[3 * i + k];
      }
    }
    std::vector<double> gb((size_t)N, 0.0), bm((size_t)N, 0.0);
    {
      std::vector<double> nx((size_t)N * 3);
      idft_geometry_node_points(&d, nx.data());
      for (int64_t n = 0; n < N; ++n) {
        const int64_t i = n % t.N1, j = (n / t.N1) % t.N1, k = n / ((int64_t)t.N1 * t.N1);
        if (!(i == 0 || j == 0 || k == 0 || i == t.N1 - 1 || j == t.N1 - 1 || k == t.N1 - 1)) continue;
        bm[n] = 1.0;
        if (bc_host) { gb[n] = bc_host[n]; continue; }
        const double* x = &nx[3 * n];
        const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        gb[n] = Q / r + (dip[0] * x[0] + dip[1] * x[1] + dip[2] * x[2]) / (r * r * r);
      }
    }
    DevBuf<double> gd, bmask, w4, f, kg, b(N), xt(N), zeros(nQ);
    gd.upload(gb.data(), N, s);
    bmask.upload(bm.data(), N, s);
    std::vector<double> c4((size_t)nQ);
    for (int64_t i = 0; i < nQ; ++i) c4[i] = 4.0 * M_PI * cq_host[i];
    DevBuf<double> c4d;
    c4d.upload(c4.data(), nQ, s);
    w4.resize(nQ);
    launch_mul_q(nQ, t.nq, c4d.p, jxw.p, w4.p, s);
    f.resize(N);
    integrate_parent(w4.p, f.p);
    launch_lincomb<double>(nQ, 1, zeros.p, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr,
                           nullptr, s);
    CellOp<double> kin;
    kin.W = zeros.p;
    kin.kin_scale = 0.25 * t.h;
    kg.resize(N);
    cell_apply1(kin, bmask.p, S.p, gd.p, kg.p);  // S (1/2 K) g
    launch_lincomb<double>(N, 1, b.p, 0.0, f.p, 0.5, nullptr, nullptr, 0.0, nullptr, S.p, s);
    launch_lincomb<double>(N, 1, b.p, 1.0, kg.p, -1.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    double rr = 0.0;
    const int it = pcg1([&](const double* in, double* out) { cell_apply1(kin, S.p, S.p, in, out); },
                        b.p, Djac.p, tol, maxit, xt.p, &rr);
    DevBuf<double> phi(N);
    // S is zero on the boundary, so phi = S phi~ there is g alone; in the
    // interior the boundary term g is zero
    launch_lincomb<double>(N, 1, phi.p, 0.0, xt.p, 1.0, nullptr, nullptr, 0.0, nullptr, S.p, s);
    launch_lincomb<double>(N, 1, phi.p, 1.0, gd.p, 1.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    if (phi_nodal_host) phi.download(phi_nodal_host, N, s);
    if (phi_q_host) {
      DevBuf<double> pq(nQ);
      launch_nodal_to_quad(g, dt.Lq.p, phi.p, pq.p, s);
      pq.download(phi_q_host, nQ, s);
    }
    sync();
    if (pst) {
      pst->iterations = it;
      pst->rel_residual = rr;
      pst->charge = Q;
      for (int k = 0; k < 3; ++k) pst->dipole[k] = dip[k];
    }
    if (it >= maxit && rr > tol) throw Error(IDFT_ENOCONV, "Poisson CG did not converge");
  }
  // L2 projection of quadrature values v onto the parent nodes with zero
  // boundary: (S M S) x~ = S (int N v), x = S x~, M the consistent mass at the
  // parent quadrature (the cell operator with weights JxW, no kinetic part).
  void l2_project_parent(const double* vq_dev, double* x_dev) {
    const int64_t N = t.N;
    DevBuf<double> w(nQ), f(N), b(N), xt(N), ones(nQ), jw(nQ);
    launch_mul_q(nQ, t.nq, vq_dev, jxw.p, w.p, s);
    integrate_parent(w.p, f.p);
    launch_lincomb<double>(N, 1, b.p, 0.0, f.p, 1.0, nullptr, nullptr, 0.0, nullptr, S.p, s);
    fill_ones(ones.p, nQ);
    launch_mul_q(nQ, t.nq, ones.p, jxw.p, jw.p, s);
    CellOp<double> mass;
    mass.W = jw.p;
    mass.kin_scale = 0.0;
    DevBuf<double> one(N);
    fill_ones(one.p, N);
    double rr = 0.0;
    const int it = pcg1([&](const double* in, double* out) { cell_apply1(mass, S.p, S.p, in, out); },
                        b.p, one.p, 1e-14, 500, xt.p, &rr);
    if (it >= 500 && rr > 1e-14) throw Error(IDFT_ENOCONV, "L2 projection CG did not converge");
    launch_lincomb<double>(N, 1, x_dev, 0.0, xt.p, 1.0, nullptr, nullptr, 0.0, nullptr, S.p, s);
  }
  // Parent nodal field -> child Q1 nodes: values at the child Gauss points
  // times `scale`, integrated against the child basis, child consistent mass
  // (the tensor product of 1D Q1 masses) inverted by line solves.
  void parent_to_child(const double* x_dev, double scale, double* out_dev) {
    DevBuf<double> ones(t.N);
    fill_ones(ones.p, t.N);
    launch_lincomb<double>(t.NV, 1, out_dev, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr,
                           nullptr, s);
    const double jc = scale * std::pow(0.25 * t.h, 3);
    for (int colour = 0; colour < 8; ++colour)
      launch_force(g, dt, x_dev, ones.p, 1, 0, jc, jxw.p, out_dev, colour, s);
    std::vector<double> cp, inv;
    const double hc = 0.5 * t.h;
    child_mass_factors(t.NV1, hc, cp, inv);
    DevBuf<double> cpd, invd;
    cpd.upload(cp.data(), cp.size(), s);
    invd.upload(inv.data(), inv.size(), s);
    launch_child_mass_solve(t.NV1, hc, cpd.p, invd.p, out_dev, s);
    sync();
  }

  // ------------------------------------------- eigenvector checkpointing
  // With desc.psi_checkpoint the FP64 block X lives only while a pass needs
  // it: between the forward and the backward pass (and between evaluations)
  // it is kept as an FP32 copy, and the eigensolver's and the adjoint's work
  // buffers are released, so the peak device memory is the larger of the two
  // passes' working sets instead of their sum.
  void release_eigen_work() {
    X2.release(); HX.release(); Xtmp.release();
    F0.release(); F1.release(); F2.release(); Rf.release();
    haveHX = false;
  }
  void release_adjoint_work() {
    mw64.r1.release(); mw64.r2.release(); mw64.y.release(); mw64.w1.release();
    mw64.w.release(); mw64.rhs.release();
    mx.release(); mtmp.release(); rref.release(); xdelta.release(); T2buf.release();
    psiAll.release(); mvPsi.own.release(); mvX.own.release();
    Yc64.release(); Yc32.release();
  }
  void checkpoint_x() {
    Xck.resize(NB(B));
    launch_convert_d2f((int64_t)NB(B), Xck.p, X.p, s);
    sync();
    X.release();
    release_eigen_work();
    checkpointed = true;
  }
  void restore_x() {
    if (!checkpointed) return;
    X.resize(NB(B));
    launch_convert_f2d((int64_t)NB(B), X.p, Xck.p, s);
    checkpointed = false;
  }
  // Runs f with X present, restored from the checkpoint for the call only.
  template <typename F>
  void with_x(F&& f) {
    const bool was = checkpointed;
    restore_x();
    f();
    if (was) {
      sync();
      X.release();
      checkpointed = true;
    }
  }
  // The FP64 block of the last forward pass, recomputed from its FP32
  // checkpoint by one ChFSI pass (filter, orthonormalisation, Rayleigh-Ritz)
  // on the same Hamiltonian: the filter takes the rounding of the copy out
  // again.  The occupancies of the forward pass are kept.
  void recompute_x() {
    restore_x();
    haveHX = false;
    chfsi_pass();
    st.psi_recomputes++;
    release_eigen_work();
  }

  // ------------------------------------------------------ getForceVector
  void forward(const double* pot, double* lossOut) {
    const double t0 = wall_ms();
    if (counter > 3) {
      const double prev = tolCheb;
      tolCheb = std::min(d.cheb_tol, lossPrev / d.adaptive_factor_cheb);
      tolCheb = std::min(tolCheb, prev);
    } else {
      tolCheb = std::min(d.cheb_tol, d.initial_tol_cheb);
    }
    if (vxcFreq > 0) potEval.assign(pot, pot + t.NV);
    set_potential(pot);
    restore_x();  // warm start from the checkpoint of the previous evaluation
    solve_passes(tolCheb, d.max_cheb_passes, nullptr, nullptr);
    st.cheb_tol_used = tolCheb;
    sync();
    const double t1 = wall_ms();
    rho.resize(nQ); u.resize(nQ);
    density_into(rho.p);
    partial.resize((size_t)kLossBlocks * 4);
    small.resize(16);
    launch_loss(nQ, t.nq, rho.p, rhoT.p, weight.p, veff.p, jxw.p, u.p, partial.p, s);
    launch_sum_partials(kLossBlocks, 4, partial.p, small.p, s);
    double v4[4];
    small.download(v4, 4, s);
    sync();
    loss = v4[0];
    st.loss_unweighted = v4[1];
    st.l1_density_error = v4[2];
    st.int_rho_veff = v4[3];
    st.t_eigen_ms = t1 - t0;
    st.t_density_ms = wall_ms() - t1;
    if (lossOut) *lossOut = loss;
    haveForward = true;
    if (d.psi_checkpoint) checkpoint_x();
  }

  void backward(double* forceOut) {
    if (!haveForward) throw Error(IDFT_ESTATE, "backward before forward");
    const double t0 = wall_ms();
    if (checkpointed) recompute_x();
    physical_psi();
    force.resize(t.NV);
    launch_lincomb<double>(t.NV, 1, force.p, 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr, nullptr, s);
    st.minres_iters_max = 0;
    st.minres_iters_total = 0;
    st.refinements = 0;
    double tForce = 0.0;
    const int defBS = d.block_size > 0 ? d.block_size : B;
    int jvec = 0;
    st.adjoint_blocks = 0;
    while (jvec < B) {
      st.adjoint_blocks++;
      int bs = std::min(defBS, B - jvec);
      while (true) {
        const int last = jvec + bs - 1;
        if (last + 1 != B && std::fabs(eig[last] - eig[last + 1]) < d.degeneracy_tol && bs > 1) bs--;
        else break;
      }
      mvPsi.ncols = bs; mvPsi.ctx = this;
      mvPsi.own.resize(NB(bs));
      launch_copy_cols(t.N, B, jvec, bs, psiAll.p, mvPsi.p(), s);
      std::vector<double> occb(bs), shifts(bs);
      for (int j = 0; j < bs; ++j) { shifts[j] = eig[jvec + j]; occb[j] = occ[jvec + j]; }
      // degeneracy map (InverseDFTSolverFunction.cpp:37-77)
      std::map<int, std::set<int>> groups;
      std::map<int, int> toGroup;
      int count = 0;
      for (int i = 0; i < bs; ++i) {
        auto it = toGroup.find(i);
        int gi;
        if (it != toGroup.end()) gi = it->second;
        else { gi = count++; groups[gi].insert(i); toGroup[i] = gi; }
        for (int j = 0; j < bs; ++j)
          if (std::fabs(shifts[i] - shifts[j]) < d.degeneracy_tol) { groups[gi].insert(j); toGroup[j] = gi; }
      }
      std::vector<int32_t> pairs;
      for (int i = 0; i < bs; ++i)
        for (int j : groups[toGroup[i]]) { pairs.push_back(i); pairs.push_back(j); }
      update_input_psi(&mvPsi, occb.data(), u.p, nullptr, pairs.data(), (int)pairs.size() / 2,
                       fermi, shifts.data(), bs);
      adjointTol = std::min(adjointTol, tolCheb / d.adaptive_factor_adjoint);
      st.adjoint_tol_used = adjointTol;
      mvX.ncols = bs; mvX.ctx = this;
      mvX.own.resize(NB(bs));
      minres(&mvX, adjointTol, d.adjoint_max_iter, 1, nullptr);
      launch_zero_rows(t.N, bs, mvX.p(), adjMask.p, s);
      const double tf = wall_ms();
      const int mode = d.gradient_mode == IDFT_GRADIENT_REFERENCE ? 0 : 1;
      const double jc = std::pow(0.25 * t.h, 3);
      for (int colour = 0; colour < 8; ++colour)
        launch_force(g, dt, mvPsi.p(), mvX.p(), bs, mode, jc, jxw.p, force.p, colour, s);
      sync();
      tForce += wall_ms() - tf;
      jvec += bs;
    }
    force.download(forceOut, t.NV, s);
    sync();
    st.t_adjoint_ms = wall_ms() - t0 - tForce;
    st.t_force_ms = tForce;
    lossPrev = loss;
    // ISF:1250-1254: every vxcFreq-th evaluation writes its input potential
    if (vxcFreq > 0 && counter % vxcFreq == 0 && (int64_t)potEval.size() == t.NV) {
      const std::string file = vxcFolder + "/" + vxcPostfix + "_" + std::to_string(counter);
      if (idft_write_vxc_text(&d, file.c_str(), potEval.data(), 1) != IDFT_OK)
        throw Error(IDFT_EINVAL, "cannot write v_xc file " + file);
    }
    counter++;
    st.n_evaluations = counter;
    haveForward = false;
    if (d.psi_checkpoint) {
      release_adjoint_work();
      checkpoint_x();
    }
  }
};

const BfgsCheckpoint& idft::bfgs_checkpoint_of(const idft_ctx* c) { return c->ckpt; }

// ======================================================================= ABI
namespace {

template <typename F>
int guarded(idft_ctx* c, F&& f) {
  try {
    f();
    return IDFT_OK;
  } catch (const Error& e) {
    if (c) c->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    if (c) c->err = e.what();
    return IDFT_EINVAL;
  }
}

idft_ctx* create_ctx(const idft_desc* d, const double* pot_base_q, const double* rho_target_q,
                     const double* weight_q, const double* rho_target_nodal) {
  idft_ctx* c = new idft_ctx();
  try {
    c->d = *d;
    c->t.init(*d);
    const Tables& t = c->t;
    c->g = Geo{t.nc, t.p, t.n1, t.N1, t.q1, t.NV1, t.N, t.ncell};
    c->B = d->n_eig;
    c->nQ = t.ncell * (int64_t)t.nq;
    c->mixed = d->precision == IDFT_MIXED;
    c->verbose = std::getenv("IDFT_VERBOSE") != nullptr;
    IDFT_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    hipStream_t s = c->s;
    c->dt.Lq.upload(t.Lq.data(), t.Lq.size(), s);
    c->dt.Lc.upload(t.Lc.data(), t.Lc.size(), s);
    c->dt.M1.upload(t.M1.data(), t.M1.size(), s);
    c->dt.D1.upload(t.D1.data(), t.D1.size(), s);
    c->dt.Wq.upload(t.Wq.data(), t.Wq.size(), s);
    c->dt.Wc.upload(t.Wc.data(), t.Wc.size(), s);
    c->check_sf_tables();
    upload_sf_constants();
    // default cell product: sum-factorised (no stored H_e, 4 columns per
    // work-group), the faster kernel at p = 4 and 5 for 12 to 80 columns,
    // if its cell scratch fits 32-bit indexing; the row pass over stored
    // tiles otherwise (idft_autotune_apply refines both)
    if (cell_scratch_elems(c->g, d->n_eig) < (int64_t)0xffffffffLL) {
      for (GemmShape* sh : {&c->shape64, &c->shape32}) {
        sh->kernel = CK_SUM_FACTOR;
        sh->vectors_per_tile = 4;
        sh->cells_per_wg = 1;
      }
    }
    // uniform-cell JxW at the parent Gauss points
    const double hh = 0.5 * t.h, J = hh * hh * hh;
    std::vector<double> jw(t.nq);
    for (int qz = 0; qz < t.q1; ++qz)
      for (int qy = 0; qy < t.q1; ++qy)
        for (int qx = 0; qx < t.q1; ++qx)
          jw[(qz * t.q1 + qy) * t.q1 + qx] = J * t.gq_w[qx] * t.gq_w[qy] * t.gq_w[qz];
    c->jxw.upload(jw.data(), jw.size(), s);
    // lumped mass and stiffness diagonal
    std::vector<double> M(t.N, 0.0), K(t.N, 0.0);
    const int n1 = t.n1;
    std::vector<double> Kloc(t.ne);
    for (int cc = 0; cc < n1; ++cc)
      for (int b = 0; b < n1; ++b)
        for (int a = 0; a < n1; ++a)
          Kloc[(cc * n1 + b) * n1 + a] =
              hh * (t.D1[a * n1 + a] * t.M1[b * n1 + b] * t.M1[cc * n1 + cc] +
                    t.M1[a * n1 + a] * t.D1[b * n1 + b] * t.M1[cc * n1 + cc] +
                    t.M1[a * n1 + a] * t.M1[b * n1 + b] * t.D1[cc * n1 + cc]);
    for (int64_t e = 0; e < t.ncell; ++e) {
      const int64_t cx = e % t.nc, cy = (e / t.nc) % t.nc, cz = e / ((int64_t)t.nc * t.nc);
      for (int cc = 0; cc < n1; ++cc)
        for (int b = 0; b < n1; ++b)
          for (int a = 0; a < n1; ++a) {
            const int64_t n = ((cz * t.p + cc) * t.N1 + (cy * t.p + b)) * (int64_t)t.N1 + (cx * t.p + a);
            M[n] += J * t.gll_w[a] * t.gll_w[b] * t.gll_w[cc];
            K[n] += Kloc[(cc * n1 + b) * n1 + a];
          }
    }
    std::vector<double> Sv(t.N), sq(t.N), isq(t.N), Dj(t.N), Djs(t.N);
    std::vector<float> S32(t.N);
    std::vector<unsigned char> bnd(t.N), am(t.N);
    for (int64_t n = 0; n < t.N; ++n) {
      const int64_t i = n % t.N1, j = (n / t.N1) % t.N1, k = n / ((int64_t)t.N1 * t.N1);
      bnd[n] = (i == 0 || j == 0 || k == 0 || i == t.N1 - 1 || j == t.N1 - 1 || k == t.N1 - 1);
      am[n] = rho_target_nodal[n] < d->rho_tol_constraints;
      sq[n] = std::sqrt(M[n]);
      isq[n] = 1.0 / sq[n];
      Sv[n] = bnd[n] ? 0.0 : isq[n];
      S32[n] = (float)Sv[n];
      Dj[n] = 2.0 * M[n] / K[n];
      Djs[n] = std::sqrt(Dj[n]);
    }
    c->hMdiag = M;
    c->Mdiag.upload(M.data(), t.N, s);
    c->S.upload(Sv.data(), t.N, s);
    c->S32.upload(S32.data(), t.N, s);
    c->sqrtM.upload(sq.data(), t.N, s);
    c->invSqrtM.upload(isq.data(), t.N, s);
    c->Djac.upload(Dj.data(), t.N, s);
    c->DjacSqrt.upload(Djs.data(), t.N, s);
    c->bnd.upload(bnd.data(), t.N, s);
    c->adjMask.upload(am.data(), t.N, s);
    c->potBase.upload(pot_base_q, c->nQ, s);
    c->rhoT.upload(rho_target_q, c->nQ, s);
    c->weight.upload(weight_q, c->nQ, s);
    c->adjointTol = d->adjoint_initial_tol;
    c->pot0.assign(t.NV, 0.0);
    c->potSol.assign(t.NV, 0.0);
    c->sync();
  } catch (...) {
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
    throw;
  }
  return c;
}

}  // namespace

extern "C" {

const char* idft_backend(void) { return "hip-gfx950"; }

int idft_hip_init(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return IDFT_EHIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return IDFT_EHIP;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return IDFT_EINVAL;
  // device address of the allocation-free entry point's fixed scratch
  try {
    (void)noalloc_scratch();
  } catch (const Error& e) {
    return e.code;
  }
  return IDFT_OK;
}

int idft_create(const idft_desc* d, const double* pot_base_q, const double* rho_target_q,
                const double* weight_q, const double* rho_target_nodal, idft_ctx** out) {
  if (!out) return IDFT_EINVAL;
  *out = nullptr;
  if (!desc_valid(d) || !pot_base_q || !rho_target_q || !weight_q || !rho_target_nodal)
    return IDFT_EINVAL;
  if (d->precision != IDFT_FP64 && d->precision != IDFT_MIXED) return IDFT_EINVAL;
  try {
    *out = create_ctx(d, pot_base_q, rho_target_q, weight_q, rho_target_nodal);
  } catch (const Error& e) {
    return e.code;
  } catch (...) {
    return IDFT_EHIP;
  }
  return IDFT_OK;
}

void idft_destroy(idft_ctx* c) {
  if (!c) return;
  if (c->s) (void)hipStreamSynchronize(c->s);
  if (c->blas) (void)rocblas_destroy_handle(c->blas);
  hipStream_t s = c->s;
  delete c;
  if (s) (void)hipStreamDestroy(s);
}

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
  return guarded(c, [&] { c->set_potential(pot); });
}

int idft_set_veff_quad(idft_ctx* c, const double* vq) {
  if (!c || !vq) return IDFT_EINVAL;
  return guarded(c, [&] {
    c->veff.upload(vq, c->nQ, c->s);
    c->build_h();
    c->sync();
  });
}

int idft_get_eigenvalues(const idft_ctx* c, double* ev) {
  if (!c || !ev) return IDFT_EINVAL;
  if (!c->haveRitz || c->eig.empty()) return IDFT_ESTATE;
  std::copy(c->eig.begin(), c->eig.end(), ev);
  return IDFT_OK;
}

int idft_get_veff_quad(const idft_ctx* cc, double* vq) {
  idft_ctx* c = const_cast<idft_ctx*>(cc);
  if (!c || !vq) return IDFT_EINVAL;
  if (!c->haveH) return IDFT_ESTATE;
  return guarded(c, [&] {
    c->veff.download(vq, c->nQ, c->s);
    c->sync();
  });
}

int idft_eigensolve(idft_ctx* c, double tol, int32_t maxp, double* ev, double* mr, int32_t* passes) {
  if (!c) return IDFT_EINVAL;
  return guarded(c, [&] {
    int p = 0;
    c->restore_x();
    c->solve_passes(tol, maxp, &p, mr);
    if (passes) *passes = p;
    if (ev) std::copy(c->eig.begin(), c->eig.end(), ev);
  });
}

int idft_get_eigenvectors(const idft_ctx* cc, double* psi) {
  idft_ctx* c = const_cast<idft_ctx*>(cc);
  if (!c || !psi) return IDFT_EINVAL;
  if (!c->haveX) return IDFT_ESTATE;
  return guarded(c, [&] {
    // between passes with psi_checkpoint: the FP32 checkpoint's values
    c->with_x([&] {
      c->physical_psi();
      c->psiAll.download(psi, c->NB(c->B), c->s);
      c->sync();
    });
  });
}

int idft_set_eigenvectors(idft_ctx* c, const double* psi) {
  if (!c || !psi) return IDFT_EINVAL;
  return guarded(c, [&] {
    c->checkpointed = false;
    c->X.upload(psi, c->NB(c->B), c->s);
    launch_lincomb<double>(c->t.N, c->B, c->X.p, 0.0, c->X.p, 1.0, nullptr, nullptr, 0.0, nullptr,
                           c->sqrtM.p, c->s);
    c->zero_boundary(c->X.p, c->B);
    c->sync();
    c->haveX = true;
    c->haveRitz = false;
    c->haveHX = false;
    c->needRR = true;
  });
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
  if (!c->haveX || c->occ.empty()) {
    c->err = "no eigenvectors/occupancies";
    return IDFT_ESTATE;
  }
  return guarded(c, [&] {
    DevBuf<double> r(c->nQ);
    c->with_x([&] { c->density_into(r.p); });
    r.download(rho, c->nQ, c->s);
    c->sync();
  });
}

int idft_apply_hamiltonian(idft_ctx* c, const double* x, double* y, int32_t ncols, int32_t prec) {
  if (!c || !x || !y || ncols < 1 || ncols > kMaxCols) return IDFT_EINVAL;
  if (!c->haveH) { c->err = "no Hamiltonian"; return IDFT_ESTATE; }
  return guarded(c, [&] {
    const size_t n = c->NB(ncols);
    DevBuf<double> xd, yd(n);
    xd.upload(x, n, c->s);
    if (prec == IDFT_MIXED) {
      DevBuf<float> xf(n), yf(n);
      launch_convert_d2f((int64_t)n, xf.p, xd.p, c->s);
      c->happly32(xf.p, yf.p, ncols, 1.0);
      launch_convert_f2d((int64_t)n, yd.p, yf.p, c->s);
    } else {
      c->happly64(xd.p, yd.p, ncols, 1.0);
    }
    yd.download(y, n, c->s);
    c->sync();
  });
}

int idft_chebyshev_filter(idft_ctx* c, double* x, int32_t ncols, int32_t deg, double a, double b,
                          double a0, int32_t prec) {
  if (!c || !x || ncols < 1 || ncols > kMaxCols || deg < 1) return IDFT_EINVAL;
  if (!c->haveH) { c->err = "no Hamiltonian"; return IDFT_ESTATE; }
  return guarded(c, [&] {
    DevBuf<double> xd;
    xd.upload(x, c->NB(ncols), c->s);
    c->chebyshev_filter_X(xd.p, ncols, deg, a, b, a0, prec == IDFT_MIXED);
    xd.download(x, c->NB(ncols), c->s);
    c->sync();
  });
}

int idft_mv_create(idft_ctx* c, int32_t ncols, idft_mv** out) {
  if (!c || !out || ncols < 1 || ncols > kMaxCols) return IDFT_EINVAL;
  return guarded(c, [&] {
    idft_mv* v = new idft_mv();
    v->ctx = c;
    v->ncols = ncols;
    v->own.resize(c->NB(ncols));
    launch_lincomb<double>(c->t.N, ncols, v->p(), 0.0, nullptr, 0.0, nullptr, nullptr, 0.0, nullptr,
                           nullptr, c->s);
    c->sync();
    *out = v;
  });
}

int idft_mv_wrap(idft_ctx* c, double* dptr, int32_t ncols, idft_mv** out) {
  if (!c || !dptr || !out || ncols < 1 || ncols > kMaxCols) return IDFT_EINVAL;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, dptr) != hipSuccess || attr.type != hipMemoryTypeDevice) {
    c->err = "idft_mv_wrap: not a device pointer";
    return IDFT_EINVAL;
  }
  idft_mv* v = new idft_mv();
  v->ctx = c;
  v->ncols = ncols;
  v->ext = dptr;
  *out = v;
  return IDFT_OK;
}

double* idft_mv_device_ptr(const idft_mv* v) { return v ? v->p() : nullptr; }

int idft_fma_probe(int32_t precision, int32_t chains, double* tflops) {
  if (!tflops || (precision != IDFT_FP64 && precision != IDFT_MIXED) ||
      (chains != 4 && chains != 8 && chains != 16))
    return IDFT_EINVAL;
  try {
    *tflops = launch_fma_probe(precision == IDFT_FP64, chains, nullptr);
    IDFT_HIP(hipDeviceSynchronize());
  } catch (const Error& e) {
    return e.code;
  } catch (...) {
    return IDFT_EHIP;
  }
  return IDFT_OK;
}

int idft_device_memory(int64_t* cur, int64_t* peak, int32_t reset) {
  if (cur) *cur = devmem_current().load();
  if (peak) *peak = devmem_peak().load();
  if (reset) devmem_peak().store(devmem_current().load());
  return IDFT_OK;
}

void idft_mv_destroy(idft_mv* v) { delete v; }

int idft_mv_set(idft_ctx* c, idft_mv* v, const double* h) {
  if (!c || !v || !h) return IDFT_EINVAL;
  return guarded(c, [&] {
    IDFT_HIP(hipMemcpyAsync(v->p(), h, c->NB(v->ncols) * sizeof(double), hipMemcpyHostToDevice,
                            c->s));
    c->sync();
  });
}

int idft_mv_get(idft_ctx* c, const idft_mv* v, double* h) {
  if (!c || !v || !h) return IDFT_EINVAL;
  return guarded(c, [&] {
    IDFT_HIP(hipMemcpyAsync(h, v->p(), c->NB(v->ncols) * sizeof(double), hipMemcpyDeviceToHost,
                            c->s));
    c->sync();
  });
}

int idft_adj_update_input_psi(idft_ctx* c, idft_mv* psi, const double* occ, const double* u_q,
                              const int32_t* pairs, int32_t np, double fermi, double* eigv,
                              int32_t block) {
  if (!c || !psi || !occ || !u_q || (!pairs && np) || !eigv || np < 0) return IDFT_EINVAL;
  return guarded(c, [&] {
    c->update_input_psi(psi, occ, nullptr, u_q, pairs, np, fermi, eigv, block);
    c->sync();
  });
}

int idft_adj_compute_rhs(idft_ctx* c, idft_mv* ndbc, idft_mv* x, idft_mv* rhs) {
  if (!c || !x || !rhs) return IDFT_EINVAL;
  (void)ndbc;
  return guarded(c, [&] {
    if (rhs->ncols != c->adj.B) throw Error(IDFT_EINVAL, "block size mismatch");
    c->compute_rhs(x, rhs->p());
  });
}

int idft_adj_vmult(idft_ctx* c, idft_mv* Ax, idft_mv* x, int32_t block) {
  if (!c || !Ax || !x) return IDFT_EINVAL;
  return guarded(c, [&] {
    if (block != c->adj.B || x->ncols != block || Ax->ncols != block)
      throw Error(IDFT_EINVAL, "block size mismatch");
    c->vmult(Ax->p(), x->p(), block);
    c->sync();
  });
}

int idft_adj_vmult_noalloc(idft_ctx* c, idft_mv* Ax, idft_mv* x, int32_t block) {
  if (!c || !Ax || !x) return IDFT_EINVAL;
  return guarded(c, [&] {
    if (block != c->adj.B || x->ncols != block || Ax->ncols != block)
      throw Error(IDFT_EINVAL, "block size mismatch");
    c->vmult_noalloc(Ax->p(), x->p(), block);
    c->sync();
  });
}

int idft_adj_precondition_jacobi(idft_ctx* c, idft_mv* dst, const idft_mv* src, double) {
  if (!c || !dst || !src || dst->ncols != src->ncols) return IDFT_EINVAL;
  return guarded(c, [&] {
    launch_lincomb<double>(c->t.N, src->ncols, dst->p(), 0.0, src->p(), 1.0, nullptr, nullptr, 0.0,
                           nullptr, c->Djac.p, c->s);
    c->sync();
  });
}

int idft_adj_precondition_jacobi_sqrt(idft_ctx* c, idft_mv* dst, const idft_mv* src, double) {
  if (!c || !dst || !src || dst->ncols != src->ncols) return IDFT_EINVAL;
  return guarded(c, [&] {
    launch_lincomb<double>(c->t.N, src->ncols, dst->p(), 0.0, src->p(), 1.0, nullptr, nullptr, 0.0,
                           nullptr, c->DjacSqrt.p, c->s);
    c->sync();
  });
}

int idft_adj_distribute_x(idft_ctx* c) {
  if (!c) return IDFT_EINVAL;
  return guarded(c, [&] { c->distribute_x(); });
}

int idft_adj_dot_quad_wise(idft_ctx* c, idft_mv* v1, idft_mv* v2, double* out) {
  if (!c || !v1 || !v2 || !out) return IDFT_EINVAL;
  return guarded(c, [&] {
    if (!c->adj.ready) throw Error(IDFT_ESTATE, "no adjoint problem");
    std::vector<double> o = c->dot_quad_wise(v1->p(), v2->p(), c->adj.B);
    std::copy(o.begin(), o.end(), out);
  });
}

int idft_minres_solve(idft_ctx* c, idft_mv* x, idft_mv* ndbc, double tol, int32_t maxit,
                      int32_t distribute, int32_t* iters) {
  if (!c || !x || maxit < 0) return IDFT_EINVAL;
  (void)ndbc;
  return guarded(c, [&] {
    c->st.minres_iters_max = 0;
    c->st.minres_iters_total = 0;
    c->st.refinements = 0;
    c->minres(x, tol, maxit, distribute, iters);
  });
}

int idft_profile_apply(idft_ctx* c, int32_t ncols, int32_t prec, int32_t reps, double* ms,
                       double* ms_gemm) {
  if (!c || !ms || ncols < 1 || ncols > kMaxCols || reps < 1) return IDFT_EINVAL;
  if (!c->haveH) { c->err = "no Hamiltonian"; return IDFT_ESTATE; }
  return guarded(c, [&] {
    *ms = c->profile_ms(ncols, prec == IDFT_MIXED, reps, false, ms_gemm);
    c->release_profile_block();
  });
}

int idft_set_vxc_output(idft_ctx* c, const char* folder, const char* postfix, int32_t frequency) {
  if (!c || frequency < 0 || (frequency > 0 && (!folder || !postfix))) return IDFT_EINVAL;
  c->vxcFreq = frequency;
  c->vxcFolder = frequency > 0 ? folder : "";
  c->vxcPostfix = frequency > 0 ? postfix : "";
  return IDFT_OK;
}

int idft_bfgs_set_checkpoint(idft_ctx* c, const char* path, int32_t every, int32_t resume) {
  if (!c || (path && every < 1)) return IDFT_EINVAL;
  c->ckpt.path = path ? path : "";
  c->ckpt.every = path ? every : 1;
  c->ckpt.resume = path && resume != 0;
  return IDFT_OK;
}

int idft_set_apply_shape(idft_ctx* c, int32_t prec, int32_t cells_per_wg,
                         int32_t vectors_per_tile, int32_t kernel, int32_t gather_rows) {
  if (!c || cells_per_wg < 1 || cells_per_wg > 64 || gather_rows < 0 || gather_rows > 4096 ||
      kernel < CK_ROW_PASS || kernel > CK_SUM_FACTOR)
    return IDFT_EINVAL;
  if (kernel == CK_SUM_FACTOR
          ? !sf_vc_supported(vectors_per_tile, cells_per_wg, prec != IDFT_MIXED)
          : (vectors_per_tile < 16 || vectors_per_tile > kMaxCols || vectors_per_tile % 16))
    return IDFT_EINVAL;
  GemmShape& t = prec == IDFT_MIXED ? c->shape32 : c->shape64;
  t.cells_per_wg = cells_per_wg;
  t.vectors_per_tile = vectors_per_tile;
  t.kernel = kernel;
  t.gather_rows = gather_rows;
  return IDFT_OK;
}

int idft_autotune_apply(idft_ctx* c, int32_t ncols, int32_t prec, int32_t reps,
                        int32_t* cells_per_wg, int32_t* vectors_per_tile, int32_t* kernel,
                        int32_t* gather_rows, double* ms_best) {
  if (!c || ncols < 1 || ncols > kMaxCols || reps < 1) return IDFT_EINVAL;
  if (!c->haveH) { c->err = "no Hamiltonian"; return IDFT_ESTATE; }
  const bool f32 = prec == IDFT_MIXED;
  if (f32 && !c->mixed) { c->err = "FP32 Hamiltonian not built"; return IDFT_ESTATE; }
  return guarded(c, [&] {
    GemmShape& target = f32 ? c->shape32 : c->shape64;
    const GemmShape saved = target;
    GemmShape best = saved;
    double bestMs = 1e300;
    const int vmax = ((ncols + 15) / 16) * 16;
    std::vector<int> widths;
    for (int vt : {16, 32, 64, vmax}) {
      const int e = std::min(vt, vmax);
      if (std::find(widths.begin(), widths.end(), e) == widths.end()) widths.push_back(e);
    }
    // score: a plain application (eigensolver) plus one with node and
    // column scales on the input (MINRES steps)
    auto trial = [&] {
      double ms = 0.0;
      try {
        ms = c->profile_ms(ncols, f32, reps, false, nullptr) +
             c->profile_ms(ncols, f32, reps, true, nullptr);
      } catch (...) {
        target = saved;
        throw;
      }
      if (ms < bestMs) { bestMs = ms; best = target; }
    };
    target.gather_rows = 0;
    target.kernel = CK_SUM_FACTOR;
    for (int cg : {2, 4, 8}) {
      if (cg > 2 && cg / 2 >= ncols) continue;  // mostly empty column groups
      // 2 columns per work-group gather each node row of the block once per
      // 2 columns.  Between the solvers' other vector kernels this measured
      // 1.3x slower than 4 columns at B = 12 (FP64 MINRES, kernel trace), a
      // cost that a timing loop of back-to-back applications does not show.
      if (cg == 2 && ncols >= 8) continue;
      target.vectors_per_tile = cg;
      target.cells_per_wg = 1;
      trial();
    }
    const int tpMax = f32 ? 32 : 16;  // widest tile of the tile-pass kernel
    for (int vt : widths) {
      for (int k = CK_ROW_PASS; k <= (vt <= tpMax ? CK_TILE_PASS : CK_ROW_PASS); ++k) {
        for (int cp : {1, 2, 4, 8}) {
          target.cells_per_wg = cp;
          target.vectors_per_tile = vt;
          target.kernel = k;
          trial();
        }
      }
    }
    target = best;
    for (int gr : {1, 2, 4, 8}) {
      target.gather_rows = gr;
      trial();
    }
    target = best;
    c->release_unused_h();
    c->release_profile_block();
    if (cells_per_wg) *cells_per_wg = best.cells_per_wg;
    if (vectors_per_tile) *vectors_per_tile = best.vectors_per_tile;
    if (kernel) *kernel = best.kernel;
    if (gather_rows) *gather_rows = best.gather_rows;
    if (ms_best) *ms_best = bestMs;
  });
}

// ------------------------------------------------ set-up stage (idft_setup.h)
}  // extern "C"
namespace {
// A context that only carries the mesh of d (zero potential and density).
std::unique_ptr<idft_ctx> mesh_ctx(const idft_desc* d) {
  idft_desc m = *d;
  m.n_eig = 2;
  m.n_electrons = 2.0;
  m.precision = IDFT_FP64;
  if (!desc_valid(&m)) throw Error(IDFT_EINVAL, "invalid descriptor");
  Tables t;
  t.init(m);
  const int64_t nQ = t.ncell * (int64_t)t.nq;
  std::vector<double> zq((size_t)nQ, 0.0), oq((size_t)nQ, 1.0), on((size_t)t.N, 1.0);
  return std::unique_ptr<idft_ctx>(create_ctx(&m, zq.data(), zq.data(), oq.data(), on.data()));
}
template <typename F>
int guarded_free(F&& f) {
  try {
    f();
    return IDFT_OK;
  } catch (const Error& e) {
    return e.code;
  } catch (const std::exception&) {
    return IDFT_EINVAL;
  }
}
}  // namespace
extern "C" {

int idft_poisson_solve(const idft_desc* d, const double* charge_q, const double* boundary_nodal,
                       double tol, int32_t max_iter, double* phi_q, double* phi_nodal,
                       idft_poisson_stats* st) {
  if (!d || !charge_q || !(tol > 0.0) || max_iter < 1) return IDFT_EINVAL;
  return guarded_free([&] {
    auto c = mesh_ctx(d);
    c->poisson(charge_q, boundary_nodal, tol, max_iter, phi_q, phi_nodal, st);
  });
}

int idft_l2_project_parent(const idft_desc* d, const double* v_q, double* x_nodal) {
  if (!d || !v_q || !x_nodal) return IDFT_EINVAL;
  return guarded_free([&] {
    auto c = mesh_ctx(d);
    DevBuf<double> vq, x(c->t.N);
    vq.upload(v_q, c->nQ, c->s);
    c->l2_project_parent(vq.p, x.p);
    x.download(x_nodal, c->t.N, c->s);
    c->sync();
  });
}

int idft_parent_to_child(const idft_desc* d, const double* x_nodal, double scale, double* vxc_child) {
  if (!d || !x_nodal || !vxc_child) return IDFT_EINVAL;
  return guarded_free([&] {
    auto c = mesh_ctx(d);
    DevBuf<double> x, out(c->t.NV);
    x.upload(x_nodal, c->t.N, c->s);
    c->parent_to_child(x.p, scale, out.p);
    out.download(vxc_child, c->t.NV, c->s);
    c->sync();
  });
}

int idft_initial_vxc(const idft_desc* d, const idft_guess_params* gp, const double* rho_spin_q,
                     const double* sigma_q, const double* v_hartree_q, double* v_lda_q,
                     double* vxc_child) {
  if (!d || !gp || !rho_spin_q || (gp->lb94 && !sigma_q) ||
      (gp->fermi_amaldi && (!v_hartree_q || !(gp->n_electrons > 0.0))))
    return IDFT_EINVAL;
  return guarded_free([&] {
    auto c = mesh_ctx(d);
    const int64_t nQ = c->nQ;
    DevBuf<double> rho, sig, vh, v(nQ);
    rho.upload(rho_spin_q, nQ, c->s);
    if (gp->lb94) sig.upload(sigma_q, nQ, c->s);
    if (gp->fermi_amaldi) vh.upload(v_hartree_q, nQ, c->s);
    GuessArgs ga{gp->lb94, gp->fermi_amaldi, gp->tau_vx, gp->tau_fa, gp->factor_fa,
                 gp->spin_factor, gp->n_electrons};
    launch_guess(nQ, ga, rho.p, sig.p, vh.p, v.p, c->s);
    if (v_lda_q) v.download(v_lda_q, nQ, c->s);
    if (vxc_child) {
      DevBuf<double> x(c->t.N), out(c->t.NV);
      c->l2_project_parent(v.p, x.p);
      c->parent_to_child(x.p, 1.0 - gp->factor_lda, out.p);
      out.download(vxc_child, c->t.NV, c->s);
    }
    c->sync();
  });
}

int idft_forward(idft_ctx* c, const double* pot, double* loss) {
  if (!c || !pot) return IDFT_EINVAL;
  return guarded(c, [&] { c->forward(pot, loss); });
}

int idft_backward(idft_ctx* c, double* force) {
  if (!c || !force) return IDFT_EINVAL;
  return guarded(c, [&] { c->backward(force); });
}

int idft_get_force(idft_ctx* c, double* pot, double* force, double* loss) {
  if (!c || !pot || !force || !loss) return IDFT_EINVAL;
  return guarded(c, [&] {
    const double t0 = wall_ms();
    c->forward(pot, loss);
    c->backward(force);
    c->st.t_total_ms = wall_ms() - t0;
  });
}

int idft_dot(idft_ctx* c, const double* v1, const double* v2, int64_t n, int32_t block, double* out) {
  if (!c || !v1 || !v2 || !out || block < 1) return IDFT_EINVAL;
  for (int i = 0; i < block; ++i) out[i] = 0.0;
  for (int64_t i = 0; i < n; ++i) out[i % block] += v1[i] * v2[i];
  return IDFT_OK;
}

int idft_set_initial_guess(idft_ctx* c, const double* pot) {
  if (!c || !pot) return IDFT_EINVAL;
  c->pot0.assign(pot, pot + c->t.NV);
  return IDFT_OK;
}
int idft_get_initial_guess(const idft_ctx* c, double* pot) {
  if (!c || !pot) return IDFT_EINVAL;
  std::copy(c->pot0.begin(), c->pot0.end(), pot);
  return IDFT_OK;
}
int idft_set_solution(idft_ctx* c, const double* pot) {
  if (!c || !pot) return IDFT_EINVAL;
  c->potSol.assign(pot, pot + c->t.NV);
  return IDFT_OK;
}
int idft_get_solution(const idft_ctx* c, double* pot) {
  if (!c || !pot) return IDFT_EINVAL;
  std::copy(c->potSol.begin(), c->potSol.end(), pot);
  return IDFT_OK;
}

}  // extern "C"
