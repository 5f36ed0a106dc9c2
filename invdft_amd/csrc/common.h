// Internal declarations of the HIP inverse-DFT library (libidft.so).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/idft.h"
#include "../../include/idft_hip.h"
#include "../../include/idft_io.h"

namespace idft {

// Error carried from a failing HIP call or an invalid state up to the C ABI.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define IDFT_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      throw ::idft::Error(IDFT_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device bytes held by DevBuf allocations of this library, process-wide: the
// current total and its peak since the last reset (idft_device_memory).
inline std::atomic<int64_t>& devmem_current() {
  static std::atomic<int64_t> v{0};
  return v;
}
inline std::atomic<int64_t>& devmem_peak() {
  static std::atomic<int64_t> v{0};
  return v;
}
inline void devmem_add(int64_t bytes) {
  const int64_t now = devmem_current().fetch_add(bytes) + bytes;
  int64_t pk = devmem_peak().load();
  while (now > pk && !devmem_peak().compare_exchange_weak(pk, now)) {
  }
}

// Owning device buffer (hipMalloc'd, freed on destruction).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { resize(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      devmem_add(-(int64_t)(n * sizeof(T)));
    }
    p = nullptr; n = 0;
  }
  // Grow-only reallocation; contents are not preserved.
  void resize(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) { p = nullptr; throw Error(IDFT_ENOMEM, "hipMalloc failed"); }
    n = count;
    devmem_add((int64_t)(n * sizeof(T)));
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    resize(count);
    IDFT_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T* h, size_t count, hipStream_t s) const {
    IDFT_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// Host-side finite-element tables of one (n_cells, p) discretisation.
struct Tables {
  int nc = 0, p = 0, n1 = 0, q1 = 0, ne = 0, nq = 0, N1 = 0, NV1 = 0;
  int64_t N = 0, ncell = 0, NV = 0;
  double L = 0, h = 0;
  std::vector<double> gll_x, gll_w, gq_x, gq_w;
  std::vector<double> Lq;      // [q1][n1] basis at parent Gauss points
  std::vector<double> Lc;      // [4][n1]  basis at the child Gauss points
  std::vector<double> M1, D1;  // [n1][n1] 1D mass and stiffness on [-1,1]
  std::vector<double> Wq;      // [q1][3] child Q1 basis (3 nodes / parent axis) at parent points
  std::vector<double> Wc;      // [4][3]  child Q1 basis at the child Gauss points
  void init(const idft_desc& d);
};

bool desc_valid(const idft_desc* d);

// ---- small dense linear algebra on the host (B x B, FP64)
void sym_eig(int n, std::vector<double> A, std::vector<double>& evals, std::vector<double>& V);
bool cholesky_lower(int n, std::vector<double>& S);
uint64_t splitmix64(uint64_t& s);
double rand_pm1(uint64_t& s);
double fd_occ(double e, double mu, double T);
double fd_occ_der(double e, double mu, double T);
double fermi_level(const std::vector<double>& e, double ne, double T);

// ---- device kernels (kernels.hip) -------------------------------------------
// Geometry passed to kernels by value.
struct Geo {
  int nc, p, n1, N1, q1, NV1;
  int64_t N, ncell;
};

// Constant tables (uploaded once per context).
struct DevTables {
  DevBuf<double> Lq;  // [p+3][p+1] parent basis at parent Gauss points
  DevBuf<double> Lc;  // [4][p+1]   parent basis at child Gauss points
  DevBuf<double> M1, D1;  // [p+1][p+1] 1D mass / stiffness on [-1,1]
  DevBuf<double> Wq;  // [p+3][3] child Q1 basis (3 child nodes per parent axis) at parent points
  DevBuf<double> Wc;  // [4][3]   child Q1 basis at child Gauss points
};

constexpr int kMaxCols = 80;  // widest block the cell apply handles in one pass

// Per-column MINRES state slots: st[slot * B + col].
enum MinresSlot {
  MS_BETA, MS_OLDB, MS_DBAR, MS_EPSLN, MS_PHIBAR, MS_CS, MS_SN, MS_ALFA, MS_DONE, MS_IT, MS_TOL,
  MS_BETA1, MS_CV, MS_CVN, MS_CR1, MS_CA2, MS_CWV, MS_CWW1, MS_CWW, MS_CX,
  MS_Q2, MS_Q3,  // D^2 r2.r2 and D r2.r1 of the current step (fused steps)
  MS_COUNT
};

int64_t hamiltonian_tile_elems(int p);
// H_e = K_e/2 + V_e(veff) for all cells in the MFMA tile layout (double and,
// when h32 != nullptr, float).  jxw[q] is the (uniform) parent JxW.
void launch_build_hamiltonian(const Geo& g, const DevTables& t, double hh, const double* veff,
                              const double* jxw, double* h64, float* h32, hipStream_t s);
// Output combination of one operator application:
//   Y = beta Y + alpha S[n] (sum_e P_e^T H_e P_e S V)[n] + a ca xn X + b cb Z + c cc W
// (null coefficient arrays read as 1, null blocks as 0), where the operator
// input is V = xcol[j] xn[n] Xin (xcol, xn may be null).  With dot_partial,
// dot_partial[blk][j] receives the block partials of V.Y (gather_blocks
// blocks) and V, if given, the scaled input.
template <typename T>
struct ApplyEpilogue {
  T* Y = nullptr;
  double alpha = 1.0, beta = 0.0;
  const T* X = nullptr; double a = 0.0; const double* ca = nullptr;
  const T* Z = nullptr; double b = 0.0; const double* cb = nullptr;
  const T* W = nullptr; double c = 0.0; const double* cc = nullptr;
  const double* xn = nullptr;    // node scale of the operator input and of X
  const double* xcol = nullptr;  // column scale of the operator input
  T* V = nullptr;                // stores the scaled input xcol xn Xin (needs dot_partial)
  double* dot_partial = nullptr;
};
constexpr int kGatherBlocks = 16384;
// grid cap of the fused MINRES gather (its work-groups loop over node chunks),
// which bounds its partial rows per step to 3 x B x 4096
constexpr int kMinresGatherBlocks = 4096;
// rows of the gather's dot partials for vectors of vbytes-byte elements
int gather_blocks(const Geo& g, int B, int rows, int vbytes);
// Cell-product kernels.
enum CellKernel {
  CK_ROW_PASS = 0,   // stored H_e, one wave per 16-row block (mirror tiles read again)
  CK_TILE_PASS = 1,  // stored H_e, every stored tile read once per row half
  CK_SUM_FACTOR = 2  // no stored H_e: 1D contractions with the per-cell quadrature weights
};
// Launch shape of the cell product (chosen by the autotuner).
struct GemmShape {
  int cells_per_wg = 1;       // stored-H kernels: cells processed in sequence by one work-group;
                              // sum-factorised kernel: columns per item (1)
  int vectors_per_tile = 80;  // stored-H kernels: columns per tile (rounded to 16; >= B: one
                              // pass); sum-factorised kernel: columns per work-group (2, 4, 8)
  int gather_rows = 0;        // node chunks (256 / (B / VW) nodes each) per gather work-group
                              // (0: about kGatherBlocks groups)
  int kernel = CK_ROW_PASS;   // CellKernel; the tile pass applies to 16-column tiles (FP32 also 32)
};
// The cell operator at one precision: the stored upper tiles of H_e (row and
// tile pass; null until a stored-H kernel first runs) and, for the
// sum-factorised kernel, the quadrature weights w_e[cell][q] = v_eff JxW with
// the kinetic scale h/4.  On the uniform mesh
//   H_e = (h/4) (D (x) M (x) M + M (x) D (x) M + M (x) M (x) D) + L^T diag(w_e) L
// with the 1D tables of sf_tables.h.
template <typename T>
struct CellOp {
  const T* H = nullptr;
  const T* W = nullptr;
  double kin_scale = 0.0;
};
// Cell operator in two launches: per-cell H_e (S xcol Xin)_e into the scratch
// Yc (cell_scratch_elems(g, B) elements), then the node gather + epilogue.
// Returns the number of gather blocks (rows of dot_partial).
// T: cell operator, cell scratch Yc and arithmetic; TV: the node-major
// vectors (Xin and the epilogue's blocks).  Sg (type T) scales the input of
// the cell products, S (type TV) the gathered sums.
template <typename T, typename TV>
int launch_cell_apply(const Geo& g, const CellOp<T>& op, const T* Sg, const TV* S, const TV* Xin,
                      T* Yc, int B, const ApplyEpilogue<TV>& ep, const GemmShape& shape,
                      hipStream_t s);
int64_t cell_scratch_elems(const Geo& g, int B);
// The FP64 even-odd coefficients of the sum-factorised kernel into constant
// memory of the current device (once per device; every context calls it).
void upload_sf_constants();
// Capacity (FP64 elements, 64 MiB) of the fixed device scratch of the
// allocation-free operator entry point, and its device address on the current
// device (noalloc.hip; looked up once per device).
constexpr int64_t kNoallocScratchElems = int64_t(8) << 20;
double* noalloc_scratch();
// Phase 1 alone (per-cell products into Yc), for profiling.
template <typename T>
void launch_cell_gemm(const Geo& g, const CellOp<T>& op, const T* S, const T* X,
                      const double* xcol, T* Yc, int B, const GemmShape& shape, hipStream_t s);
// Columns per work-group / per item the sum-factorised kernel is built for.
bool sf_cols_supported(int cols, bool fp64);
bool sf_vc_supported(int cols, int vc, bool fp64);
// Y = beta Y + ns[n] (a ca[j] X + b cb[j] Z); null pointers read as 1 (X/Z as 0).
template <typename T>
void launch_lincomb(int64_t N, int B, T* Y, double beta, const T* X, double a, const double* ca,
                    const T* Z, double b, const double* cb, const double* ns, hipStream_t s);
template <typename T>
void launch_lincomb3(int64_t N, int B, T* Y, double beta, const T* X, double a, const double* ca,
                     const T* Z, double b, const double* cb, const T* W, double c,
                     const double* cc, const double* ns, hipStream_t s);
// Y = a ca[j] X + b cb[j] Z in FP64, rounded once to Y's type (FP64 X and Z
// into FP32 Y; FP64 X and FP32 Z into FP64 Y).
void launch_lincomb_d2f(int64_t N, int B, float* Y, const double* X, double a, const double* ca,
                        const double* Z, double b, const double* cb, hipStream_t s);
void launch_lincomb_df2d(int64_t N, int B, double* Y, const double* X, double a,
                         const double* ca, const float* Z, double b, const double* cb,
                         hipStream_t s);
void launch_convert_d2f(int64_t n, float* y, const double* x, hipStream_t s);
void launch_convert_f2d(int64_t n, double* y, const float* x, hipStream_t s);
int coldot_blocks(int64_t N, int B);
// partial[blk][j] = sum_n w[n] X[n][j] Y[n][j] (w may be null); returns #blocks
template <typename T>
int launch_coldot_partial(int64_t N, int B, const T* X, const T* Y, const double* w,
                          double* partial, hipStream_t s);
void launch_sum_partials(int nblk, int B, const double* partial, double* out, hipStream_t s);
int gram_partial_blocks(int64_t N, int B);
// G = X^T Y (B x B)
void launch_gram(int64_t N, int B, const double* X, const double* Y, double* partial, double* G,
                 hipStream_t s);
// Y = beta Y + ns[n] (X Q)[n]  (Q row-major B x B on device, X != Y)
void launch_right_mult(int64_t N, int B, const double* X, const double* Q, double* Y,
                       const double* ns, double beta, hipStream_t s);
int launch_residual_partial(int64_t N, int B, const double* HX, const double* X,
                            const double* ev, double* partial, hipStream_t s);
void launch_vxc_to_quad(const Geo& g, const DevTables& t, const double* pot,
                        const double* potBase, double* veff, hipStream_t s);
// rho(q) = sum_j 2 occ[j] (sum_a N_a(q) S[n_a] X[n_a][j])^2
void launch_density(const Geo& g, const DevTables& t, const double* X, const double* S, int B,
                    const double* occ, double* rho, hipStream_t s);
// u = w (rho_t - rho/2); partial[blk][4] = (loss, loss_unweighted, l1, int rho v)
void launch_loss(int64_t nQ, int nq, const double* rho, const double* rhoT, const double* w,
                 const double* veff, const double* jxw, double* u, double* partial,
                 hipStream_t s);
constexpr int kLossBlocks = 256;
// out[cell][k] = sum_q wq[cell][q] A_{i_k}(q) B_{j_k}(q), pairs (i_k, j_k)
void launch_pair_quad(const Geo& g, const DevTables& t, const double* A, const double* Bv, int B,
                      const int* pairs, int np, const double* wq, int64_t wstride, double* out,
                      hipStream_t s);
void launch_copy_cols(int64_t N, int Bs, int off, int Bd, const double* src, double* dst,
                      hipStream_t s);
// dst[n][k] = src[n][idx[k]], k < Bd (gather) / dst[n][idx[k]] = src[n][k], k < Bs (scatter)
void launch_gather_cols(int64_t N, int Bs, const double* src, int Bd, const int* idx, double* dst,
                        hipStream_t s);
void launch_scatter_cols(int64_t N, int Bs, const double* src, int Bd, const int* idx, double* dst,
                         hipStream_t s);
// out[q] = u[q] * jxw[q % nq]
void launch_mul_q(int64_t nQ, int nq, const double* u, const double* jxw, double* out,
                  hipStream_t s);
// partial[blk][k] = sum_n w[n] X[n][i_k] Y[n][j_k] (w may be null)
int launch_pairdot_partial(int64_t N, int B, const double* X, const double* Y, const double* w,
                           const int* pairs, int np, double* partial, hipStream_t s);
// Y += N^T diag(wq) N X over the cells of one colour
void launch_quad_operator(const Geo& g, const DevTables& t, const double* X, int B,
                          const double* wq, double* Y, int colour, hipStream_t s);
// force[child node] += int N^child sum_j Psi_j P_j over one colour; mode 0 = child
// Gauss points (weight jxwChild), mode 1 = parent Gauss points (weights jxwParent[q])
void launch_force(const Geo& g, const DevTables& t, const double* Psi, const double* P, int B,
                  int mode, double jxwChild, const double* jxwParent, double* force, int colour,
                  hipStream_t s);
void launch_zero_rows(int64_t N, int B, double* X, const unsigned char* mask, hipStream_t s);
// MINRES
// nq = 2: partial also holds (D r)^2 partials at row offset nblk (into Q2)
void launch_minres_start(int B, int nblk, const double* partial, const double* negEig,
                         double* st, hipStream_t s, int nq = 1);
void launch_minres_alpha(int B, int nblk, const double* partial, double* st, hipStream_t s);
template <typename T>
int launch_minres_lanczos(int64_t N, int B, const T* y, const T* r2, T* rnew, const double* D,
                          const double* st, double* partial, hipStream_t s);
// nq = 3: partial also holds the Q2 and Q3 partials at row offsets nblk, 2 nblk
void launch_minres_givens(int B, int nblk, const double* partial, const double* negEig,
                          double* st, hipStream_t s, int nq = 1);
template <typename T>
void launch_minres_update(int64_t N, int B, const T* r, const double* D, T* w1, const T* w, T* x,
                          const double* st, hipStream_t s);
void launch_copy_done(int B, const double* st, int* flags, hipStream_t s);
// TFLOP/s of `chains` independent FMA chains per thread over 32 work-groups
// per CU (FP64 or FP32, unpacked), timed with HIP events on stream s
double launch_fma_probe(bool fp64, int chains, hipStream_t s);
// Fused MINRES step for sum-factorised cell products (shape.kernel ==
// CK_SUM_FACTOR), four launches per step instead of five and four block
// streams fewer (9 instead of 13, plus one update launch per run of steps):
//   start_dots   partials of r.D r and (D r)^2 of the right-hand side (2 x nblk
//                rows), reduced by launch_minres_start(nq = 2) into beta and Q2
//   cell_dot     cell product of v = cv D r2 (node scale D, column scale cv)
//                with the per-cell dots x_e.(H_e x_e) (column-major, B x ncell)
//   alpha_cell   alpha = sum of the cell dots + cv (cvn Q2 + cr1 Q3)
//   gather       r1 <- S sum + (cvn D + ca2) r2 + cr1 r1 with the three next
//                dots (3 x nblk rows), reduced by launch_minres_givens(nq = 3);
//                with w != nullptr also the previous step's update
//                w <- cwv D r1 + cww1 w + cww w1, x += cx w (read before r1 is
//                overwritten), so that launch_minres_update runs only after the
//                last step of a run of steps
int launch_minres_start_dots(int64_t N, int B, const double* r, const double* D, double* partial,
                             hipStream_t s);
template <typename T, typename TV>
void launch_cell_dot(const Geo& g, const CellOp<T>& op, const T* Sg, const TV* X, int B,
                     const double* xn, const double* xcol, T* Yc, double* cell_dots,
                     const GemmShape& shape, hipStream_t s);
void launch_minres_alpha_cell(int B, int ncell, const double* cell_dots, double* st, hipStream_t s);
template <typename TY>
int launch_minres_gather(const Geo& g, const TY* Yc, const double* S, const double* D,
                         const double* r2, double* r1, int B, const double* st, double* partial,
                         int gather_rows, hipStream_t s, double* w = nullptr,
                         const double* w1 = nullptr, double* x = nullptr);

// L-BFGS checkpoint settings of a context (idft_bfgs_set_checkpoint).
struct BfgsCheckpoint {
  std::string path;  // empty: off
  int every = 1;
  bool resume = false;
};
const BfgsCheckpoint& bfgs_checkpoint_of(const idft_ctx* c);

}  // namespace idft
