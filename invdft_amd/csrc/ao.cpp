// Set-up stage entry points of the HIP library that need no mesh context:
// the atomic-orbital basis (reader, device evaluation, target density), the
// loss weights and the exchange-correlation potentials (include/idft_setup.h).
#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/idft_setup.h"
#include "common.h"
#include "setup.h"

using namespace idft;

struct idft_ao_basis {
  struct Fn {
    double c[3];
    int l = 0, m = 0;
    std::vector<double> alpha, cn;  // exponents, contraction coefficient x primitive norm
  };
  std::vector<Fn> f;
  std::vector<double> norm;  // basisNormConst per function
};

namespace {

const double kPi = 3.14159265358979323846;
// DFT-FE's Angstrom -> Bohr factor (C_AngToBohr).
const double kAngToBohr = 1.0 / 0.52917721067;

template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    return e.code;
  } catch (const std::exception&) {
    return IDFT_EINVAL;
  }
}

double odd_factorial(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// Norm of one radial primitive r^l exp(-a r^2) (getNormConsts,
// gaussianFunctionManager.cpp:251-262).
double primitive_norm(double a, int l) {
  const double ov = odd_factorial(2 * l + 1) * std::sqrt(kPi) /
                    (std::pow(2.0, 2 * l + 3.5) * std::pow(a, l + 1.5));
  return 1.0 / std::sqrt(ov);
}

// Angular normalisation C_l|m| D_m (Blm/Clm and Dm, :264-287).
double angular_norm(int l, int m) {
  const int am = std::abs(m);
  double b = std::sqrt((2.0 * l + 1) / 2.0);
  for (int k = 1; k <= am; ++k) b /= std::sqrt((l - k + 1.0) * (l + k));
  return b * (m == 0 ? 1.0 / std::sqrt(2.0 * kPi) : 1.0 / std::sqrt(kPi));
}

// Polynomials in (x, y, z) as monomial -> coefficient.
using Mono = std::array<int, 3>;
using Poly = std::map<Mono, double>;
Poly pmul(const Poly& a, const Poly& b) {
  Poly r;
  for (const auto& [ma, ca] : a)
    for (const auto& [mb, cb] : b) r[{ma[0] + mb[0], ma[1] + mb[1], ma[2] + mb[2]}] += ca * cb;
  return r;
}
Poly padd(const Poly& a, double sa, const Poly& b, double sb) {
  Poly r;
  for (const auto& [m, c] : a) r[m] += sa * c;
  for (const auto& [m, c] : b) r[m] += sb * c;
  return r;
}

// r^l P_l^|m|(z/r) {cos, sin}(|m| phi) (no Condon-Shortley phase) as a
// polynomial: (r sin theta)^|m| {cos, sin}(|m| phi) = {Re, Im}(x + i y)^|m|,
// times A_l = r^(l-|m|) d^|m| P_l / dx^|m| (z / r), which follows the
// derivative form of the Legendre recurrence
//   (k - m + 1) A_{k+1} = (2k + 1) z A_k - (k + m) r^2 A_{k-1},
// A_m = (2m - 1)!!, A_{m-1} = 0.
Poly solid_harmonic(int l, int m) {
  const int am = std::abs(m);
  const Poly z{{{0, 0, 1}, 1.0}}, r2{{{2, 0, 0}, 1.0}, {{0, 2, 0}, 1.0}, {{0, 0, 2}, 1.0}};
  Poly prev, cur{{{0, 0, 0}, odd_factorial(2 * am - 1)}};
  for (int k = am; k < l; ++k) {
    Poly next = padd(pmul(z, cur), (2.0 * k + 1) / (k - am + 1), pmul(r2, prev),
                     -(double)(k + am) / (k - am + 1));
    prev = cur;
    cur = next;
  }
  Poly T;
  if (m == 0) {
    T[{0, 0, 0}] = 1.0;
  } else {
    double binom = 1.0;
    for (int k = 0; k <= am; ++k) {
      // C(am, k) x^(am-k) (i y)^k
      const bool real = k % 2 == 0;
      const double sgn = ((k / 2) % 2 == 0) ? 1.0 : -1.0;
      if ((m > 0) == real) T[{am - k, k, 0}] += sgn * binom;
      binom = binom * (am - k) / (k + 1);
    }
  }
  Poly Y = pmul(cur, T);
  const double an = angular_norm(l, m);
  Poly out;
  for (const auto& [mm, c] : Y)
    if (c != 0.0) out[mm] = c * an;
  return out;
}

bool to_number(const std::string& w, double& v) {
  char* end = nullptr;
  v = std::strtod(w.c_str(), &end);
  return !w.empty() && end && *end == '\0';
}

struct ShellData {
  int l;
  std::vector<double> a, c;
};

// readGaussianFiles (:720-825)
void read_basis_file(const std::string& path, std::vector<ShellData>& out) {
  std::ifstream in(path);
  if (!in) throw Error(IDFT_EINVAL, "cannot open basis file " + path);
  std::string line;
  std::getline(in, line);  // the first line names the element
  const std::string kL = "SPDFGHI";
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string chars;
    if (!(ls >> chars)) continue;
    for (char& ch : chars) ch = (char)std::toupper((unsigned char)ch);
    if (chars.find_first_not_of(kL) != std::string::npos)
      throw Error(IDFT_EINVAL, "undefined L character(s) in " + path);
    std::string ns;
    ls >> ns;
    char* end = nullptr;
    const long n = std::strtol(ns.c_str(), &end, 10);
    if (ns.empty() || !end || *end != '\0' || n < 1)
      throw Error(IDFT_EINVAL, "number of contracted Gaussians is not an integer in " + path);
    std::vector<double> a(n);
    std::vector<std::vector<double>> c(n, std::vector<double>(chars.size()));
    for (long i = 0; i < n; ++i) {
      if (!std::getline(in, line) || line.empty())
        throw Error(IDFT_EINVAL, "missing or empty contraction row in " + path);
      std::istringstream lc(line);
      std::string w;
      if (!(lc >> w) || !to_number(w, a[i])) throw Error(IDFT_EINVAL, "exponent is not a number in " + path);
      for (size_t j = 0; j < chars.size(); ++j)
        if (!(lc >> w) || !to_number(w, c[i][j]))
          throw Error(IDFT_EINVAL, "coefficient is not a number in " + path);
    }
    for (size_t j = 0; j < chars.size(); ++j) {
      ShellData s;
      s.l = (int)kL.find(chars[j]);
      s.a = a;
      for (long i = 0; i < n; ++i) s.c.push_back(c[i][j]);
      out.push_back(std::move(s));
    }
  }
}

// The basis on the device (AoDev arrays), monomial expansions included.
struct AoUpload {
  DevBuf<double> center, alpha, cn, coef, norm;
  DevBuf<int> pstart, mstart, mpow;
  AoDev view{};
  void load(const idft_ao_basis& b, hipStream_t s) {
    const int nb = (int)b.f.size();
    std::vector<double> cen, al, c, cf;
    std::vector<int> ps{0}, ms{0}, mp;
    std::map<std::pair<int, int>, Poly> cache;
    for (const auto& f : b.f) {
      cen.insert(cen.end(), f.c, f.c + 3);
      al.insert(al.end(), f.alpha.begin(), f.alpha.end());
      c.insert(c.end(), f.cn.begin(), f.cn.end());
      ps.push_back((int)al.size());
      auto key = std::make_pair(f.l, f.m);
      if (!cache.count(key)) cache[key] = solid_harmonic(f.l, f.m);
      for (const auto& [mm, v] : cache[key]) {
        mp.insert(mp.end(), mm.begin(), mm.end());
        cf.push_back(v);
      }
      ms.push_back((int)cf.size());
    }
    center.upload(cen.data(), cen.size(), s);
    alpha.upload(al.data(), al.size(), s);
    cn.upload(c.data(), c.size(), s);
    coef.upload(cf.data(), cf.size(), s);
    norm.upload(b.norm.data(), nb, s);
    pstart.upload(ps.data(), ps.size(), s);
    mstart.upload(ms.data(), ms.size(), s);
    mpow.upload(mp.data(), mp.size(), s);
    view = AoDev{nb, center.p, pstart.p, alpha.p, cn.p, mstart.p, mpow.p, coef.p, norm.p};
  }
};

struct Stream {
  hipStream_t s = nullptr;
  Stream() { IDFT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~Stream() { if (s) (void)hipStreamDestroy(s); }
  void sync() { IDFT_HIP(hipStreamSynchronize(s)); }
};

struct Blas {
  rocblas_handle h = nullptr;
  explicit Blas(hipStream_t s) {
    if (rocblas_create_handle(&h) != rocblas_status_success) throw Error(IDFT_EHIP, "rocblas_create_handle failed");
    rocblas_set_stream(h, s);
  }
  ~Blas() { if (h) rocblas_destroy_handle(h); }
};

QuadGeo quad_geo(const Tables& t) {
  QuadGeo q{};
  q.nc = t.nc; q.q1 = t.q1; q.L = t.L; q.h = t.h;
  if (t.q1 > 16) throw Error(IDFT_EINVAL, "too many quadrature points per axis");
  for (int i = 0; i < t.q1; ++i) q.gx[i] = t.gq_x[i];
  return q;
}

}  // namespace

extern "C" {

int idft_ao_basis_read(const char* coords_file, const char* basis_dir, int32_t angstrom,
                       idft_ao_basis** out) {
  if (!coords_file || !out) return IDFT_EINVAL;
  *out = nullptr;
  return guarded([&] {
    std::ifstream in(coords_file);
    if (!in) throw Error(IDFT_EINVAL, std::string("cannot open ") + coords_file);
    auto b = std::make_unique<idft_ao_basis>();
    std::map<std::string, std::vector<ShellData>> files;
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string sym, w[3], file;
      if (!(ls >> sym)) continue;
      double c[3];
      if (!(ls >> w[0] >> w[1] >> w[2] >> file)) throw Error(IDFT_EINVAL, "malformed atom line");
      for (int k = 0; k < 3; ++k)
        if (!to_number(w[k], c[k])) throw Error(IDFT_EINVAL, "coordinate is not a number");
      if (angstrom)
        for (double& v : c) v *= kAngToBohr;
      std::string path = file;
      if (basis_dir && *basis_dir && file[0] != '/') path = std::string(basis_dir) + "/" + file;
      if (!files.count(path)) read_basis_file(path, files[path]);
      for (const ShellData& sh : files[path]) {
        std::vector<int> ms;
        if (sh.l == 1) ms = {1, -1, 0};  // the reference's x, y, z order of p shells
        else
          for (int m = -sh.l; m <= sh.l; ++m) ms.push_back(m);
        for (int m : ms) {
          idft_ao_basis::Fn f;
          std::copy(c, c + 3, f.c);
          f.l = sh.l;
          f.m = m;
          f.alpha = sh.a;
          for (size_t k = 0; k < sh.a.size(); ++k) f.cn.push_back(sh.c[k] * primitive_norm(sh.a[k], sh.l));
          b->f.push_back(std::move(f));
        }
      }
    }
    if (b->f.empty()) throw Error(IDFT_EINVAL, "no atoms read");
    b->norm.assign(b->f.size(), 1.0);
    *out = b.release();
    return IDFT_OK;
  });
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
    if (line.empty()) return IDFT_EINVAL;  // readMatrix rejects empty lines
    std::istringstream ls(line);
    int col = 0;
    double v;
    while (ls >> w) {
      if (!to_number(w, v) || row >= n || col >= n) return IDFT_EINVAL;
      out[(size_t)row * n + col++] = v;
    }
    if (col != n) return IDFT_EINVAL;
    ++row;
  }
  return row == n ? IDFT_OK : IDFT_EINVAL;
}

int idft_ao_values(const idft_ao_basis* b, int64_t np, const double* xyz, double* vals) {
  if (!b || !xyz || !vals || np < 0) return IDFT_EINVAL;
  return guarded([&] {
    Stream st;
    AoUpload up;
    up.load(*b, st.s);
    const int nb = (int)b->f.size();
    DevBuf<double> pts, phi((size_t)std::max<int64_t>(np, 1) * nb);
    pts.upload(xyz, (size_t)np * 3, st.s);
    launch_ao_values_at(up.view, np, pts.p, phi.p, st.s);
    std::vector<double> h((size_t)np * nb);
    phi.download(h.data(), h.size(), st.s);
    st.sync();
    for (int i = 0; i < nb; ++i)
      for (int64_t p = 0; p < np; ++p) vals[(size_t)i * np + p] = h[(size_t)p * nb + i];
    return IDFT_OK;
  });
}

int idft_ao_density_quad(const idft_desc* d, idft_ao_basis* b, int32_t normalize, int32_t n_spin,
                         const double* P, double* rho_q, double* grad_q, double* s_diag) {
  if (!desc_valid(d) || !b || !P || !rho_q || n_spin < 1 || n_spin > 2) return IDFT_EINVAL;
  return guarded([&] {
    Tables t;
    t.init(*d);
    const QuadGeo qg = quad_geo(t);
    const int nb = (int)b->f.size();
    const int64_t nQ = t.ncell * (int64_t)t.nq;
    Stream st;
    Blas blas(st.s);
    // JxW at the parent Gauss points (uniform cells)
    const double hh = 0.5 * t.h, J = hh * hh * hh;
    std::vector<double> jw(t.nq);
    for (int q = 0; q < t.nq; ++q)
      jw[q] = J * t.gq_w[q % t.q1] * t.gq_w[(q / t.q1) % t.q1] * t.gq_w[q / (t.q1 * t.q1)];
    DevBuf<double> jxw;
    jxw.upload(jw.data(), jw.size(), st.s);
    // rows per chunk: phi, T and (with gradients) 3 phi per row
    const int64_t per_row = (int64_t)nb * 8 * (grad_q ? 5 : 2);
    const int64_t chunk = std::max<int64_t>(4096, std::min<int64_t>(nQ, (int64_t(3) << 30) / per_row));
    DevBuf<double> phi((size_t)chunk * nb), T((size_t)chunk * nb), gphi, rho((size_t)chunk),
        grad;
    if (grad_q) { gphi.resize((size_t)chunk * nb * 3); grad.resize((size_t)chunk * 3); }
    // pass 1: self-overlaps S_ii with the current normalisation set to 1
    b->norm.assign(nb, 1.0);
    AoUpload up;
    up.load(*b, st.s);
    constexpr int kMaxBlk = 1024;
    DevBuf<double> partial((size_t)kMaxBlk * nb);
    std::vector<double> S(nb, 0.0), hp((size_t)kMaxBlk * nb);
    for (int64_t r0 = 0; r0 < nQ; r0 += chunk) {
      const int64_t rows = std::min(chunk, nQ - r0);
      launch_ao_values(qg, up.view, r0, rows, phi.p, nullptr, st.s);
      const int nblk = launch_ao_selfoverlap(qg, nb, r0, rows, phi.p, jxw.p, partial.p, kMaxBlk, st.s);
      partial.download(hp.data(), (size_t)nblk * nb, st.s);
      st.sync();
      for (int k = 0; k < nblk; ++k)
        for (int i = 0; i < nb; ++i) S[i] += hp[(size_t)k * nb + i];
    }
    if (s_diag) std::copy(S.begin(), S.end(), s_diag);
    if (normalize)
      for (int i = 0; i < nb; ++i) b->norm[i] = 1.0 / std::sqrt(S[i]);
    up.load(*b, st.s);
    // pass 2: T = phi P_sym (rocBLAS), rho = rowsum(T o phi), grad = 2 T gphi
    DevBuf<double> Pd((size_t)nb * nb);
    std::vector<double> Ps((size_t)nb * nb);
    for (int s = 0; s < n_spin; ++s) {
      const double* Pm = P + (size_t)s * nb * nb;
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) Ps[(size_t)i * nb + j] = Pm[(size_t)std::min(i, j) * nb + std::max(i, j)];
      Pd.upload(Ps.data(), Ps.size(), st.s);
      for (int64_t r0 = 0; r0 < nQ; r0 += chunk) {
        const int64_t rows = std::min(chunk, nQ - r0);
        launch_ao_values(qg, up.view, r0, rows, phi.p, grad_q ? gphi.p : nullptr, st.s);
        // row-major T[rows][nb] = phi[rows][nb] P: column-major T^T = P phi^T
        const double one = 1.0, zero = 0.0;
        if (rocblas_dgemm(blas.h, rocblas_operation_none, rocblas_operation_none, nb, (int)rows, nb,
                          &one, Pd.p, nb, phi.p, nb, &zero, T.p, nb) != rocblas_status_success)
          throw Error(IDFT_EHIP, "rocblas_dgemm failed");
        launch_ao_rows(nb, rows, T.p, phi.p, grad_q ? gphi.p : nullptr, rho.p,
                       grad_q ? grad.p : nullptr, st.s);
        IDFT_HIP(hipMemcpyAsync(rho_q + (size_t)s * nQ + r0, rho.p, rows * sizeof(double),
                                hipMemcpyDeviceToHost, st.s));
        if (grad_q)
          IDFT_HIP(hipMemcpyAsync(grad_q + ((size_t)s * nQ + r0) * 3, grad.p, rows * 3 * sizeof(double),
                                  hipMemcpyDeviceToHost, st.s));
        st.sync();
      }
    }
    return IDFT_OK;
  });
}

int idft_loss_weights(int64_t n, const double* rho, double sf, double a1, double a2, double tau,
                      double* w) {
  if (n < 0 || !rho || !w) return IDFT_EINVAL;
  return guarded([&] {
    Stream st;
    DevBuf<double> r, out((size_t)std::max<int64_t>(n, 1));
    r.upload(rho, (size_t)n, st.s);
    launch_loss_weights(n, r.p, sf, a1, a2, tau, out.p, st.s);
    out.download(w, (size_t)n, st.s);
    st.sync();
    return IDFT_OK;
  });
}

void idft_guess_params_default(idft_guess_params* gp) {
  if (!gp) return;
  gp->lb94 = 1; gp->tau_vx = 1e-2; gp->fermi_amaldi = 1; gp->tau_fa = 1e-2;
  gp->factor_fa = 1.0; gp->factor_lda = 0.0; gp->spin_factor = 2.0; gp->n_electrons = 0.0;
}

int idft_xc_potentials(int64_t n, const double* r, const double* sigma, int32_t lb94, double* vx,
                       double* vc) {
  if (n < 0 || !r || !vx || !vc || (lb94 && !sigma)) return IDFT_EINVAL;
  return guarded([&] {
    Stream st;
    DevBuf<double> rd, sd, x((size_t)std::max<int64_t>(n, 1)), c((size_t)std::max<int64_t>(n, 1));
    rd.upload(r, (size_t)n, st.s);
    if (lb94) sd.upload(sigma, (size_t)n, st.s);
    launch_xc(n, rd.p, sd.p, lb94, x.p, c.p, st.s);
    x.download(vx, (size_t)n, st.s);
    c.download(vc, (size_t)n, st.s);
    st.sync();
    return IDFT_OK;
  });
}

}  // extern "C"
