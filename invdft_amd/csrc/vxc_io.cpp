// v_xc checkpoint files (text records of the reference, and a binary
// variant).  Host code only; the v_xc mesh is the uniform child grid of
// (2 n_cells + 1)^3 nodes on [-L, L]^3, so the reference's R-tree
// nearest-node search reduces to rounding each coordinate to the grid.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "common.h"
#include "../../include/idft_io.h"

namespace {

struct VxcGrid {
  int64_t nv1 = 0, nv = 0;
  double L = 0.0, hc = 0.0;
  explicit VxcGrid(const idft_desc& d) {
    nv1 = 2 * (int64_t)d.n_cells + 1;
    nv = nv1 * nv1 * nv1;
    L = d.half_width;
    hc = 2.0 * L / (double)(nv1 - 1);
  }
  double coord(int64_t i) const { return -L + hc * (double)i; }
};

bool spin_ok(int32_t n_spin) { return n_spin == 1 || n_spin == 2; }

const char kVxcMagic[8] = {'I', 'D', 'F', 'T', 'V', 'X', 'C', '1'};

}  // namespace

extern "C" {

// writeChildMeshDataToFile (InverseDFTSolverFunction.cpp:593-645): node id,
// its coordinates, then the value of every spin channel.
int idft_write_vxc_text(const idft_desc* d, const char* path, const double* pot,
                        int32_t n_spin) {
  if (!idft::desc_valid(d) || !path || !pot || !spin_ok(n_spin)) return IDFT_EINVAL;
  const VxcGrid g(*d);
  std::FILE* f = std::fopen(path, "w");
  if (!f) return IDFT_EINVAL;
  for (int64_t n = 0; n < g.nv; ++n) {
    const int64_t i = n % g.nv1, j = (n / g.nv1) % g.nv1, k = n / (g.nv1 * g.nv1);
    std::fprintf(f, "%lld %.17g %.17g %.17g", (long long)n, g.coord(i), g.coord(j), g.coord(k));
    for (int s = 0; s < n_spin; ++s) std::fprintf(f, " %.17g", pot[(int64_t)s * g.nv + n]);
    std::fputc('\n', f);
  }
  const bool ok = std::ferror(f) == 0;
  return (std::fclose(f) == 0 && ok) ? IDFT_OK : IDFT_EINVAL;
}

// readVxcDataFromFileWithSearch (InverseDFTEngine.cpp:2982-3142).
int idft_read_vxc_text(const idft_desc* d, const char* path, double dist_tol, int32_t n_spin,
                       double* pot) {
  if (!idft::desc_valid(d) || !path || !pot || !spin_ok(n_spin) || !(dist_tol > 0.0))
    return IDFT_EINVAL;
  const VxcGrid g(*d);
  std::ifstream in(path);
  if (!in) return IDFT_EINVAL;
  std::vector<char> assigned((size_t)g.nv, 0);
  for (int64_t n = 0; n < (int64_t)n_spin * g.nv; ++n) pot[n] = 0.0;
  const double lo = -g.L - 1e-2, hi = g.L + 1e-2;  // bounding box widened by 1e-2
  for (int64_t rec = 0; rec < g.nv; ++rec) {
    double id, x, y, z, v[2] = {0.0, 0.0};
    if (!(in >> id >> x >> y >> z)) break;
    for (int s = 0; s < n_spin; ++s)
      if (!(in >> v[s])) return IDFT_EINVAL;
    if (x < lo || y < lo || z < lo || x > hi || y > hi || z > hi) continue;
    int64_t idx[3];
    const double p[3] = {x, y, z};
    double dist2 = 0.0;
    for (int a = 0; a < 3; ++a) {
      int64_t i = (int64_t)std::llround((p[a] + g.L) / g.hc);
      i = i < 0 ? 0 : (i >= g.nv1 ? g.nv1 - 1 : i);
      idx[a] = i;
      const double dd = p[a] - g.coord(i);
      dist2 += dd * dd;
    }
    if (!(std::sqrt(dist2) < dist_tol)) continue;
    const int64_t node = idx[0] + g.nv1 * (idx[1] + g.nv1 * idx[2]);
    if (assigned[(size_t)node]) return IDFT_EINVAL;  // two records nearest to one node
    assigned[(size_t)node] = 1;
    for (int s = 0; s < n_spin; ++s) pot[(int64_t)s * g.nv + node] = v[s];
  }
  for (char a : assigned)
    if (!a) return IDFT_EINVAL;  // a node without a record
  return IDFT_OK;
}

int idft_write_vxc_binary(const idft_desc* d, const char* path, const double* pot,
                          int32_t n_spin) {
  if (!idft::desc_valid(d) || !path || !pot || !spin_ok(n_spin)) return IDFT_EINVAL;
  const VxcGrid g(*d);
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return IDFT_EINVAL;
  const int32_t nc = d->n_cells;
  const double L = d->half_width;
  const int64_t nv = g.nv;
  bool ok = std::fwrite(kVxcMagic, 1, 8, f) == 8 && std::fwrite(&nc, sizeof nc, 1, f) == 1 &&
            std::fwrite(&n_spin, sizeof n_spin, 1, f) == 1 &&
            std::fwrite(&L, sizeof L, 1, f) == 1 && std::fwrite(&nv, sizeof nv, 1, f) == 1 &&
            std::fwrite(pot, sizeof(double), (size_t)(n_spin * nv), f) == (size_t)(n_spin * nv);
  ok = (std::fclose(f) == 0) && ok;
  return ok ? IDFT_OK : IDFT_EINVAL;
}

int idft_read_vxc_binary(const idft_desc* d, const char* path, int32_t n_spin, double* pot) {
  if (!idft::desc_valid(d) || !path || !pot || !spin_ok(n_spin)) return IDFT_EINVAL;
  const VxcGrid g(*d);
  std::FILE* f = std::fopen(path, "rb");
  if (!f) return IDFT_EINVAL;
  char magic[8];
  int32_t nc = 0, ns = 0;
  double L = 0.0;
  int64_t nv = 0;
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kVxcMagic, 8) == 0 &&
            std::fread(&nc, sizeof nc, 1, f) == 1 && std::fread(&ns, sizeof ns, 1, f) == 1 &&
            std::fread(&L, sizeof L, 1, f) == 1 && std::fread(&nv, sizeof nv, 1, f) == 1;
  ok = ok && nc == d->n_cells && ns == n_spin && L == d->half_width && nv == g.nv;
  ok = ok && std::fread(pot, sizeof(double), (size_t)(n_spin * nv), f) == (size_t)(n_spin * nv);
  std::fclose(f);
  return ok ? IDFT_OK : IDFT_EINVAL;
}

}  // extern "C"
