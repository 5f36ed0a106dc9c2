// Device kernels of the MI355X inverse-DFT hot path (gfx950 / CDNA4).
//
// Layouts
//   multivectors  node-major X[node * B + col]
//   quad data     [cell][q], q = (qz*Q + qy)*Q + qx
//   H tiles       H_e is symmetric, so per cell only the 16x16 tiles (rt, kg)
//                 with rt <= kg of the NEP x NEP matrix (NEP = ne rounded up
//                 to 16) are stored, row-major over the upper triangle
//                 (tile_index), each as [lane][s] so that lane l of a wave
//                 holds the A operand of v_mfma_*_16x16x4 for four
//                 consecutive k-steps: H[rt*16 + l%16][kg*16 + s*4 + l/16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <type_traits>

#include "common.h"
#include "sf_tables.h"

namespace idft {

namespace {

constexpr int kThreads = 256;

// Total of one value per thread over a work-group of NT threads (fixed
// order: each wave by shuffles, then the NT / 64 wave sums in order), returned
// to every thread.  red holds NT / 64 doubles; a caller that reuses red for a
// second total synchronises in between.
template <int NT>
__device__ inline double block_total(double t, double* red) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if ((tid & 63) == 0) red[tid >> 6] = t;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) s += red[w];
  return s;
}

// Sum of partial[b * B + j] over b < nblk by one work-group of NT threads
// (fixed order: strided per-thread sums, then block_total).
template <int NT = kThreads>
__device__ inline double block_sum_col(const double* __restrict__ partial, int nblk, int B, int j,
                                       double* red) {
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int b = tid; b < nblk; b += NT) s += partial[(int64_t)b * B + j];
  return block_total<NT>(s, red);
}

// Sum of n contiguous values by one work-group of NT threads (all threads
// call it).
template <int NT = kThreads>
__device__ inline double block_sum_contig(const double* __restrict__ v, int n, double* red) {
  const int tid = threadIdx.x;
  // four independent partial sums, so that four loads are in flight
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int b = tid;
  for (; b + 3 * NT < n; b += 4 * NT) {
    s0 += v[b];
    s1 += v[b + NT];
    s2 += v[b + 2 * NT];
    s3 += v[b + 3 * NT];
  }
  for (; b < n; b += NT) s0 += v[b];
  return block_total<NT>((s0 + s1) + (s2 + s3), red);
}

// Work-group size of the per-column reductions of long partial rows (cell
// dots, gather partials, Gram partials): 16 waves, about 12 values per thread
// for the 12-16 k values of a column.
constexpr int kRedThreads = 1024;

inline void check_u32(int64_t n) {
  if (n >= (int64_t)0xffffffffLL) throw Error(IDFT_EINVAL, "multivector too large for 32-bit indexing");
}

__host__ __device__ constexpr int ne_of(int p) { return (p + 1) * (p + 1) * (p + 1); }
__host__ __device__ constexpr int nep_of(int p) { return (ne_of(p) + 15) / 16 * 16; }
// stored tiles per cell and the position of tile (r, k), r <= k, among them
__host__ __device__ constexpr int ntiles_of(int p) { return (nep_of(p) / 16) * (nep_of(p) / 16 + 1) / 2; }
__host__ __device__ constexpr int tile_index(int rt_count, int r, int k) {
  return r * rt_count - r * (r - 1) / 2 + (k - r);
}

__device__ inline int64_t cell_node(const Geo& g, int64_t cx, int64_t cy, int64_t cz, int l) {
  const int a = l % g.n1, b = (l / g.n1) % g.n1, c = l / (g.n1 * g.n1);
  return ((cz * g.p + c) * g.N1 + (cy * g.p + b)) * (int64_t)g.N1 + (cx * g.p + a);
}

// Cell of colour `colour` handled by work-group w.  The 8 XCDs receive
// work-groups round-robin; each XCD is given a contiguous run of cells so its
// L2 sees spatially adjacent cells.
// The grid has 8 * ceil(nw / 8) work-groups, so (xcd, slot) covers every k.
__device__ inline int64_t colour_cell(const Geo& g, int colour, int64_t w, int64_t nw) {
  const int64_t per = (nw + 7) / 8;
  const int64_t xcd = w % 8, slot = w / 8;
  const int64_t k = xcd * per + slot;  // position inside the colour
  if (k >= nw) return -1;
  const int64_t hc = (g.nc + 1 - (colour & 1)) / 2;
  const int64_t hy = (g.nc + 1 - ((colour >> 1) & 1)) / 2;
  const int64_t kx = k % hc, ky = (k / hc) % hy, kz = k / (hc * hy);
  const int64_t cx = 2 * kx + (colour & 1), cy = 2 * ky + ((colour >> 1) & 1),
                cz = 2 * kz + ((colour >> 2) & 1);
  return (cz * g.nc + cy) * g.nc + cx;
}

inline unsigned colour_grid(int64_t nw) { return (unsigned)(8 * ((nw + 7) / 8)); }

inline int64_t colour_count(const Geo& g, int colour) {
  const int64_t hx = (g.nc + 1 - (colour & 1)) / 2;
  const int64_t hy = (g.nc + 1 - ((colour >> 1) & 1)) / 2;
  const int64_t hz = (g.nc + 1 - ((colour >> 2) & 1)) / 2;
  return hx * hy * hz;
}

// ------------------------------------------------------------ MFMA apply
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Output rows held by lane group lk (= lane / 16), item i, of the 16x16x4 MFMA:
// f32 keeps 4 consecutive rows per group, f64 interleaves them (measured on gfx950).
template <> struct Mfma<float> {
  typedef v4f acc_t;
  static __device__ inline int row(int lk, int i) { return 4 * lk + i; }
  static __device__ inline acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<double> {
  typedef v4d acc_t;
  static __device__ inline int row(int lk, int i) { return lk + 4 * i; }
  static __device__ inline acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

template <typename T> struct Vec4;
template <> struct Vec4<float> { typedef float4 type; };
template <> struct Vec4<double> { typedef double4 type; };

__host__ __device__ constexpr int ldb_of(int bt) { return (bt % 32 == 16) ? bt : bt + 16; }

// Stored tile of block row i, k-group k of H_e (Hc = the cell's stored
// tiles), one vector load per lane: tile (i, k) for k >= i, else its mirror
// (k, i), which the caller transposes (lds_transpose).
template <typename T, int RT>
__device__ inline typename Vec4<T>::type fetch_tile(const T* __restrict__ Hc, int i, int k,
                                                     int lane) {
  typedef typename Vec4<T>::type vec_t;
  const int t = k >= i ? tile_index(RT, i, k) : tile_index(RT, k, i);
  return reinterpret_cast<const vec_t*>(Hc)[t * 64 + lane];
}

// The transposed tile read straight from memory: lane l needs
// T[4s + l/16][l%16], which sits in lane (4s + l/16) + 16 (l%4), slot
// (l%16)/4 of the mirror tile (k, i) -- four scalar loads of the same 1 KiB
// (FP32) tile.  Used where the LDS square would cost occupancy.
template <typename T, int RT>
__device__ inline typename Vec4<T>::type fetch_tile_transposed(const T* __restrict__ Hc, int i,
                                                                int k, int lane) {
  typedef typename Vec4<T>::type vec_t;
  const T* t = Hc + tile_index(RT, k, i) * 256 + ((lane & 15) >> 2);
  const int lb = (lane >> 4) + 16 * (lane & 3);
  vec_t v;
  v.x = t[lb * 4];
  v.y = t[(lb + 4) * 4];
  v.z = t[(lb + 8) * 4];
  v.w = t[(lb + 12) * 4];
  return v;
}

// A operand of the transposed tile: lane l holds m = T[l%16][4s + l/16]
// (s = 0..3) and needs T[4s + l/16][l%16].  Through a per-wave 16 x 17
// (padded) LDS square: write row-major, read column-major.  LDS operations of
// one wave execute in order; the wavefront fences keep the compiler from
// moving the reads above the writes.
template <typename T>
__device__ inline typename Vec4<T>::type lds_transpose(typename Vec4<T>::type m, T* sq,
                                                        int lane) {
  typedef typename Vec4<T>::type vec_t;
  const int a = lane & 15, b = lane >> 4;
  sq[a * 17 + b] = m.x;
  sq[a * 17 + 4 + b] = m.y;
  sq[a * 17 + 8 + b] = m.z;
  sq[a * 17 + 12 + b] = m.w;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  vec_t t;
  t.x = sq[b * 17 + a];
  t.y = sq[(4 + b) * 17 + a];
  t.z = sq[(8 + b) * 17 + a];
  t.w = sq[(12 + b) * 17 + a];
  return t;
}

// The cell's input tile in LDS: xs[l][j] = S[n] xn[n] X[n][c0 + j] xcol[c0 + j]
// (node n of local index l; rows l >= NE and columns j >= bw are zero).
template <typename T, typename TV, int P, int BT, int LDB, int NTH = kThreads>
__device__ inline void load_cell_input(const Geo& g, int64_t cx, int64_t cy, int64_t cz,
                                       const T* __restrict__ S, const TV* __restrict__ X,
                                       const double* __restrict__ xn,
                                       const double* __restrict__ xcol, int B, int c0, int bw,
                                       T* xs) {
  constexpr int NE = ne_of(P), NEP = nep_of(P), TOT = NEP * BT;
  // U entries per thread per round, all loads of a round issued before the
  // first product: the optional scales are loaded inside uniform branches
  // that hold nothing else, so no product is pulled in after its own load
  constexpr int U = 4;
  for (int i0 = threadIdx.x; i0 < TOT; i0 += NTH * U) {
    int64_t n[U];
    int col[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * NTH;
      const int l = i / BT, j = i % BT;
      ok[u] = i < TOT && l < NE && j < bw;
      n[u] = ok[u] ? cell_node(g, cx, cy, cz, l) : 0;
      col[u] = ok[u] ? c0 + j : 0;
    }
    T sv[U];
    TV xv[U];
    double nv[U], cv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      sv[u] = S[n[u]];
      xv[u] = X[n[u] * B + col[u]];
    }
    if (xn) {
#pragma unroll
      for (int u = 0; u < U; ++u) nv[u] = xn[n[u]];
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) nv[u] = 1.0;
    }
    if (xcol) {
#pragma unroll
      for (int u = 0; u < U; ++u) cv[u] = xcol[col[u]];
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) cv[u] = 1.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * NTH;
      if (i < TOT) {
        T v = T(0);
        if (ok[u]) {
          if constexpr (std::is_same<T, TV>::value) {
            v = sv[u] * xv[u] * (T)nv[u] * (T)cv[u];
          } else {  // wider input vectors: scale in FP64, round once
            v = (T)((double)sv[u] * (double)xv[u] * nv[u] * cv[u]);
          }
        }
        xs[(i / BT) * LDB + i % BT] = v;
      }
    }
  }
}

// Phase 1 of the cell operator.  One work-group handles `cpw` consecutive
// cells and the column tile [blockIdx.y * 16 NCT, +16 NCT) of the block.  Per
// cell the tile of (S X)_e is gathered into LDS and each wave computes whole
// 16-row blocks of H_e (S X)_e, streaming the tiles of its block row from
// HBM (upper-triangle tiles as stored, lower-triangle ones as LDS transposes
// of their mirror tiles, which the mirror's own row pass reads too, so the
// second read is mostly served by the caches), KPF k-groups in flight per
// wave.
// Output: Yc[e][l][j] = (H_e (S X)_e)[l][j].  Work-groups are dealt to the 8
// XCDs in contiguous runs of cells.
template <typename T, typename TV, int P, int NCT>
__global__ __launch_bounds__(kThreads) void k_cell_gemm(Geo g, const T* __restrict__ H,
                                                        const T* __restrict__ S,
                                                        const TV* __restrict__ X,
                                                        const double* __restrict__ xn,
                                                        const double* __restrict__ xcol,
                                                        T* __restrict__ Yc, int B, int cpw) {
  constexpr int NE = ne_of(P), NEP = nep_of(P), RT = NEP / 16, BT = NCT * 16;
  constexpr int NTILE = ntiles_of(P);
  constexpr int LDB = ldb_of(BT);
  // k-groups of H_e in flight per wave (deeper prefetch costs occupancy and
  // measured slower at p = 4, 5)
  constexpr int KPF = 3;
  // lower-triangle tiles through per-wave LDS transpose squares, unless
  // these would push the work-group past half the 160 KiB LDS (two
  // work-groups per CU), as for FP64 with 80-column tiles
  constexpr bool kLdsT = (NEP * LDB + 4 * 272) * (int)sizeof(T) <= 81920 ||
                         NEP * LDB * (int)sizeof(T) > 81920;
  __shared__ T xs[NEP * LDB];
  __shared__ T tsq[kLdsT ? 4 : 1][kLdsT ? 16 * 17 : 1];
  typedef typename Vec4<T>::type vec_t;
  // stored tile (transposed by the caller) or the transposed tile itself
  auto fetch = [](auto ldsT, const T* Hc, int i, int k, int lane) -> vec_t {
    if constexpr (decltype(ldsT)::value) return fetch_tile<T, RT>(Hc, i, k, lane);
    else return k >= i ? fetch_tile<T, RT>(Hc, i, k, lane)
                       : fetch_tile_transposed<T, RT>(Hc, i, k, lane);
  };
  const int64_t nwg = (g.ncell + cpw - 1) / cpw;
  const int64_t per = (nwg + 7) / 8;
  const int64_t wg = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (wg >= nwg) return;
  const int c0 = blockIdx.y * BT;
  const int bw = min(BT, B - c0);  // columns of this tile
  const int tid = threadIdx.x;
  const int wave = tid / 64, lane = tid % 64;
  const int lr = lane % 16, lk = lane / 16;
  typedef typename Mfma<T>::acc_t acc_t;
  const int64_t e0 = wg * cpw, e1 = min((int64_t)g.ncell, e0 + cpw);
  for (int64_t e = e0; e < e1; ++e) {
    const int64_t cx = e % g.nc, cy = (e / g.nc) % g.nc, cz = e / ((int64_t)g.nc * g.nc);
    if (e > e0) __syncthreads();  // previous cell's tiles fully consumed
    load_cell_input<T, TV, P, BT, LDB>(g, cx, cy, cz, S, X, xn, xcol, B, c0, bw, xs);
    __syncthreads();
    const T* Hc = H + (size_t)e * NTILE * 256;
    T* yc = Yc + (size_t)e * NE * B + c0;
    for (int rt = wave; rt < RT; rt += 4) {
      acc_t acc[NCT];
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) acc[ct] = acc_t{0, 0, 0, 0};
      vec_t buf[KPF];
#pragma unroll
      for (int kk = 0; kk < KPF; ++kk)
        buf[kk] = kk < RT ? fetch(std::integral_constant<bool, kLdsT>{}, Hc, rt, kk, lane) : vec_t{};
#pragma unroll 1
      for (int kg0 = 0; kg0 < RT; kg0 += KPF) {
#pragma unroll
        for (int kk = 0; kk < KPF; ++kk) {
          const int kg = kg0 + kk;
          if (kg < RT) {
            vec_t a = buf[kk];
            if (kg + KPF < RT) buf[kk] = fetch(std::integral_constant<bool, kLdsT>{}, Hc, rt, kg + KPF, lane);
            if constexpr (kLdsT) {
              if (kg < rt) a = lds_transpose<T>(a, tsq[wave], lane);
            }
            const T* xk = xs + (kg * 16 + lk) * LDB + lr;
            const T av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
#pragma unroll
              for (int ct = 0; ct < NCT; ++ct)
                acc[ct] = Mfma<T>::run(av[s], xk[s * 4 * LDB + ct * 16], acc[ct]);
            }
          }
        }
      }
      // lane holds rows rt*16 + Mfma<T>::row(lk, i), column ct*16 + lr
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        const int col = ct * 16 + lr;
        if (col >= bw) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = rt * 16 + Mfma<T>::row(lk, i);
          if (r < NE) yc[(size_t)r * B + col] = acc[ct][i];
        }
      }
    }
  }
}

// Compile-time row / column of stored tile t (row-major upper triangle).
__host__ __device__ constexpr int tile_row(int rt_count, int t) {
  int r = 0;
  while (r + 1 < rt_count && tile_index(rt_count, r + 1, r + 1) <= t) ++r;
  return r;
}
__host__ __device__ constexpr int tile_col(int rt_count, int t) {
  return tile_row(rt_count, t) + (t - tile_index(rt_count, tile_row(rt_count, t),
                                                 tile_row(rt_count, t)));
}

template <int I>
using ic = std::integral_constant<int, I>;
template <int B0, int E0, typename F>
__device__ inline void static_for(F&& f) {
  if constexpr (B0 < E0) {
    f(ic<B0>{});
    static_for<B0 + 1, E0>(f);
  }
}

// Tiles of the stored upper triangle that touch block rows [r0, r1): the
// j-th (in storage order) of those with r or k in the range, and their count.
__host__ __device__ constexpr int half_tile_count(int rt_count, int r0, int r1) {
  int n = 0;
  for (int t = 0; t < rt_count * (rt_count + 1) / 2; ++t) {
    const int r = tile_row(rt_count, t), k = tile_col(rt_count, t);
    if ((r >= r0 && r < r1) || (k >= r0 && k < r1)) ++n;
  }
  return n;
}
__host__ __device__ constexpr int half_tile(int rt_count, int r0, int r1, int j) {
  for (int t = 0; t < rt_count * (rt_count + 1) / 2; ++t) {
    const int r = tile_row(rt_count, t), k = tile_col(rt_count, t);
    if ((r >= r0 && r < r1) || (k >= r0 && k < r1)) {
      if (j == 0) return t;
      --j;
    }
  }
  return -1;
}

constexpr int kSymmThreads = 512;

// Phase 1, tile pass (8 waves per work-group).  The block rows are split in
// two halves; waves 0-3 produce the rows of the first half, waves 4-7 those
// of the second.  A half's four waves split, in compile-time lists (every
// accumulator index static), the stored tiles (r, k) that touch the half:
// T (S X)_k is added to row r if r is in the half and, for r < k,
// T^T (S X)_r (LDS transpose) to row k if k is in the half.  A tile is read
// once per half it touches -- the two halves read the shared ones at the same
// time on the same CU, so the second read is a cache hit -- and each wave
// keeps partial sums of its half's rows only.  The four partials of each half
// are added in LDS in wave order (deterministic) and written to Yc.
template <typename T, typename TV, int P, int NCT>
__global__ __launch_bounds__(kSymmThreads) void k_cell_symm(Geo g, const T* __restrict__ H,
                                                            const T* __restrict__ S,
                                                            const TV* __restrict__ X,
                                                            const double* __restrict__ xn,
                                                            const double* __restrict__ xcol,
                                                            T* __restrict__ Yc, int B, int cpw) {
  constexpr int NE = ne_of(P), NEP = nep_of(P), RT = NEP / 16, BT = NCT * 16;
  constexpr int NTILE = ntiles_of(P);
  constexpr int LDB = ldb_of(BT);
  constexpr int KPF = 3;
  constexpr int RH = (RT + 1) / 2;  // block rows of the first half (and at most per half)
  __shared__ T xs[NEP * LDB];  // the input tile, then the partial-sum buffer
  __shared__ T tsq[8][16 * 17];
  typedef typename Vec4<T>::type vec_t;
  typedef typename Mfma<T>::acc_t acc_t;
  const int64_t nwg = (g.ncell + cpw - 1) / cpw;
  const int64_t per = (nwg + 7) / 8;
  const int64_t wg = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (wg >= nwg) return;
  const int c0 = blockIdx.y * BT;
  const int bw = min(BT, B - c0);
  const int tid = threadIdx.x;
  const int wave = tid / 64, lane = tid % 64;
  const int lr = lane % 16, lk = lane / 16;
  const int64_t e0 = wg * cpw, e1 = min((int64_t)g.ncell, e0 + cpw);
  for (int64_t e = e0; e < e1; ++e) {
    const int64_t cx = e % g.nc, cy = (e / g.nc) % g.nc, cz = e / ((int64_t)g.nc * g.nc);
    if (e > e0) __syncthreads();  // previous cell's output read from xs
    load_cell_input<T, TV, P, BT, LDB, kSymmThreads>(g, cx, cy, cz, S, X, xn, xcol, B, c0, bw,
                                                     xs);
    __syncthreads();
    const vec_t* Hc = reinterpret_cast<const vec_t*>(H + (size_t)e * NTILE * 256) + lane;
    acc_t acc[RH][NCT];
#pragma unroll
    for (int r = 0; r < RH; ++r)
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) acc[r][ct] = acc_t{0, 0, 0, 0};
    // tiles [J0, J1) of the list of half [R0, R1)
    auto run = [&](auto r0c, auto r1c, auto j0c, auto j1c) {
      constexpr int R0 = decltype(r0c)::value, R1 = decltype(r1c)::value;
      constexpr int J0 = decltype(j0c)::value, CNT = decltype(j1c)::value - J0;
      vec_t buf[KPF];
      static_for<0, (KPF < CNT ? KPF : CNT)>([&](auto jc) {
        constexpr int tp = half_tile(RT, R0, R1, J0 + decltype(jc)::value);
        buf[decltype(jc)::value] = Hc[tp * 64];
      });
      static_for<0, CNT>([&](auto jc) {
        constexpr int J = decltype(jc)::value, t = half_tile(RT, R0, R1, J0 + J);
        constexpr int r = tile_row(RT, t), k = tile_col(RT, t);
        const vec_t a = buf[J % KPF];
        if constexpr (J + KPF < CNT) {
          constexpr int tn = half_tile(RT, R0, R1, J0 + J + KPF);
          buf[J % KPF] = Hc[tn * 64];
        }
        if constexpr (r >= R0 && r < R1) {  // row r += T (S X)_k
          const T* xk = xs + (k * 16 + lk) * LDB + lr;
          const T av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct)
              acc[r - R0][ct] = Mfma<T>::run(av[s4], xk[s4 * 4 * LDB + ct * 16], acc[r - R0][ct]);
        }
        if constexpr (r < k && k >= R0 && k < R1) {  // row k += T^T (S X)_r
          const vec_t at = lds_transpose<T>(a, tsq[wave], lane);
          const T* xr = xs + (r * 16 + lk) * LDB + lr;
          const T av[4] = {at.x, at.y, at.z, at.w};
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct)
              acc[k - R0][ct] = Mfma<T>::run(av[s4], xr[s4 * 4 * LDB + ct * 16], acc[k - R0][ct]);
        }
      });
    };
    constexpr int C0 = half_tile_count(RT, 0, RH), C1 = half_tile_count(RT, RH, RT);
    switch (wave) {
      case 0: run(ic<0>{}, ic<RH>{}, ic<0>{}, ic<C0 / 4>{}); break;
      case 1: run(ic<0>{}, ic<RH>{}, ic<C0 / 4>{}, ic<C0 / 2>{}); break;
      case 2: run(ic<0>{}, ic<RH>{}, ic<C0 / 2>{}, ic<(3 * C0) / 4>{}); break;
      case 3: run(ic<0>{}, ic<RH>{}, ic<(3 * C0) / 4>{}, ic<C0>{}); break;
      case 4: run(ic<RH>{}, ic<RT>{}, ic<0>{}, ic<C1 / 4>{}); break;
      case 5: run(ic<RH>{}, ic<RT>{}, ic<C1 / 4>{}, ic<C1 / 2>{}); break;
      case 6: run(ic<RH>{}, ic<RT>{}, ic<C1 / 2>{}, ic<(3 * C1) / 4>{}); break;
      default: run(ic<RH>{}, ic<RT>{}, ic<(3 * C1) / 4>{}, ic<C1>{}); break;
    }
    // partial sums of each half's four waves, added in wave order through xs
    // (the halves own disjoint rows, so they reduce side by side)
    const int r0 = wave < 4 ? 0 : RH, nr = wave < 4 ? RH : RT - RH;
    for (int w = 0; w < 4; ++w) {
      __syncthreads();
      if (wave % 4 == w) {
#pragma unroll
        for (int r = 0; r < RH; ++r) {
          if (r >= nr) break;
#pragma unroll
          for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              T* p = xs + ((r0 + r) * 16 + Mfma<T>::row(lk, i)) * LDB + ct * 16 + lr;
              *p = w == 0 ? acc[r][ct][i] : *p + acc[r][ct][i];
            }
        }
      }
    }
    __syncthreads();
    T* yc = Yc + (size_t)e * NE * B + c0;
    for (int i = tid; i < NE * BT; i += kSymmThreads) {
      const int l = i / BT, j = i % BT;
      if (j < bw) yc[(size_t)l * B + j] = xs[l * LDB + j];
    }
  }
}

// ----------------------------------------------- sum-factorised cell product
// Phase 1 without stored cell matrices.  On the uniform mesh
//   H_e = (h/4)(D(x)M(x)M + M(x)D(x)M + M(x)M(x)D) + L^T diag(w_e) L,  L = Lq(x)Lq(x)Lq,
// so H_e X_e is a chain of 1D contractions along x, y and z, each in the
// even-odd form below: 11.9 k FMA and 6.6 k adds per column at p = 5 (23 k FMA
// without the even-odd split, 47 k for the dense product), and per cell only
// the (p+3)^3 weights w_e are read from HBM instead of n_e(n_e+1)/2 entries of
// H_e.  The 1D tables are compile-time constants (SfC<P>): after unrolling
// every FP32 coefficient is an immediate operand, and the FP64 ones come from
// a constant-memory table (eo_coef below); h/4 scales the kinetic sum once.
//
// One work-group handles one cell and CG columns; an item is one line of the
// cell's tensor along one axis for VC consecutive columns (VC-wide LDS
// accesses).  Intermediates live in LDS as [entry][CG / VC] packs, with padded
// line strides (N1 and Q rounded up to odd: 5, 7 at p = 4; 7, 9 at p = 5) so
// that the 64 lanes of a wave spread over the banks (an even stride such as 6
// or 8 at p = 4 folds the lines of a wave onto a few banks).  The exception is
// region D in FP64 at p = 5 (Q = 8): unpadded lines, entry x of line (c, y)
// at slot x ^ y, which spreads the lanes over the banks as the stride 9 does,
// and puts the lines that one stage-6 item reads 512 B apart, so that the
// compiler pairs those accesses (ds_read2st64 / ds_write2st64).  It also
// takes 20.4 KB instead of 21.9 KB per work-group, which does not raise the
// occupancy: the kernel's 71 registers allow 7 waves per SIMD either way (with
// the stage-6 weights loaded after stage 3 it fits 64 registers and 8
// work-groups per CU, but measured slower: the weights' latency shows).  The kinetic
// chain runs first and the x-line input stays in registers for the potential
// chain, and the y and z stages overwrite their input in place (each item
// reads its whole line before writing the same positions; in stage 5 the
// item (c, qx) owns the column [c][*][qx] of D, which holds t1 along b and
// then t2 along qy).  Region Bk is dead once stage 3 has run, so D shares its
// memory:
//   1  x lines (c,b):  ka = M x -> A,  kb = D x -> Bk      (x gathered from HBM)
//   2  y lines (c,a):  k1 = M ka -> A,  k2 = D ka + M kb -> Bk
//   3  z lines (b,a):  yk = D k1 + M k2 -> A
//   4  x lines (c,b):  t1[c][b][qx] = Lq x -> D             (x from registers)
//   5  y lines (c,qx): t2[c][qy][qx] = Lq t1 -> D
//   6  z lines (qy,qx): s2 = Lq^T (w .* (Lq t2)) -> D
//   7  y lines (c,qx): s1[c][b][qx] = Lq^T s2 -> D
//   8  x lines (c,b):  Yc[e][(c,b,a)][j] = (Lq^T s1)[a] + (h/4) yk[a]
// At p = 5 that is 252 + 432 = 684 entries per column, 252 + 384 = 636 with
// the unpadded FP64 region D.
__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <typename T, int VC>
struct alignas(sizeof(T) * VC) Pack {
  T v[VC];
};
template <typename T, int VC>
__device__ inline void pack_zero(Pack<T, VC>& p) {
#pragma unroll
  for (int u = 0; u < VC; ++u) p.v[u] = T(0);
}
template <typename T, int VC>
__device__ inline void pack_fma(Pack<T, VC>& acc, T c, const Pack<T, VC>& x) {
#pragma unroll
  for (int u = 0; u < VC; ++u) acc.v[u] += c * x.v[u];
}

// The 1D tables as operators A (R x C): K = 0 Lq (Gauss points from nodes),
// 1 Lq^T, 2 M, 3 D.  The GLL nodes and the Gauss points lie symmetrically
// about the cell centre, so each A is centro-symmetric,
// A[R-1-i][C-1-j] = A[i][j] (to rounding of the tables).
template <int P, int K>
struct SfTab {
  static constexpr int kP = P, kK = K;
  static constexpr int N1 = P + 1, Q = P + 3;
  static constexpr int R = K == 0 ? Q : N1;
  static constexpr int C = K == 1 ? Q : N1;
  __host__ __device__ static constexpr double at(int i, int j) {
    return K == 0   ? SfC<P>::v.L[i * N1 + j]
           : K == 1 ? SfC<P>::v.L[j * N1 + i]
           : K == 2 ? SfC<P>::v.M[i * N1 + j]
                    : SfC<P>::v.D[i * N1 + j];
  }
};
// Even-odd form of y = A x (Kopriva's even-odd decomposition).  With
// x+_j = x_j + x_{C-1-j} and x-_j = x_j - x_{C-1-j} for j < C/2 (the middle
// entry of an odd C kept alone as x+_{C/2}), row i < R/2 takes
//   e_i = sum_j (A_ij + A_i,C-1-j)/2 x+_j [+ A_i,mid x_mid]
//   o_i = sum_j (A_ij - A_i,C-1-j)/2 x-_j
// and y_i = e_i + o_i, y_{R-1-i} = e_i - o_i; the middle row of an odd R is
// e alone.  R C / 2 multiply-adds instead of R C.  The coefficients are
// constant expressions, so each is an immediate operand after unrolling.
template <typename T, class A, int I, int J>
__device__ constexpr T eo_even() {
  constexpr double c = 0.5 * (A::at(I, J) + A::at(I, A::C - 1 - J));
  return (T)c;
}
template <typename T, class A, int I, int J>
__device__ constexpr T eo_odd() {
  constexpr double c = 0.5 * (A::at(I, J) - A::at(I, A::C - 1 - J));
  return (T)c;
}
template <typename T, class A, int I>
__device__ constexpr T eo_mid() {
  constexpr double c = A::at(I, A::C / 2);
  return (T)c;
}
// FP64 coefficients from constant memory.  A double has no literal operand,
// so an immediate FP64 coefficient costs two scalar moves per use; the FP64
// kernels instead read the even-odd coefficients of all four operators from
// one table in constant memory, which the compiler fetches with wide scalar
// loads (4 doubles per load).  Per operator K (R x C) and row I < ceil(R/2):
// the floor(C/2) even coefficients, the floor(C/2) odd ones, then the middle
// one if C is odd.
template <int P>
struct EoTable {
  static constexpr int R(int k) { return k == 0 ? P + 3 : P + 1; }
  static constexpr int C(int k) { return k == 1 ? P + 3 : P + 1; }
  static constexpr int stride(int k) { return 2 * (C(k) / 2) + C(k) % 2; }
  static constexpr int base(int k) {
    int b = 0;
    for (int j = 0; j < k; ++j) b += ((R(j) + 1) / 2) * stride(j);
    return b;
  }
  static constexpr int kSize = base(4);
  double v[kSize];
};
template <int P>
constexpr EoTable<P> make_eo_table() {
  EoTable<P> t{};
  for (int k = 0; k < 4; ++k) {
    const int r = EoTable<P>::R(k), c = EoTable<P>::C(k), st = EoTable<P>::stride(k);
    for (int i = 0; i < (r + 1) / 2; ++i) {
      double* row = t.v + EoTable<P>::base(k) + i * st;
      auto at = [&](int ii, int jj) {
        return k == 0   ? SfC<P>::v.L[ii * (P + 1) + jj]
               : k == 1 ? SfC<P>::v.L[jj * (P + 1) + ii]
               : k == 2 ? SfC<P>::v.M[ii * (P + 1) + jj]
                        : SfC<P>::v.D[ii * (P + 1) + jj];
      };
      for (int j = 0; j < c / 2; ++j) {
        row[j] = 0.5 * (at(i, j) + at(i, c - 1 - j));
        row[c / 2 + j] = 0.5 * (at(i, j) - at(i, c - 1 - j));
      }
      if (c % 2) row[2 * (c / 2)] = at(i, c / 2);
    }
  }
  return t;
}
// Written from the host once per device (upload_sf_constants): with an
// initialiser the compiler would fold the values back into immediates.  The
// kernels take the table's device address as an argument (sf_eo_table): an
// access through the symbol itself goes through the global offset table, one
// more dependent scalar load wherever the compiler reloads the coefficients.
__constant__ EoTable<2> kEo2;
__constant__ EoTable<3> kEo3;
__constant__ EoTable<4> kEo4;
__constant__ EoTable<5> kEo5;
// Coefficient W (0 even, 1 odd, 2 middle) of row I, column J of operator A:
// an immediate in FP32, a scalar load from the table eo (EoTable<P>::v) in FP64.
template <typename T, class A, int I, int J, int W>
__device__ inline T eo_coef(const double* __restrict__ eo) {
  if constexpr (std::is_same<T, double>::value) {
    using E = EoTable<A::kP>;
    constexpr int c = E::C(A::kK);
    constexpr int idx = E::base(A::kK) + I * E::stride(A::kK) +
                        (W == 0 ? J : W == 1 ? c / 2 + J : 2 * (c / 2));
    return eo[idx];
  } else if constexpr (W == 0) {
    return eo_even<T, A, I, J>();
  } else if constexpr (W == 1) {
    return eo_odd<T, A, I, J>();
  } else {
    return eo_mid<T, A, I>();
  }
}
template <int C, typename T, int VC>
__device__ inline void eo_split(const Pack<T, VC> (&x)[C], Pack<T, VC> (&xp)[(C + 1) / 2],
                                Pack<T, VC> (&xm)[C / 2]) {
#pragma unroll
  for (int j = 0; j < C / 2; ++j)
#pragma unroll
    for (int u = 0; u < VC; ++u) {
      xp[j].v[u] = x[j].v[u] + x[C - 1 - j].v[u];
      xm[j].v[u] = x[j].v[u] - x[C - 1 - j].v[u];
    }
  if constexpr (C % 2 == 1) xp[C / 2] = x[C / 2];
}
// Adds row I's even and odd sums of A applied to the split input (xp, xm).
template <typename T, class A, int I, int VC>
__device__ inline void eo_row(const Pack<T, VC> (&xp)[(A::C + 1) / 2],
                              const Pack<T, VC> (&xm)[A::C / 2], Pack<T, VC>& e,
                              Pack<T, VC>& o, const double* __restrict__ eo) {
  static_for<0, A::C / 2>([&](auto j) {
    constexpr int J = decltype(j)::value;
    pack_fma(e, eo_coef<T, A, I, J, 0>(eo), xp[J]);
    if constexpr (2 * I + 1 != A::R) pack_fma(o, eo_coef<T, A, I, J, 1>(eo), xm[J]);
  });
  if constexpr (A::C % 2 == 1) pack_fma(e, eo_coef<T, A, I, 0, 2>(eo), xp[A::C / 2]);
}
// The R outputs of one or more even-odd contractions: rows(i, e, o) adds the
// sums of row i, put(i, y) receives output i (i an integral constant).
template <int R, typename T, int VC, typename Rows, typename Put>
__device__ inline void eo_out(Rows&& rows, Put&& put) {
  static_for<0, (R + 1) / 2>([&](auto i) {
    constexpr int I = decltype(i)::value;
    Pack<T, VC> e, o;
    pack_zero(e);
    pack_zero(o);
    rows(i, e, o);
    if constexpr (2 * I + 1 == R) {
      put(i, e);
    } else {
      Pack<T, VC> y0, y1;
#pragma unroll
      for (int u = 0; u < VC; ++u) {
        y0.v[u] = e.v[u] + o.v[u];
        y1.v[u] = e.v[u] - o.v[u];
      }
      put(i, y0);
      put(ic<R - 1 - I>{}, y1);
    }
  });
}
#define IDFT_IC(x) decltype(x)::value
// Entry A of a line from its even-odd split (eo_split): the inverse of the
// split, to rounding.
template <int C, int A, typename T, int VC>
__device__ inline T eo_unsplit(const Pack<T, VC> (&xp)[(C + 1) / 2], const Pack<T, VC> (&xm)[C / 2],
                               int u) {
  if constexpr (2 * A + 1 == C) {
    return xp[A].v[u];
  } else if constexpr (A < C / 2) {
    return T(0.5) * (xp[A].v[u] + xm[A].v[u]);
  } else {
    constexpr int J = C - 1 - A;
    return T(0.5) * (xp[J].v[u] - xm[J].v[u]);
  }
}

template <typename T, typename TV, int P, int CG, int VC, bool XN>
__global__ __launch_bounds__(kThreads) void k_cell_sf(Geo g, T hq, const T* __restrict__ W,
                                                      const T* __restrict__ S,
                                                      const TV* __restrict__ X,
                                                      const double* __restrict__ xn,
                                                      const double* __restrict__ xcol,
                                                      T* __restrict__ Yc, int B,
                                                      double* __restrict__ dotp,
                                                      const double* __restrict__ eo) {
  constexpr int N1 = P + 1, Q = P + 3, N2 = N1 * N1, NE = N2 * N1, Q2 = Q * Q;
  constexpr int QP = Q | 1, N1P = N1 | 1, NV = CG / VC;
  // x-line items per thread (lines x column packs over the work-group)
  constexpr int KI = (N2 * NV + kThreads - 1) / kThreads;
  typedef Pack<T, VC> pk;
  typedef SfTab<P, 0> TL;
  typedef SfTab<P, 1> TLt;
  typedef SfTab<P, 2> TM;
  typedef SfTab<P, 3> TD;
  // regions, each [entry][NV] packs:
  //   A  [(c,b)][N1P]      ka -> k1 -> yk            (kinetic chain, in place)
  //   Bk [(c,b)][N1P]      kb -> k2                  (stages 1-3)
  //   D  [c][Q][QS]        t1 -> t2 -> s2 -> s1      (stages 4-8, over Bk)
  // SWZ: D unpadded (QS = Q = 8), entry (c, y, x) at slot x ^ y of its line
  constexpr bool SWZ = sizeof(T) == 8 && Q == 8;
  constexpr int QS = SWZ ? Q : QP;
  __shared__ __align__(16) pk rA[N2 * N1P * NV];
  __shared__ __align__(16) pk rBD[cmax(N2 * N1P, N1 * Q * QS) * NV];
  pk* const rB = rBD;
  // Byte offsets into D of entry (c, y, x), column pack jv, from a per-item
  // base and the one index that is a compile-time constant at the access
  // (so that the rest folds into the LDS instruction's offset field):
  //   dy(c, x, jv) then at_y(base, y);  dx(c, y, jv) then at_x(base, x);
  //   dc(y, x, jv) then at_c(base, c).
  // With SWZ the slot x ^ y is an XOR of the offset's slot bits (the slot
  // field, NV sizeof(pk) bytes per slot and 8 slots, holds no other bits).
  constexpr uint32_t SB = NV * sizeof(pk);
  const auto dy = [](uint32_t c, uint32_t x, uint32_t jv) -> uint32_t {
    return (c * Q * QS + x) * SB + jv * (uint32_t)sizeof(pk);
  };
  const auto at_y = [](uint32_t base, uint32_t y) -> uint32_t {
    return SWZ ? (base ^ (y * SB)) + y * QS * SB : base + y * QS * SB;
  };
  const auto dx = [](uint32_t c, uint32_t y, uint32_t jv) -> uint32_t {
    return ((c * Q + y) * QS + (SWZ ? y : 0u)) * SB + jv * (uint32_t)sizeof(pk);
  };
  const auto at_x = [](uint32_t base, uint32_t x) -> uint32_t {
    return SWZ ? base ^ (x * SB) : base + x * SB;
  };
  const auto dc = [](uint32_t y, uint32_t x, uint32_t jv) -> uint32_t {
    return (y * QS + (SWZ ? (x ^ y) : x)) * SB + jv * (uint32_t)sizeof(pk);
  };
  const auto at_c = [](uint32_t base, uint32_t c) -> uint32_t { return base + c * Q * QS * SB; };
  char* const rDb = reinterpret_cast<char*>(rBD);
  const auto dref = [rDb](uint32_t off) -> pk& { return *reinterpret_cast<pk*>(rDb + off); };
  // (cell, column group) pairs dealt to the 8 XCDs in contiguous runs, the
  // groups of one cell at consecutive slots of one XCD, so that they share
  // that XCD's L2 for the cell's rows of X and of the cell scratch (each group
  // holds CG of the B columns of a row).  Every index below fits in 32 bits
  // (the host checks N B and cells n_e B against 2^32).
  const uint32_t ncell = (uint32_t)g.ncell, nc = (uint32_t)g.nc, np = (uint32_t)g.p;
  const uint32_t NN = (uint32_t)g.N1, ub = (uint32_t)B;
  const uint32_t ngrp = (ub + CG - 1) / CG, npair = ncell * ngrp;
  const uint32_t per = (npair + 7) / 8;
  const uint32_t k = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (k >= npair) return;
  const uint32_t e = k / ngrp;
  const int c0 = (int)(k % ngrp) * CG;
  const int bw = min(CG, B - c0);
  const int tid = threadIdx.x;
  const uint32_t cx = e % nc, cy = (e / nc) % nc, cz = e / (nc * nc);
  const T* we = W + e * (uint32_t)(Q2 * Q);
  // the column scales of this thread's columns (every item of a thread has
  // column pack jv = tid % NV, since kThreads is a multiple of NV), loaded
  // first: a later load would make their first use wait for the weights too
  double xcl[VC];
#pragma unroll
  for (int u = 0; u < VC; ++u) {
    const int j = min((tid % NV) * VC + u, bw - 1);
    xcl[u] = xcol ? xcol[c0 + j] : 1.0;
  }
  // the quadrature weights of this thread's stage-6 items, loaded now so that
  // their latency hides behind stages 1-5 (item it6 = tid + k6 kThreads takes
  // w_e[qz][ln6] for its line ln6 = it6 / NV)
  constexpr int K6 = (Q2 * NV + kThreads - 1) / kThreads;
  T wpre[K6][Q];
#pragma unroll
  for (int k6 = 0; k6 < K6; ++k6) {
    const int it6 = tid + k6 * kThreads;
    const int ln6 = it6 < Q2 * NV ? it6 / NV : 0;
#pragma unroll
    for (int qz = 0; qz < Q; ++qz) wpre[k6][qz] = we[qz * Q2 + ln6];
  }
  // the input lines stay in registers from stage 1 to stage 4 (even-odd split)
  pk keep_p[KI][(N1 + 1) / 2], keep_m[KI][N1 / 2];
  // stage 1: x lines ln = c*N1 + b of the gathered input -> ka = M x, kb = D x
#pragma unroll
  for (int ki = 0; ki < KI; ++ki) {
    const int it = tid + ki * kThreads;
    if (it >= N2 * NV) break;
    const int jv = it % NV, ln = it / NV;
    pk v[N1];
    const uint32_t b = ln % N1, c = ln / N1;
    const uint32_t n0 = ((cz * np + c) * NN + (cy * np + b)) * NN + cx * np;
    // the line's N1 nodes are consecutive: one base pointer per array, the
    // nodes at immediate offsets (the X rows B apart).  All loads of the line
    // are issued before the first product (the node scale xn is a
    // compile-time choice, so no branch separates them).
    const T* Sp = S + n0;
#pragma unroll
    for (int u = 0; u < VC; ++u) {
      const int j = jv * VC + u;
      if (j < bw) {
        const TV* Xp = X + (size_t)n0 * ub + (uint32_t)(c0 + j);
        const double xc = xcl[u];
        T sv[N1];
        TV xv[N1];
        double nv[N1];
#pragma unroll
        for (int a = 0; a < N1; ++a) {
          sv[a] = Sp[a];
          xv[a] = Xp[(size_t)a * ub];
          if constexpr (XN) nv[a] = xn[n0 + a];
          else nv[a] = 1.0;
        }
#pragma unroll
        for (int a = 0; a < N1; ++a) {
          if constexpr (std::is_same<T, TV>::value) {
            T x = sv[a] * xv[a];
            if constexpr (XN) x *= (T)nv[a];
            if (xcol) x *= (T)xc;
            v[a].v[u] = x;
          } else {  // wider input vectors: scale in FP64, round once
            double x = (double)sv[a] * (double)xv[a];
            if constexpr (XN) x *= nv[a];
            if (xcol) x *= xc;
            v[a].v[u] = (T)x;
          }
        }
      } else {
#pragma unroll
        for (int a = 0; a < N1; ++a) v[a].v[u] = T(0);
      }
    }
    eo_split(v, keep_p[ki], keep_m[ki]);
    eo_out<N1, T, VC>(
        [&](auto i, pk& s, pk& d) { eo_row<T, TM, IDFT_IC(i)>(keep_p[ki], keep_m[ki], s, d, eo); },
        [&](auto a2, const pk& r) { rA[(ln * N1P + IDFT_IC(a2)) * NV + jv] = r; });
    eo_out<N1, T, VC>(
        [&](auto i, pk& s, pk& d) { eo_row<T, TD, IDFT_IC(i)>(keep_p[ki], keep_m[ki], s, d, eo); },
        [&](auto a2, const pk& r) { rB[(ln * N1P + IDFT_IC(a2)) * NV + jv] = r; });
  }
  __syncthreads();
  // stage 2: y lines (c, a) of the kinetic chain, in place:
  // k1 = M ka (over ka), k2 = D ka + M kb (over kb)
  for (int it = tid; it < N2 * NV; it += kThreads) {
    const int jv = it % NV, l2 = it / NV;
    const int a = l2 % N1, c = l2 / N1;
    pk va[N1], vb[N1], ap[(N1 + 1) / 2], am[N1 / 2], bp[(N1 + 1) / 2], bm[N1 / 2];
#pragma unroll
    for (int b = 0; b < N1; ++b) {
      va[b] = rA[((c * N1 + b) * N1P + a) * NV + jv];
      vb[b] = rB[((c * N1 + b) * N1P + a) * NV + jv];
    }
    eo_split(va, ap, am);
    eo_split(vb, bp, bm);
    eo_out<N1, T, VC>(
        [&](auto i, pk& s, pk& d) { eo_row<T, TM, IDFT_IC(i)>(ap, am, s, d, eo); },
        [&](auto b2, const pk& r) { rA[((c * N1 + IDFT_IC(b2)) * N1P + a) * NV + jv] = r; });
    eo_out<N1, T, VC>(
        [&](auto i, pk& s, pk& d) {
          eo_row<T, TD, IDFT_IC(i)>(ap, am, s, d, eo);
          eo_row<T, TM, IDFT_IC(i)>(bp, bm, s, d, eo);
        },
        [&](auto b2, const pk& r) { rB[((c * N1 + IDFT_IC(b2)) * N1P + a) * NV + jv] = r; });
  }
  __syncthreads();
  // stage 3: z lines (b, a) of the kinetic chain: yk = D k1 + M k2 (over k1)
  for (int it = tid; it < N2 * NV; it += kThreads) {
    const int jv = it % NV, l2 = it / NV;  // l2 = b * N1 + a
    const int a = l2 % N1, b = l2 / N1;
    pk v1[N1], v2[N1], p1[(N1 + 1) / 2], m1[N1 / 2], p2[(N1 + 1) / 2], m2[N1 / 2];
#pragma unroll
    for (int c = 0; c < N1; ++c) {
      v1[c] = rA[((c * N1 + b) * N1P + a) * NV + jv];
      v2[c] = rB[((c * N1 + b) * N1P + a) * NV + jv];
    }
    eo_split(v1, p1, m1);
    eo_split(v2, p2, m2);
    eo_out<N1, T, VC>(
        [&](auto i, pk& s, pk& d) {
          eo_row<T, TD, IDFT_IC(i)>(p1, m1, s, d, eo);
          eo_row<T, TM, IDFT_IC(i)>(p2, m2, s, d, eo);
        },
        [&](auto c2, const pk& r) { rA[((IDFT_IC(c2) * N1 + b) * N1P + a) * NV + jv] = r; });
  }
  __syncthreads();
  // stage 4: x lines of the potential chain from the kept input:
  // t1[c][b][qx] = Lq x (over k2, which stage 3 has consumed)
#pragma unroll
  for (int ki = 0; ki < KI; ++ki) {
    const int it = tid + ki * kThreads;
    if (it >= N2 * NV) break;
    const int jv = it % NV, ln = it / NV;
    const int b = ln % N1, c = ln / N1;
    const uint32_t base = dx(c, b, jv);
    eo_out<Q, T, VC>(
        [&](auto i, pk& s, pk& d) { eo_row<T, TL, IDFT_IC(i)>(keep_p[ki], keep_m[ki], s, d, eo); },
        [&](auto q, const pk& r) { dref(at_x(base, IDFT_IC(q))) = r; });
  }
  __syncthreads();
  // stage 5: y lines (c, qx): t2[c][qy][qx] = Lq t1, in place (over t1)
  for (int it = tid; it < N1 * Q * NV; it += kThreads) {
    const int jv = it % NV, ln = it / NV;
    const int qx = ln % Q, c = ln / Q;
    const uint32_t base = dy(c, qx, jv);
    pk v[N1], vp[(N1 + 1) / 2], vm[N1 / 2];
#pragma unroll
    for (int b = 0; b < N1; ++b) v[b] = dref(at_y(base, b));
    eo_split(v, vp, vm);
    eo_out<Q, T, VC>(
        [&](auto i, pk& s, pk& d) { eo_row<T, TL, IDFT_IC(i)>(vp, vm, s, d, eo); },
        [&](auto qy, const pk& r) { dref(at_y(base, IDFT_IC(qy))) = r; });
  }
  __syncthreads();
  // stage 6: z lines (qy, qx) through the quadrature weights, in place:
  // s2 = Lq^T (w .* (Lq t2)) (over t2)
#pragma unroll
  for (int k6 = 0; k6 < K6; ++k6) {
    const int it = tid + k6 * kThreads;
    if (it >= Q2 * NV) break;
    const int jv = it % NV, ln = it / NV;  // ln = qy * Q + qx
    const int qx = ln % Q, qy = ln / Q;
    const uint32_t base = dc(qy, qx, jv);
    pk v[N1], vp[(N1 + 1) / 2], vm[N1 / 2];
#pragma unroll
    for (int c = 0; c < N1; ++c) v[c] = dref(at_c(base, c));
    eo_split(v, vp, vm);
    pk u[Q], up[(Q + 1) / 2], um[Q / 2];
    eo_out<Q, T, VC>([&](auto i, pk& s, pk& d) { eo_row<T, TL, IDFT_IC(i)>(vp, vm, s, d, eo); },
                     [&](auto qz, const pk& r) {
                       const T wq = wpre[k6][IDFT_IC(qz)];
#pragma unroll
                       for (int uu = 0; uu < VC; ++uu) u[IDFT_IC(qz)].v[uu] = r.v[uu] * wq;
                     });
    eo_split(u, up, um);
    eo_out<N1, T, VC>([&](auto i, pk& s, pk& d) { eo_row<T, TLt, IDFT_IC(i)>(up, um, s, d, eo); },
                      [&](auto c, const pk& r) { dref(at_c(base, IDFT_IC(c))) = r; });
  }
  __syncthreads();
  // stage 7: y lines back (c, qx), in place: s1[c][b][qx] = Lq^T s2 (over s2)
  for (int it = tid; it < N1 * Q * NV; it += kThreads) {
    const int jv = it % NV, ln = it / NV;
    const int qx = ln % Q, c = ln / Q;
    const uint32_t base = dy(c, qx, jv);
    pk v[Q], vp[(Q + 1) / 2], vm[Q / 2];
#pragma unroll
    for (int qy = 0; qy < Q; ++qy) v[qy] = dref(at_y(base, qy));
    eo_split(v, vp, vm);
    eo_out<N1, T, VC>(
        [&](auto i, pk& s, pk& d) { eo_row<T, TLt, IDFT_IC(i)>(vp, vm, s, d, eo); },
        [&](auto b, const pk& r) { dref(at_y(base, IDFT_IC(b))) = r; });
  }
  __syncthreads();
  // stage 8: x lines back (c, b) plus the kinetic part, into the cell scratch;
  // with dotp (node-scaled MINRES inputs only, XN), also the cell's x.(H_e x)
  // per column, into dotp[column * ncell + cell] (x from the kept input lines:
  // each item holds the same line as in stage 1)
  double dotv[VC];
#pragma unroll
  for (int u = 0; u < VC; ++u) dotv[u] = 0.0;
#pragma unroll
  for (int ki = 0; ki < KI; ++ki) {
    const int it = tid + ki * kThreads;
    if (it >= N2 * NV) break;
    const int jv = it % NV, ln = it / NV;
    if (jv * VC >= bw) continue;
    const int b = ln % N1, c = ln / N1;
    const uint32_t base = dx(c, b, jv);
    pk v[Q], vp[(Q + 1) / 2], vm[Q / 2];
#pragma unroll
    for (int qx = 0; qx < Q; ++qx) v[qx] = dref(at_x(base, qx));
    eo_split(v, vp, vm);
    T* yp = Yc + (size_t)((e * NE + ln * N1) * ub + c0 + jv * VC);
    eo_out<N1, T, VC>([&](auto i, pk& s, pk& d) { eo_row<T, TLt, IDFT_IC(i)>(vp, vm, s, d, eo); },
                      [&](auto a, const pk& r) {
                        constexpr int A = IDFT_IC(a);
                        const pk kin = rA[(ln * N1P + A) * NV + jv];
#pragma unroll
                        for (int u = 0; u < VC; ++u) {
                          const T y = hq * kin.v[u] + r.v[u];
                          if (jv * VC + u < bw) yp[(size_t)A * ub + u] = y;
                          if constexpr (XN) {
                            if (dotp)
                              dotv[u] += (double)eo_unsplit<N1, A>(keep_p[ki], keep_m[ki], u) *
                                         (double)y;
                          }
                        }
                      });
  }
  if (XN && dotp) {
    // per column jv: the threads tid = jv (mod NV) (kThreads and 64 are
    // multiples of NV), summed over each wave, then over the 4 waves
    __syncthreads();  // stage 8 has read rA
    double* red = reinterpret_cast<double*>(rA);
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int u = 0; u < VC; ++u) {
      double t = dotv[u];
#pragma unroll
      for (int o = 32; o >= NV; o >>= 1) t += __shfl_xor(t, o, 64);
      if (lane < NV) red[(wave * NV + lane) * VC + u] = t;
    }
    __syncthreads();
    if (tid < NV) {
#pragma unroll
      for (int u = 0; u < VC; ++u) {
        const int j = tid * VC + u;
        if (j < bw) {
          double t = 0.0;
#pragma unroll
          for (int w = 0; w < kThreads / 64; ++w) t += red[(w * NV + tid) * VC + u];
          dotp[(size_t)(uint32_t)(c0 + j) * ncell + e] = t;
        }
      }
    }
  }
}
#undef IDFT_IC

// Phase 2: every node sums its contributions from the (at most 8) cells that
// share it -- found arithmetically on the structured mesh, so no atomics and
// no colouring -- and combines them with up to three other blocks:
//   Y = beta Y + alpha S[n] sum + a ca[j] xn[n] X + b cb[j] Z + c cc[j] W
// (xn, the operator input's node scale, optional).  With `partial`, the
// per-column block partials of V.Y (FP64) for the scaled operator input
// V = xcol[j] xn[n] Xin, optionally stored.  Nodes are taken in their
// linear order n = (iz N1 + iy) N1 + ix: block b covers R consecutive chunks
// of per = 256 / (B / VW) nodes, [(b R + r) per, (b R + r + 1) per), so a
// chunk may cross x rows and every lane has work; thread t owns column group
// t % (B / VW) and position t / (B / VW) within each chunk.
// Unsigned 32-bit division by a run-time constant d >= 2 through a
// multiplier computed on the host (Granlund-Montgomery): with
// s = ceil(log2 d) and m = floor(2^32 (2^s - d) / d) + 1,
// n / d = (t + ((n - t) >> 1)) >> (s - 1), t = mulhi(n, m), for every n.
struct FastDiv {
  uint32_t m, s;
};
inline FastDiv make_fastdiv(uint32_t d) {
  uint32_t s = 0;
  while ((uint64_t(1) << s) < d) ++s;
  const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1;
  return FastDiv{(uint32_t)m, s};
}
__device__ inline uint32_t fast_div(uint32_t n, FastDiv f) {
  const uint32_t t = __umulhi(n, f.m);
  return (t + ((n - t) >> 1)) >> (f.s - 1);
}

// VW consecutive elements of a node-major row (16-byte aligned for VW > 1).
template <typename T, int VW>
__device__ inline void vload(const T* __restrict__ p, T (&o)[VW]) {
  if constexpr (VW == 1) {
    o[0] = p[0];
  } else if constexpr (VW == 2) {
    typedef typename std::conditional<std::is_same<T, float>::value, float2, double2>::type v2;
    const v2 v = *reinterpret_cast<const v2*>(p);
    o[0] = v.x; o[1] = v.y;
  } else {
    const typename Vec4<T>::type v = *reinterpret_cast<const typename Vec4<T>::type*>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  }
}
template <typename T, int VW>
__device__ inline void vstore(T* __restrict__ p, const T (&o)[VW]) {
  if constexpr (VW == 1) {
    p[0] = o[0];
  } else if constexpr (VW == 2) {
    typedef typename std::conditional<std::is_same<T, float>::value, float2, double2>::type v2;
    *reinterpret_cast<v2*>(p) = v2{o[0], o[1]};
  } else {
    typedef typename Vec4<T>::type v4;
    *reinterpret_cast<v4*>(p) = v4{o[0], o[1], o[2], o[3]};
  }
}

// Sum of the (at most 8) cell contributions of node nidx for the VW columns
// from j0: the cells sharing the node are found arithmetically on the
// structured mesh (nodes n = (iz N1 + iy) N1 + ix).
// The cells sharing node index i along one axis: (c0, a0) and, when the node
// lies on an interior cell face (n = 2), (c1, a1), cell c and local index a,
// in the order (left cell, right cell).
struct AxisCells {
  uint32_t c0, a0, c1, a1;
  int n;
};
template <int P>
__device__ inline AxisCells axis_cells(uint32_t i, uint32_t nc) {
  const uint32_t q = i / P, r = i - q * P;
  AxisCells A;
  if (r != 0) { A.c0 = A.c1 = q; A.a0 = A.a1 = r; A.n = 1; }
  else if (q == 0) { A.c0 = A.c1 = 0; A.a0 = A.a1 = 0; A.n = 1; }
  else if (q == nc) { A.c0 = A.c1 = nc - 1; A.a0 = A.a1 = P; A.n = 1; }
  else { A.c0 = q - 1; A.a0 = P; A.c1 = q; A.a1 = 0; A.n = 2; }
  return A;
}
template <typename TY, int P, int VW>
__device__ inline void gather_node(const Geo& g, FastDiv divN1, uint32_t B, uint32_t nidx,
                                   const TY* __restrict__ Yc, uint32_t j0, double (&sum)[VW]) {
  constexpr uint32_t n1 = P + 1, ne = n1 * n1 * n1;
  const uint32_t N1 = g.N1, nc = g.nc;
  const uint32_t row = fast_div(nidx, divN1), ix = nidx - row * N1;
  const uint32_t iz = fast_div(row, divN1), iy = row - iz * N1;
  // contributions (cell, local index) along each axis (selects, no arrays:
  // run-time-indexed arrays would live in scratch memory)
  const AxisCells X = axis_cells<P>(ix, nc), Y = axis_cells<P>(iy, nc), Z = axis_cells<P>(iz, nc);
  // all 8 contribution slots are loaded before the first is summed, so
  // the loads are in flight together
  TY c8[8][VW];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int kz = k >> 2, ky = (k >> 1) & 1, kx = k & 1;
    // no branch: a contribution the node does not have loads the node's
    // first one (same cache line) and is then zeroed, so that all eight
    // loads are issued before the first wait (behind a per-load branch
    // the compiler waits for each load in turn)
    const bool ok = kz < Z.n && ky < Y.n && kx < X.n;
    const bool sz = ok && kz, sy = ok && ky, sx = ok && kx;
    // the host checks cells n_e B < 2^32
    const uint32_t e = ((sz ? Z.c1 : Z.c0) * nc + (sy ? Y.c1 : Y.c0)) * nc + (sx ? X.c1 : X.c0);
    const uint32_t l = ((sz ? Z.a1 : Z.a0) * n1 + (sy ? Y.a1 : Y.a0)) * n1 + (sx ? X.a1 : X.a0);
    TY t[VW];
    vload<TY, VW>(Yc + ((e * ne + l) * B + j0), t);
#pragma unroll
    for (int v = 0; v < VW; ++v) c8[k][v] = ok ? t[v] : TY(0);
  }
#pragma unroll
  for (int v = 0; v < VW; ++v) sum[v] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int v = 0; v < VW; ++v) sum[v] += (double)c8[k][v];
}

template <typename TY, typename T, int P, int VW>
__global__ __launch_bounds__(kThreads) void k_cell_gather(
    Geo g, FastDiv divN1, uint32_t B, uint32_t R, const TY* __restrict__ Yc, const T* __restrict__ S,
    T* __restrict__ Y, double alpha, double beta, const T* __restrict__ X, double a,
    const double* __restrict__ ca, const T* __restrict__ Z, double b,
    const double* __restrict__ cb, const T* __restrict__ W, double c,
    const double* __restrict__ cc, const T* __restrict__ Xin, const double* __restrict__ xn,
    const double* __restrict__ xcol, T* __restrict__ V, double* __restrict__ partial) {
  __shared__ double red[kThreads * VW];
  const uint32_t N1 = g.N1;
  const uint32_t nb = B / VW;  // column groups of a node
  const uint32_t tid = threadIdx.x, per = kThreads / nb;
  const uint32_t jg = tid % nb, off = tid / nb, j0 = jg * VW;
  double dot[VW];
#pragma unroll
  for (int v = 0; v < VW; ++v) dot[v] = 0.0;
  if (off < per) {
    double caj[VW], cbj[VW], ccj[VW], xs[VW];
#pragma unroll
    for (int v = 0; v < VW; ++v) {
      caj[v] = X ? a * (ca ? ca[j0 + v] : 1.0) : 0.0;
      cbj[v] = Z ? b * (cb ? cb[j0 + v] : 1.0) : 0.0;
      ccj[v] = W ? c * (cc ? cc[j0 + v] : 1.0) : 0.0;
      xs[v] = xcol ? xcol[j0 + v] : 1.0;
    }
    const uint32_t nn = N1 * N1 * N1;
    for (uint32_t rr = 0; rr < R; ++rr) {
      const uint32_t nidx = (blockIdx.x * R + rr) * per + off;
      if (nidx >= nn) break;
      double sum[VW];
      gather_node<TY, P, VW>(g, divN1, B, nidx, Yc, j0, sum);
      const uint32_t n = nidx;
      const uint32_t i = n * B + j0;
      const double sn = xn ? xn[n] : 1.0;
      const double sa = alpha * (double)S[n];
      T xv[VW], zv[VW], wv[VW], yv[VW], xin[VW], yo[VW];
      if (X) vload<T, VW>(X + i, xv);
      if (Z) vload<T, VW>(Z + i, zv);
      if (W) vload<T, VW>(W + i, wv);
      if (beta != 0.0) vload<T, VW>(Y + i, yv);
      if (partial) vload<T, VW>(Xin + i, xin);
#pragma unroll
      for (int v = 0; v < VW; ++v) {
        double t = sa * sum[v];
        if (X) t += caj[v] * (sn * (double)xv[v]);
        if (Z) t += cbj[v] * (double)zv[v];
        if (W) t += ccj[v] * (double)wv[v];
        if (beta != 0.0) t += beta * (double)yv[v];
        yo[v] = (T)t;
      }
      if (partial) {
        T vv[VW];
#pragma unroll
        for (int v = 0; v < VW; ++v) {
          vv[v] = (T)(xs[v] * (sn * (double)xin[v]));
          dot[v] += (double)vv[v] * (double)yo[v];
        }
        if (V) vstore<T, VW>(V + i, vv);
      }
      vstore<T, VW>(Y + i, yo);
    }
  }
  if (partial) {
#pragma unroll
    for (int v = 0; v < VW; ++v) red[tid * VW + v] = dot[v];
    __syncthreads();
    if (tid < B) {
      const uint32_t g2 = tid / VW, v = tid % VW;
      double t = 0.0;
      for (uint32_t k = 0; k < per; ++k) t += red[(k * nb + g2) * VW + v];
      partial[((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * B + tid] = t;
    }
  }
}

// MINRES step, phase 2 with the Lanczos recurrence fused in (sum-factorised
// cell products, whose cell kernel has already produced v.Hv, so that the
// step's alpha is known before this kernel runs):
//   r_new = S sum + (cvn_j D r2 + ca2_j r2) + cr1_j r1     (over r1, in place)
// with cvn = -eps/beta, ca2 = -alpha/beta, cr1 = -beta/beta_old (per column,
// from the state), and block partials of the three node dots the next step
// needs: r_new.D r_new (beta^2), D^2 r_new^2 and D r_new.r2, at
// partial[(k B + j) nblk + block] for k = 0, 1, 2.  Each thread reads and writes the same entries
// of r1, so the update in place is safe.
// UPD: also the previous step's direction and solution update, deferred to
// here because this kernel reads r1 (that step's Lanczos residual) anyway:
//   w <- cwv D r1 + cww1 w + cww w1,  x += cx w   (the previous step's
// coefficients, which the state still holds: its Givens kernel wrote them
// and this step's has not run yet)
template <typename TY, typename T, int P, int VW, bool UPD>
__global__ __launch_bounds__(kThreads) void k_minres_gather(
    Geo g, FastDiv divN1, uint32_t B, uint32_t R, const TY* __restrict__ Yc,
    const T* __restrict__ S, const double* __restrict__ D, const T* __restrict__ r2,
    T* __restrict__ r1, const double* __restrict__ st, double* __restrict__ partial,
    T* __restrict__ w, const T* __restrict__ w1, T* __restrict__ xs) {
  __shared__ double red[kThreads * VW];
  const uint32_t N1 = g.N1;
  const uint32_t nb = B / VW;
  const uint32_t tid = threadIdx.x, per = kThreads / nb;
  const uint32_t jg = tid % nb, off = tid / nb, j0 = jg * VW;
  double q1[VW], q2[VW], q3[VW];
#pragma unroll
  for (int v = 0; v < VW; ++v) q1[v] = q2[v] = q3[v] = 0.0;
  if (off < per) {
    double cvn[VW], ca2[VW], cr1[VW], cwv[VW], cw1[VW], cww[VW], cx[VW];
#pragma unroll
    for (int v = 0; v < VW; ++v) {
      cvn[v] = st[MS_CVN * B + j0 + v];
      ca2[v] = st[MS_CA2 * B + j0 + v];
      cr1[v] = st[MS_CR1 * B + j0 + v];
      if constexpr (UPD) {
        cwv[v] = st[MS_CWV * B + j0 + v];
        cw1[v] = st[MS_CWW1 * B + j0 + v];
        cww[v] = st[MS_CWW * B + j0 + v];
        cx[v] = st[MS_CX * B + j0 + v];
      }
    }
    const uint32_t nn = N1 * N1 * N1;
    // groups of R consecutive chunks, strided over a grid of at most
    // kMinresGatherBlocks work-groups (fewer partial rows to reduce)
    const uint32_t ngroups = (nn + R * per - 1) / (R * per);
    for (uint32_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x)
    for (uint32_t rr = 0; rr < R; ++rr) {
      const uint32_t n = (grp * R + rr) * per + off;
      if (n >= nn) break;
      double sum[VW];
      gather_node<TY, P, VW>(g, divN1, B, n, Yc, j0, sum);
      const uint32_t i = n * B + j0;
      const double dn = D[n];
      const double sn = (double)S[n];
      T xv[VW], zv[VW], ro[VW];
      vload<T, VW>(r2 + i, xv);
      vload<T, VW>(r1 + i, zv);
      T wv[VW], w1v[VW], sv[VW];
      if constexpr (UPD) {
        vload<T, VW>(w + i, wv);
        vload<T, VW>(w1 + i, w1v);
        vload<T, VW>(xs + i, sv);
      }
#pragma unroll
      for (int v = 0; v < VW; ++v) {
        const double x = (double)xv[v];
        const double t = sn * sum[v] + (cvn[v] * dn + ca2[v]) * x + cr1[v] * (double)zv[v];
        ro[v] = (T)t;
        const double r = (double)ro[v], dr = dn * r;
        q1[v] += r * dr;
        q2[v] += dr * dr;
        q3[v] += dr * x;
        if constexpr (UPD) {
          const double wn = cwv[v] * (dn * (double)zv[v]) + cw1[v] * (double)wv[v] +
                            cww[v] * (double)w1v[v];
          wv[v] = (T)wn;
          sv[v] = (T)((double)sv[v] + cx[v] * wn);
        }
      }
      vstore<T, VW>(r1 + i, ro);
      if constexpr (UPD) {
        vstore<T, VW>(w + i, wv);
        vstore<T, VW>(xs + i, sv);
      }
    }
  }
  // partials column-major, [k][column][block], so that each column's
  // reduction reads contiguous memory
  const int64_t nblk = (int64_t)gridDim.x;
  auto emit = [&](int k, const double (&q)[VW]) {
#pragma unroll
    for (int v = 0; v < VW; ++v) red[tid * VW + v] = q[v];
    __syncthreads();
    if (tid < B) {
      const uint32_t g2 = tid / VW, v = tid % VW;
      double t = 0.0;
      for (uint32_t m = 0; m < per; ++m) t += red[(m * nb + g2) * VW + v];
      partial[((int64_t)k * B + tid) * nblk + blockIdx.x] = t;
    }
    __syncthreads();
  };
  emit(0, q1);
  emit(1, q2);
  emit(2, q3);
}

// ------------------------------------------------------ Hamiltonian build
// H_e = K_e/2 + V_e,  V_e = N^T diag(v JxW) N by sum factorisation.
template <int P>
__global__ __launch_bounds__(kThreads) void k_build_h(Geo g, const double* __restrict__ Lq,
                                                      const double* __restrict__ M1,
                                                      const double* __restrict__ D1, double hh,
                                                      const double* __restrict__ veff,
                                                      const double* __restrict__ jxw,
                                                      double* __restrict__ h64,
                                                      float* __restrict__ h32) {
  constexpr int N1 = P + 1, Q = P + 3, NE = ne_of(P), NEP = nep_of(P), RT = NEP / 16;
  constexpr int N2 = N1 * N1;
  __shared__ double W[Q * Q * Q];
  __shared__ double Pq[Q * N2];         // Pq[q][a][a'] = L_a(q) L_a'(q)
  __shared__ double T1[Q * Q * N2];     // [qz][qy][a][a']
  __shared__ double T2[Q * N2 * N2];    // [qz][b][b'][a][a']
  __shared__ double sM[N2], sD[N2];
  const int64_t e = blockIdx.x;
  const int tid = threadIdx.x;
  for (int i = tid; i < Q * Q * Q; i += kThreads) W[i] = veff[e * Q * Q * Q + i] * jxw[i];
  for (int i = tid; i < Q * N2; i += kThreads) {
    const int q = i / N2, a = (i / N1) % N1, a2 = i % N1;
    Pq[i] = Lq[q * N1 + a] * Lq[q * N1 + a2];
  }
  for (int i = tid; i < N2; i += kThreads) { sM[i] = M1[i]; sD[i] = D1[i]; }
  __syncthreads();
  for (int i = tid; i < Q * Q * N2; i += kThreads) {
    const int aa = i % N2, qzy = i / N2;
    double s = 0.0;
#pragma unroll
    for (int qx = 0; qx < Q; ++qx) s += Pq[qx * N2 + aa] * W[qzy * Q + qx];
    T1[i] = s;
  }
  __syncthreads();
  for (int i = tid; i < Q * N2 * N2; i += kThreads) {
    const int aa = i % N2, bb = (i / N2) % N2, qz = i / (N2 * N2);
    double s = 0.0;
#pragma unroll
    for (int qy = 0; qy < Q; ++qy) s += Pq[qy * N2 + bb] * T1[(qz * Q + qy) * N2 + aa];
    T2[i] = s;
  }
  __syncthreads();
  constexpr int NTILE = ntiles_of(P);
  double* o64 = h64 + (size_t)e * NTILE * 256;
  float* o32 = h32 ? h32 + (size_t)e * NTILE * 256 : nullptr;
  for (int idx = tid; idx < NTILE * 256; idx += kThreads) {
    const int s = idx % 4, lane = (idx / 4) % 64, blk = idx / 256;
    int rt = 0;
    while (rt + 1 < RT && tile_index(RT, rt + 1, rt + 1) <= blk) ++rt;
    const int kg = rt + (blk - tile_index(RT, rt, rt));
    const int row = rt * 16 + lane % 16, col = kg * 16 + s * 4 + lane / 16;
    double v = 0.0;
    if (row < NE && col < NE) {
      const int a = row % N1, b = (row / N1) % N1, c = row / N2;
      const int a2 = col % N1, b2 = (col / N1) % N1, c2 = col / N2;
      const int aa = a * N1 + a2, bb = b * N1 + b2, cc = c * N1 + c2;
      double vv = 0.0;
#pragma unroll
      for (int qz = 0; qz < Q; ++qz) vv += Pq[qz * N2 + cc] * T2[(qz * N2 + bb) * N2 + aa];
      const double k = hh * (sD[aa] * sM[bb] * sM[cc] + sM[aa] * sD[bb] * sM[cc] +
                             sM[aa] * sM[bb] * sD[cc]);
      v = 0.5 * k + vv;
    }
    o64[idx] = v;
    if (o32) o32[idx] = (float)v;
  }
}

// ------------------------------------------------------ tensor contraction
// in [N1][N1][N1][NC] -> out [Q][Q][Q][NC] with A [Q][N1] (all in LDS).
template <int N1, int Q, int NC>
__device__ void tp_eval(const double* A, const double* in, double* t1, double* t2, double* out) {
  const int tid = threadIdx.x;
  for (int i = tid; i < N1 * N1 * Q * NC; i += kThreads) {
    const int col = i % NC, qx = (i / NC) % Q, cb = i / (NC * Q);
    const double* src = in + cb * N1 * NC + col;
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < N1; ++a) s += A[qx * N1 + a] * src[a * NC];
    t1[i] = s;
  }
  __syncthreads();
  for (int i = tid; i < N1 * Q * Q * NC; i += kThreads) {
    const int col = i % NC, qx = (i / NC) % Q, qy = (i / (NC * Q)) % Q, c = i / (NC * Q * Q);
    const double* src = t1 + (c * N1 * Q + qx) * NC + col;
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < N1; ++b) s += A[qy * N1 + b] * src[b * Q * NC];
    t2[i] = s;
  }
  __syncthreads();
  for (int i = tid; i < Q * Q * Q * NC; i += kThreads) {
    const int col = i % NC, qxy = (i / NC) % (Q * Q), qz = i / (NC * Q * Q);
    const double* src = t2 + qxy * NC + col;
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < N1; ++c) s += A[qz * N1 + c] * src[c * Q * Q * NC];
    out[i] = s;
  }
  __syncthreads();
}

// Transpose: in [Q][Q][Q][NC] -> out [N1][N1][N1][NC] (integration against N).
template <int N1, int Q, int NC>
__device__ void tp_integrate(const double* A, const double* in, double* t1, double* t2,
                             double* out) {
  const int tid = threadIdx.x;
  for (int i = tid; i < Q * Q * N1 * NC; i += kThreads) {  // t1[qz][qy][a][col]
    const int col = i % NC, a = (i / NC) % N1, qzy = i / (NC * N1);
    const double* src = in + qzy * Q * NC + col;
    double s = 0.0;
#pragma unroll
    for (int qx = 0; qx < Q; ++qx) s += A[qx * N1 + a] * src[qx * NC];
    t1[i] = s;
  }
  __syncthreads();
  for (int i = tid; i < Q * N1 * N1 * NC; i += kThreads) {  // t2[qz][b][a][col]
    const int col = i % NC, a = (i / NC) % N1, b = (i / (NC * N1)) % N1, qz = i / (NC * N1 * N1);
    const double* src = t1 + (qz * Q * N1 + a) * NC + col;
    double s = 0.0;
#pragma unroll
    for (int qy = 0; qy < Q; ++qy) s += A[qy * N1 + b] * src[qy * N1 * NC];
    t2[i] = s;
  }
  __syncthreads();
  for (int i = tid; i < N1 * N1 * N1 * NC; i += kThreads) {  // out[c][b][a][col]
    const int col = i % NC, ba = (i / NC) % (N1 * N1), c = i / (NC * N1 * N1);
    const double* src = t2 + ba * NC + col;
    double s = 0.0;
#pragma unroll
    for (int qz = 0; qz < Q; ++qz) s += A[qz * N1 + c] * src[qz * N1 * N1 * NC];
    out[i] = s;
  }
  __syncthreads();
}

template <int N1, int Q, int NC>
struct TpScratch {
  static constexpr int kIn = (N1 * N1 * N1 > Q * Q * Q ? N1 * N1 * N1 : Q * Q * Q) * NC;
  static constexpr int kT1 = (N1 * N1 * Q > Q * Q * N1 ? N1 * N1 * Q : Q * Q * N1) * NC;
  static constexpr int kT2 = (N1 * Q * Q > Q * N1 * N1 ? N1 * Q * Q : Q * N1 * N1) * NC;
  static constexpr int kOut = kIn;
};

// -------------------------------------------------------------- density
template <int P, int NC>
__global__ __launch_bounds__(kThreads) void k_density(Geo g, const double* __restrict__ Lq,
                                                      const double* __restrict__ X,
                                                      const double* __restrict__ S, int B,
                                                      const double* __restrict__ occ,
                                                      double* __restrict__ rho) {
  constexpr int N1 = P + 1, Q = P + 3, NE = ne_of(P), NQ = Q * Q * Q;
  typedef TpScratch<N1, Q, NC> Sz;
  __shared__ double A[Q * N1];
  __shared__ double sin_[Sz::kIn], t1[Sz::kT1], t2[Sz::kT2], sout[Sz::kOut];
  __shared__ double acc[NQ];
  __shared__ int64_t nodes[NE];
  const int64_t e = blockIdx.x;
  const int64_t cx = e % g.nc, cy = (e / g.nc) % g.nc, cz = e / ((int64_t)g.nc * g.nc);
  const int tid = threadIdx.x;
  for (int i = tid; i < Q * N1; i += kThreads) A[i] = Lq[i];
  for (int l = tid; l < NE; l += kThreads) nodes[l] = cell_node(g, cx, cy, cz, l);
  for (int q = tid; q < NQ; q += kThreads) acc[q] = 0.0;
  __syncthreads();
  for (int j0 = 0; j0 < B; j0 += NC) {
    for (int i = tid; i < NE * NC; i += kThreads) {
      const int l = i / NC, k = i % NC, j = j0 + k;
      sin_[i] = j < B ? S[nodes[l]] * X[nodes[l] * B + j] : 0.0;
    }
    __syncthreads();
    tp_eval<N1, Q, NC>(A, sin_, t1, t2, sout);
    for (int q = tid; q < NQ; q += kThreads) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const int j = j0 + k;
        if (j < B) { const double v = sout[q * NC + k]; s += 2.0 * occ[j] * v * v; }
      }
      acc[q] += s;
    }
    __syncthreads();
  }
  for (int q = tid; q < NQ; q += kThreads) rho[e * NQ + q] = acc[q];
}

// ---------------------------------------------------- pair quadrature dots
// out[cell][k] = sum_q wq[cell][q] A_{i_k}(q) B_{j_k}(q)
template <int P, int NC>
__global__ __launch_bounds__(kThreads) void k_pair_quad(Geo g, const double* __restrict__ Lq,
                                                        const double* __restrict__ Av,
                                                        const double* __restrict__ Bv, int B,
                                                        const int* __restrict__ pairs, int np,
                                                        const double* __restrict__ wq,
                                                        int64_t wstride,
                                                        double* __restrict__ out) {
  constexpr int N1 = P + 1, Q = P + 3, NE = ne_of(P), NQ = Q * Q * Q, H = NC / 2;
  typedef TpScratch<N1, Q, NC> Sz;
  __shared__ double A[Q * N1];
  __shared__ double sin_[Sz::kIn], t1[Sz::kT1], t2[Sz::kT2], sout[Sz::kOut];
  __shared__ double w[NQ];
  __shared__ int64_t nodes[NE];
  __shared__ double red[kThreads / 64][H];
  const int64_t e = blockIdx.x;
  const int64_t cx = e % g.nc, cy = (e / g.nc) % g.nc, cz = e / ((int64_t)g.nc * g.nc);
  const int tid = threadIdx.x;
  for (int i = tid; i < Q * N1; i += kThreads) A[i] = Lq[i];
  for (int l = tid; l < NE; l += kThreads) nodes[l] = cell_node(g, cx, cy, cz, l);
  for (int q = tid; q < NQ; q += kThreads) w[q] = wq[e * wstride + q];
  __syncthreads();
  for (int k0 = 0; k0 < np; k0 += H) {
    // columns 0..H-1 <- A_i, H..NC-1 <- B_j of pairs k0..k0+H-1
    for (int i = tid; i < NE * NC; i += kThreads) {
      const int l = i / NC, k = i % NC, kp = k0 + (k % H);
      double v = 0.0;
      if (kp < np) {
        if (k < H) v = Av[nodes[l] * B + pairs[2 * kp]];
        else v = Bv[nodes[l] * B + pairs[2 * kp + 1]];
      }
      sin_[i] = v;
    }
    __syncthreads();
    tp_eval<N1, Q, NC>(A, sin_, t1, t2, sout);
    double part[H];
#pragma unroll
    for (int k = 0; k < H; ++k) part[k] = 0.0;
    for (int q = tid; q < NQ; q += kThreads) {
#pragma unroll
      for (int k = 0; k < H; ++k) part[k] += w[q] * sout[q * NC + k] * sout[q * NC + H + k];
    }
#pragma unroll
    for (int k = 0; k < H; ++k) {
      double v = part[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if ((tid & 63) == 0) red[tid / 64][k] = v;
    }
    __syncthreads();
    if (tid < H && k0 + tid < np) {
      double s = 0.0;
      for (int wv = 0; wv < kThreads / 64; ++wv) s += red[wv][tid];
      out[e * np + k0 + tid] = s;
    }
    __syncthreads();
  }
}

// --------------------------------------------- quadrature cell operator
// Y += N^T diag(wq) N X on the cells of one colour (no write conflicts).
template <int P, int NC>
__global__ __launch_bounds__(kThreads) void k_quad_operator(Geo g, const double* __restrict__ Lq,
                                                            const double* __restrict__ X, int B,
                                                            const double* __restrict__ wq,
                                                            double* __restrict__ Y, int colour,
                                                            int64_t nw) {
  constexpr int N1 = P + 1, Q = P + 3, NE = ne_of(P), NQ = Q * Q * Q;
  typedef TpScratch<N1, Q, NC> Sz;
  __shared__ double A[Q * N1];
  __shared__ double sin_[Sz::kIn], t1[Sz::kT1], t2[Sz::kT2], sout[Sz::kOut];
  __shared__ int64_t nodes[NE];
  const int64_t e = colour_cell(g, colour, blockIdx.x, nw);
  if (e < 0) return;
  const int64_t cx = e % g.nc, cy = (e / g.nc) % g.nc, cz = e / ((int64_t)g.nc * g.nc);
  const int tid = threadIdx.x;
  for (int i = tid; i < Q * N1; i += kThreads) A[i] = Lq[i];
  for (int l = tid; l < NE; l += kThreads) nodes[l] = cell_node(g, cx, cy, cz, l);
  __syncthreads();
  for (int j0 = 0; j0 < B; j0 += NC) {
    for (int i = tid; i < NE * NC; i += kThreads) {
      const int l = i / NC, j = j0 + i % NC;
      sin_[i] = j < B ? X[nodes[l] * B + j] : 0.0;
    }
    __syncthreads();
    tp_eval<N1, Q, NC>(A, sin_, t1, t2, sout);
    for (int i = tid; i < NQ * NC; i += kThreads) sout[i] *= wq[e * NQ + i / NC];
    __syncthreads();
    tp_integrate<N1, Q, NC>(A, sout, t1, t2, sin_);
    for (int i = tid; i < NE * NC; i += kThreads) {
      const int l = i / NC, j = j0 + i % NC;
      if (j < B) Y[nodes[l] * B + j] += sin_[i];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ v_xc -> quad
// Child Q1 field (3 child nodes per parent axis) to the parent Gauss points.
template <int P>
__global__ __launch_bounds__(kThreads) void k_vxc_to_quad(Geo g, const double* __restrict__ Wq,
                                                          const double* __restrict__ pot,
                                                          const double* __restrict__ potBase,
                                                          double* __restrict__ veff) {
  constexpr int Q = P + 3, NQ = Q * Q * Q;
  typedef TpScratch<3, Q, 1> Sz;
  __shared__ double A[Q * 3];
  __shared__ double sin_[Sz::kIn], t1[Sz::kT1], t2[Sz::kT2], sout[Sz::kOut];
  const int64_t e = blockIdx.x;
  const int64_t cx = e % g.nc, cy = (e / g.nc) % g.nc, cz = e / ((int64_t)g.nc * g.nc);
  const int tid = threadIdx.x;
  for (int i = tid; i < Q * 3; i += kThreads) A[i] = Wq[i];
  for (int i = tid; i < 27; i += kThreads) {
    const int a = i % 3, b = (i / 3) % 3, c = i / 9;
    sin_[i] = pot[((2 * cz + c) * g.NV1 + (2 * cy + b)) * (int64_t)g.NV1 + (2 * cx + a)];
  }
  __syncthreads();
  tp_eval<3, Q, 1>(A, sin_, t1, t2, sout);
  for (int q = tid; q < NQ; q += kThreads) veff[e * NQ + q] = potBase[e * NQ + q] + sout[q];
}

// ------------------------------------------------------------------ force
// S(q) = sum_j Psi_j(q) P_j(q) at Q points per axis (child: Q=4 via Ac; parent:
// Q=p+3 via Lq), then force[child node] += sum_q W(q) N^child(q) S(q).
template <int P, int Q, int NC>
__global__ __launch_bounds__(kThreads) void k_force(Geo g, const double* __restrict__ Aeval,
                                                    const double* __restrict__ Wchild,
                                                    const double* __restrict__ Psi,
                                                    const double* __restrict__ Pv, int B,
                                                    double jxwChild,
                                                    const double* __restrict__ jxwParent,
                                                    double* __restrict__ force, int colour,
                                                    int64_t nw) {
  constexpr int N1 = P + 1, NE = ne_of(P), NQ = Q * Q * Q, H = NC / 2;
  typedef TpScratch<N1, Q, NC> Sz;
  __shared__ double A[Q * N1];
  __shared__ double Wc[Q * 3];
  __shared__ double sin_[Sz::kIn], t1[Sz::kT1], t2[Sz::kT2], sout[Sz::kOut];
  __shared__ double Sq[NQ];
  __shared__ double u1[3 * 3 * Q], u2[3 * 3 * 3];
  __shared__ double v1[Q * Q * 3];
  __shared__ int64_t nodes[NE];
  const int64_t e = colour_cell(g, colour, blockIdx.x, nw);
  if (e < 0) return;
  const int64_t cx = e % g.nc, cy = (e / g.nc) % g.nc, cz = e / ((int64_t)g.nc * g.nc);
  const int tid = threadIdx.x;
  for (int i = tid; i < Q * N1; i += kThreads) A[i] = Aeval[i];
  for (int i = tid; i < Q * 3; i += kThreads) Wc[i] = Wchild[i];
  for (int l = tid; l < NE; l += kThreads) nodes[l] = cell_node(g, cx, cy, cz, l);
  for (int q = tid; q < NQ; q += kThreads) Sq[q] = 0.0;
  __syncthreads();
  for (int j0 = 0; j0 < B; j0 += H) {
    for (int i = tid; i < NE * NC; i += kThreads) {
      const int l = i / NC, k = i % NC, j = j0 + (k % H);
      double v = 0.0;
      if (j < B) v = (k < H) ? Psi[nodes[l] * B + j] : Pv[nodes[l] * B + j];
      sin_[i] = v;
    }
    __syncthreads();
    tp_eval<N1, Q, NC>(A, sin_, t1, t2, sout);
    for (int q = tid; q < NQ; q += kThreads) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < H; ++k) s += sout[q * NC + k] * sout[q * NC + H + k];
      Sq[q] += s;
    }
    __syncthreads();
  }
  for (int q = tid; q < NQ; q += kThreads) Sq[q] *= jxwParent ? jxwParent[q] : jxwChild;
  __syncthreads();
  // integrate against the child basis: out[c][b][a] (3 child nodes per axis)
  for (int i = tid; i < Q * Q * 3; i += kThreads) {  // v1[qz][qy][a]
    const int a = i % 3, qzy = i / 3;
    double s = 0.0;
    for (int qx = 0; qx < Q; ++qx) s += Wc[qx * 3 + a] * Sq[qzy * Q + qx];
    v1[i] = s;
  }
  __syncthreads();
  for (int i = tid; i < Q * 9; i += kThreads) {  // u1[qz][b][a]
    const int a = i % 3, b = (i / 3) % 3, qz = i / 9;
    double s = 0.0;
    for (int qy = 0; qy < Q; ++qy) s += Wc[qy * 3 + b] * v1[(qz * Q + qy) * 3 + a];
    u1[i] = s;
  }
  __syncthreads();
  for (int i = tid; i < 27; i += kThreads) {
    const int ba = i % 9, c = i / 9;
    double s = 0.0;
    for (int qz = 0; qz < Q; ++qz) s += Wc[qz * 3 + c] * u1[qz * 9 + ba];
    u2[i] = s;
  }
  __syncthreads();
  for (int i = tid; i < 27; i += kThreads) {
    const int a = i % 3, b = (i / 3) % 3, c = i / 9;
    force[((2 * cz + c) * g.NV1 + (2 * cy + b)) * (int64_t)g.NV1 + (2 * cx + a)] += u2[i];
  }
}

// --------------------------------------------------------- vector kernels
// Y = beta Y + ns[n] (a ca[j] X + b cb[j] Z + c cc[j] W); null coefficient
// pointers read as 1, null operands as 0.  N * B < 2^32 (checked by the host).
template <typename T>
__global__ void k_lincomb(uint32_t NB, uint32_t B, T* __restrict__ Y, double beta,
                          const T* __restrict__ X, double a, const double* __restrict__ ca,
                          const T* __restrict__ Z, double b, const double* __restrict__ cb,
                          const T* __restrict__ W, double c, const double* __restrict__ cc,
                          const double* __restrict__ ns) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < NB; i += gridDim.x * blockDim.x) {
    const uint32_t n = i / B, j = i - n * B;
    double v = 0.0;
    if (X) v = a * (ca ? ca[j] : 1.0) * (double)X[i];
    if (Z) v += b * (cb ? cb[j] : 1.0) * (double)Z[i];
    if (W) v += c * (cc ? cc[j] : 1.0) * (double)W[i];
    if (ns) v *= ns[n];
    if (beta != 0.0) v += beta * (double)Y[i];
    Y[i] = (T)v;
  }
}

// Y = a ca[j] X + b cb[j] Z with operands and result of different types
// (null coefficient pointers read as 1), in FP64 arithmetic and rounded once:
// the same expression as k_lincomb, so that a conversion pass before or after
// it is not needed.  Y may be X (each entry is read, then written, by one
// thread).  N * B < 2^32 (checked by the host).
template <typename TY, typename TX, typename TZ>
__global__ void k_lincomb_mixed(uint32_t NB, uint32_t B, TY* Y, const TX* X, double a,
                                const double* __restrict__ ca,
                                const TZ* __restrict__ Z, double b, const double* __restrict__ cb) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < NB; i += gridDim.x * blockDim.x) {
    const uint32_t n = i / B, j = i - n * B;
    double v = a * (ca ? ca[j] : 1.0) * (double)X[i];
    v += b * (cb ? cb[j] : 1.0) * (double)Z[i];
    Y[i] = (TY)v;
  }
}

__global__ void k_d2f(int64_t n, float* __restrict__ y, const double* __restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}
__global__ void k_f2d(int64_t n, double* __restrict__ y, const float* __restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (double)x[i];
}

// Column-wise partial dot products.  Thread t of a block owns column t % B and
// strides over nodes; FP64 accumulation; block results reduced in LDS.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_coldot(int64_t N, int B, const T* __restrict__ X,
                                                     const T* __restrict__ Y,
                                                     const double* __restrict__ w,
                                                     double* __restrict__ partial) {
  __shared__ double red[kThreads];
  const int tid = threadIdx.x;
  const int per = kThreads / B;  // nodes per block step (B <= 256)
  const int col = tid % B, off = tid / B;
  double s = 0.0;
  if (off < per) {
    for (int64_t n = (int64_t)blockIdx.x * per + off; n < N; n += (int64_t)gridDim.x * per) {
      const double x = (double)X[n * B + col], y = (double)Y[n * B + col];
      s += w ? w[n] * x * y : x * y;
    }
  }
  red[tid] = s;
  __syncthreads();
  if (tid < B) {
    double t = 0.0;
    for (int k = 0; k < per; ++k) t += red[k * B + tid];
    partial[(int64_t)blockIdx.x * B + tid] = t;
  }
}

__global__ __launch_bounds__(kThreads) void k_sum_partials(int nblk, int B,
                                                            const double* __restrict__ partial,
                                                            double* __restrict__ out) {
  __shared__ double red[kThreads];
  const int j = blockIdx.x;
  const double s = block_sum_col(partial, nblk, B, j, red);
  if (threadIdx.x == 0) out[j] = s;
}

// Sums of contiguous rows: out[j] = sum of partial[j nblk .. (j + 1) nblk).
__global__ __launch_bounds__(kRedThreads) void k_sum_rows(int nblk, const double* __restrict__ partial,
                                                           double* __restrict__ out) {
  __shared__ double red[kRedThreads / 64];
  const int j = blockIdx.x;
  const double s = block_sum_contig<kRedThreads>(partial + (int64_t)j * nblk, nblk, red);
  if (threadIdx.x == 0) out[j] = s;
}

// Gram partials: block (bx, by) accumulates the entries ij in
// [by * 256, by * 256 + 256) of G = X^T Y over the chunks of `rows` nodes
// bx, bx + gridDim.x, ..., whose rows of X and Y are staged in LDS (the loads
// of a chunk issued together).  Partials column-major, partial[ij nblk + bx],
// so that each entry's reduction reads contiguous memory.
__global__ __launch_bounds__(kThreads) void k_gram(int64_t N, int B, int rows,
                                                   const double* __restrict__ X,
                                                   const double* __restrict__ Y,
                                                   double* __restrict__ partial) {
  extern __shared__ double lds[];
  double* xs = lds;
  double* ys = lds + rows * B;
  const int tid = threadIdx.x;
  const int BB = B * B, RB = rows * B;
  const int ij = blockIdx.y * kThreads + tid;
  const int i = ij / B, j = ij % B;
  double acc = 0.0;
  for (int64_t n0 = (int64_t)blockIdx.x * rows; n0 < N; n0 += (int64_t)gridDim.x * rows) {
    const int cnt = (int)((N - n0) < rows ? (N - n0) : rows);
    const int ce = cnt * B;  // entries of this chunk (contiguous in X and Y)
    __syncthreads();
    // four entries per thread per round, their loads issued together
    for (int t0 = tid; t0 < RB; t0 += 4 * kThreads) {
      double xv[4], yv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = t0 + u * kThreads;
        const bool ok = t < ce;
        const int64_t g = ok ? n0 * B + t : 0;
        xv[u] = X[g];
        yv[u] = Y[g];
        if (!ok) xv[u] = yv[u] = 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = t0 + u * kThreads;
        if (t < RB) {
          xs[t] = xv[u];
          ys[t] = yv[u];
        }
      }
    }
    __syncthreads();
    if (ij < BB)
      for (int r = 0; r < cnt; ++r) acc += xs[r * B + i] * ys[r * B + j];
  }
  if (ij < BB) partial[(int64_t)ij * gridDim.x + blockIdx.x] = acc;
}

// Y = beta Y + ns[n] (X Q)[n]  (X != Y) over tiles of `rows` nodes: the
// tile's rows of X (contiguous) are staged in LDS with their loads issued
// together, then each thread forms the outputs t, t + 256, ... of the tile
// (node t / B, column t % B) from LDS, and the tile is written contiguously.
__global__ __launch_bounds__(kThreads) void k_right_mult(uint32_t N, uint32_t B, uint32_t rows,
                                                         const double* __restrict__ X,
                                                         const double* __restrict__ Q,
                                                         double* __restrict__ Y,
                                                         const double* __restrict__ ns,
                                                         double beta) {
  extern __shared__ double lds[];
  double* qs = lds;
  double* xt = lds + B * B;
  const uint32_t tid = threadIdx.x, RB = rows * B;
  for (uint32_t i = tid; i < B * B; i += kThreads) qs[i] = Q[i];
  for (uint32_t n0 = blockIdx.x * rows; n0 < N; n0 += gridDim.x * rows) {
    const uint32_t cnt = min(rows, N - n0), ce = cnt * B;
    const uint32_t base = n0 * B;  // N B < 2^32 (checked by the host)
    __syncthreads();  // the previous tile's outputs have read xt
    for (uint32_t t0 = tid; t0 < ce; t0 += 4 * kThreads) {
      double xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t t = t0 + u * kThreads;
        xv[u] = X[t < ce ? base + t : base];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t t = t0 + u * kThreads;
        if (t < ce) xt[t] = xv[u];
      }
    }
    __syncthreads();
    for (uint32_t t = tid; t < ce; t += kThreads) {
      const uint32_t r = t / B, j = t - r * B;
      const double* xr = xt + r * B;
      double s = 0.0;
      for (uint32_t i = 0; i < B; ++i) s += xr[i] * qs[i * B + j];
      if (ns) s *= ns[n0 + r];
      Y[base + t] = beta != 0.0 ? beta * Y[base + t] + s : s;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_residual(int64_t N, int B, const T* __restrict__ HX,
                                                       const T* __restrict__ X,
                                                       const double* __restrict__ ev,
                                                       double* __restrict__ partial) {
  __shared__ double red[kThreads];
  const int tid = threadIdx.x;
  const int per = kThreads / B;
  const int col = tid % B, off = tid / B;
  double s = 0.0;
  if (off < per) {
    const double lam = ev[col];
    for (int64_t n = (int64_t)blockIdx.x * per + off; n < N; n += (int64_t)gridDim.x * per) {
      const double r = (double)HX[n * B + col] - lam * (double)X[n * B + col];
      s += r * r;
    }
  }
  red[tid] = s;
  __syncthreads();
  if (tid < B) {
    double t = 0.0;
    for (int k = 0; k < per; ++k) t += red[k * B + tid];
    partial[(int64_t)blockIdx.x * B + tid] = t;
  }
}

__global__ void k_loss(int64_t nQ, int nq, const double* __restrict__ rho,
                       const double* __restrict__ rhoT, const double* __restrict__ w,
                       const double* __restrict__ veff, const double* __restrict__ jxw,
                       double* __restrict__ u, double* __restrict__ partial) {
  __shared__ double red[4][kThreads];
  const int tid = threadIdx.x;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int64_t q = blockIdx.x * (int64_t)kThreads + tid; q < nQ; q += (int64_t)gridDim.x * kThreads) {
    const double jw = jxw[q % nq];
    const double rd = rhoT[q] - 0.5 * rho[q];
    u[q] = rd * w[q];
    s0 += rd * rd * w[q] * jw;
    s1 += rd * rd * jw;
    s2 += fabs(rd * jw);
    s3 += veff[q] * rho[q] * jw;
  }
  red[0][tid] = s0; red[1][tid] = s1; red[2][tid] = s2; red[3][tid] = s3;
  __syncthreads();
  for (int st = kThreads / 2; st > 0; st >>= 1) {
    if (tid < st)
      for (int k = 0; k < 4; ++k) red[k][tid] += red[k][tid + st];
    __syncthreads();
  }
  if (tid < 4) partial[blockIdx.x * 4 + tid] = red[tid][0];
}

__global__ void k_zero_rows(uint32_t NB, uint32_t B, double* __restrict__ X,
                            const unsigned char* __restrict__ mask) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < NB; t += gridDim.x * blockDim.x)
    if (mask[t / B]) X[t] = 0.0;
}


// ------------------------------------------------------------- MINRES
// Per-column state st[k * B + j] (see MinresSlot); one thread per column.
__global__ __launch_bounds__(kThreads) void k_minres_start(int B, int nblk,
                                                            const double* __restrict__ partial,
                                                            const double* __restrict__ negEig,
                                                            double* __restrict__ st, int nq) {
  __shared__ double red[kThreads];
  const int j = blockIdx.x;
  const double s = block_sum_col(partial, nblk, B, j, red);
  __syncthreads();
  const double q2 = nq >= 2 ? block_sum_col(partial + (int64_t)nblk * B, nblk, B, j, red) : 0.0;
  if (threadIdx.x != 0) return;
  st[MS_Q2 * B + j] = q2;
  st[MS_Q3 * B + j] = 0.0;
  const double beta = sqrt(fmax(s, 0.0));
  st[MS_BETA * B + j] = beta;
  st[MS_OLDB * B + j] = 0.0;
  st[MS_DBAR * B + j] = 0.0;
  st[MS_EPSLN * B + j] = 0.0;
  st[MS_PHIBAR * B + j] = beta;
  st[MS_CS * B + j] = -1.0;
  st[MS_SN * B + j] = 0.0;
  st[MS_IT * B + j] = 0.0;
  st[MS_BETA1 * B + j] = beta;
  const bool done = !(beta > st[MS_TOL * B + j]);
  st[MS_DONE * B + j] = done ? 1.0 : 0.0;
  st[MS_CV * B + j] = done ? 0.0 : 1.0 / beta;
  st[MS_CVN * B + j] = st[MS_CV * B + j] * negEig[j];
  st[MS_CR1 * B + j] = 0.0;
}

__global__ __launch_bounds__(kThreads) void k_minres_alpha(int B, int nblk,
                                                            const double* __restrict__ partial,
                                                            double* __restrict__ st) {
  __shared__ double red[kThreads];
  const int j = blockIdx.x;
  const double s = block_sum_col(partial, nblk, B, j, red);
  if (threadIdx.x != 0) return;
  st[MS_ALFA * B + j] = s;
  const bool done = st[MS_DONE * B + j] != 0.0;
  st[MS_CA2 * B + j] = done ? 0.0 : -s / st[MS_BETA * B + j];
}

// alpha of a fused step: v.(H~ v) from the per-cell dots x_e.(H_e x_e)
// (x = S v the cell product's input) plus the node terms of the step's
// output, v.(cvn D r2 + cr1 r1) = cv (cvn Q2 + cr1 Q3) with v = cv D r2.
__global__ __launch_bounds__(kRedThreads) void k_minres_alpha_cell(int B, int ncell,
                                                                    const double* __restrict__ cd,
                                                                    double* __restrict__ st) {
  __shared__ double red[kRedThreads / 64];
  const int j = blockIdx.x;
  const double s = block_sum_contig<kRedThreads>(cd + (int64_t)j * ncell, ncell, red);
  if (threadIdx.x != 0) return;
  const double cv = st[MS_CV * B + j];
  const double a = s + cv * (st[MS_CVN * B + j] * st[MS_Q2 * B + j] +
                             st[MS_CR1 * B + j] * st[MS_Q3 * B + j]);
  st[MS_ALFA * B + j] = a;
  const bool done = st[MS_DONE * B + j] != 0.0;
  st[MS_CA2 * B + j] = done ? 0.0 : -a / st[MS_BETA * B + j];
}

// Partials of r.D r and (D r).(D r) per column (rows [0, nblk) and
// [nblk, 2 nblk) of partial): the start of a fused MINRES solve.
__global__ __launch_bounds__(kThreads) void k_minres_start_dots(int64_t N, int B,
                                                                const double* __restrict__ r,
                                                                const double* __restrict__ D,
                                                                double* __restrict__ partial) {
  __shared__ double red[2][kThreads];
  const int tid = threadIdx.x;
  const int per = kThreads / B;
  const int col = tid % B, off = tid / B;
  double s1 = 0.0, s2 = 0.0;
  if (off < per) {
    for (int64_t n = (int64_t)blockIdx.x * per + off; n < N; n += (int64_t)gridDim.x * per) {
      const double v = r[n * B + col], dv = D[n] * v;
      s1 += v * dv;
      s2 += dv * dv;
    }
  }
  red[0][tid] = s1;
  red[1][tid] = s2;
  __syncthreads();
  if (tid < B) {
    double t1 = 0.0, t2 = 0.0;
    for (int k = 0; k < per; ++k) {
      t1 += red[0][k * B + tid];
      t2 += red[1][k * B + tid];
    }
    partial[(int64_t)blockIdx.x * B + tid] = t1;
    partial[((int64_t)gridDim.x + blockIdx.x) * B + tid] = t2;
  }
}

// r_new = y + ca2 r2 (stored in rnew), partial dot(r_new, D r_new).  The
// preconditioned vector D r_new is not stored: the next operator
// application and update scale r_new by D on the fly.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_minres_lanczos(int64_t N, int B,
                                                             const T* __restrict__ y,
                                                             const T* __restrict__ r2,
                                                             T* __restrict__ rnew,
                                                             const double* __restrict__ D,
                                                             const double* __restrict__ st,
                                                             double* __restrict__ partial) {
  __shared__ double red[kThreads];
  const int tid = threadIdx.x;
  const int per = kThreads / B;
  const int col = tid % B, off = tid / B;
  double s = 0.0;
  if (off < per) {
    const double c = st[MS_CA2 * B + col];
    for (int64_t n = (int64_t)blockIdx.x * per + off; n < N; n += (int64_t)gridDim.x * per) {
      const int64_t i = n * B + col;
      const double rn = (double)y[i] + c * (double)r2[i];
      const T rt = (T)rn;
      rnew[i] = rt;
      s += (double)rt * (D[n] * (double)rt);
    }
  }
  red[tid] = s;
  __syncthreads();
  if (tid < B) {
    double t = 0.0;
    for (int k = 0; k < per; ++k) t += red[k * B + tid];
    partial[(int64_t)blockIdx.x * B + tid] = t;
  }
}

__global__ __launch_bounds__(kRedThreads) void k_minres_givens(int B, int nblk,
                                                                const double* __restrict__ partial,
                                                                const double* __restrict__ negEig,
                                                                double* __restrict__ st, int nq) {
  __shared__ double red[kRedThreads / 64];
  const int j = blockIdx.x;
  double s, q2 = 0.0, q3 = 0.0;
  if (nq >= 3) {  // fused steps: [k][column][block] partials
    s = block_sum_contig<kRedThreads>(partial + (int64_t)j * nblk, nblk, red);
    __syncthreads();
    q2 = block_sum_contig<kRedThreads>(partial + ((int64_t)B + j) * nblk, nblk, red);
    __syncthreads();
    q3 = block_sum_contig<kRedThreads>(partial + ((int64_t)2 * B + j) * nblk, nblk, red);
  } else {
    s = block_sum_col<kRedThreads>(partial, nblk, B, j, red);
  }
  if (threadIdx.x != 0) return;
  double* S = st;
  S[MS_Q2 * B + j] = q2;
  S[MS_Q3 * B + j] = q3;
  if (S[MS_DONE * B + j] != 0.0) {
    S[MS_CWV * B + j] = 0.0; S[MS_CWW1 * B + j] = 0.0; S[MS_CWW * B + j] = 0.0;
    S[MS_CX * B + j] = 0.0; S[MS_CV * B + j] = 0.0; S[MS_CR1 * B + j] = 0.0;
    S[MS_CVN * B + j] = 0.0;
    return;
  }
  const double oldb = S[MS_BETA * B + j];
  const double beta = sqrt(fmax(s, 0.0));
  const double alfa = S[MS_ALFA * B + j];
  const double cs = S[MS_CS * B + j], sn = S[MS_SN * B + j], dbar = S[MS_DBAR * B + j];
  const double oldeps = S[MS_EPSLN * B + j];
  const double delta = cs * dbar + sn * alfa;
  const double gbar = sn * dbar - cs * alfa;
  const double epsln = sn * beta;
  const double ndbar = -cs * beta;
  const double gamma = fmax(sqrt(gbar * gbar + beta * beta), 1e-300);
  const double ncs = gbar / gamma, nsn = beta / gamma;
  const double phibar0 = S[MS_PHIBAR * B + j];
  const double phi = ncs * phibar0, phibar = nsn * phibar0;
  S[MS_OLDB * B + j] = oldb;
  S[MS_BETA * B + j] = beta;
  S[MS_EPSLN * B + j] = epsln;
  S[MS_DBAR * B + j] = ndbar;
  S[MS_CS * B + j] = ncs;
  S[MS_SN * B + j] = nsn;
  S[MS_PHIBAR * B + j] = phibar;
  S[MS_CWV * B + j] = S[MS_CV * B + j] / gamma;  // v = cv D r: cv of this step
  S[MS_CWW1 * B + j] = -oldeps / gamma;
  S[MS_CWW * B + j] = -delta / gamma;
  S[MS_CX * B + j] = phi;
  S[MS_IT * B + j] += 1.0;
  const bool done = !(phibar > S[MS_TOL * B + j]) || beta == 0.0;
  S[MS_DONE * B + j] = done ? 1.0 : 0.0;
  S[MS_CV * B + j] = done ? 0.0 : 1.0 / beta;
  S[MS_CVN * B + j] = S[MS_CV * B + j] * negEig[j];
  S[MS_CR1 * B + j] = done ? 0.0 : -beta / oldb;
}

// w_new = cwv D r + cww1 w1 + cww w  (written over w1);  x += cx w_new,
// where cwv D r = v / gamma, v = cv D r the step's Lanczos vector.
template <typename T, int VW>
__global__ void k_minres_update(uint32_t NB, uint32_t B, const T* __restrict__ r,
                                const double* __restrict__ D, T* __restrict__ w1,
                                const T* __restrict__ w, T* __restrict__ x,
                                const double* __restrict__ st) {
  // VW consecutive columns of one node per thread (B % VW == 0), so every
  // access is one VW-wide load or store
  const uint32_t NG = NB / VW;
  for (uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x; gi < NG; gi += gridDim.x * blockDim.x) {
    const uint32_t i = gi * VW, n = i / B, j = i - n * B;
    T rv[VW], w1v[VW], wv[VW], xv[VW];
    double cwv[VW], cw1[VW], cww[VW], cx[VW];
    vload<T, VW>(r + i, rv);
    vload<T, VW>(w1 + i, w1v);
    vload<T, VW>(w + i, wv);
    vload<T, VW>(x + i, xv);
    vload<double, VW>(st + MS_CWV * B + j, cwv);
    vload<double, VW>(st + MS_CWW1 * B + j, cw1);
    vload<double, VW>(st + MS_CWW * B + j, cww);
    vload<double, VW>(st + MS_CX * B + j, cx);
    const double dn = D[n];
    T wo[VW], xo[VW];
#pragma unroll
    for (int v = 0; v < VW; ++v) {
      const double wn = cwv[v] * (dn * (double)rv[v]) + cw1[v] * (double)w1v[v] +
                        cww[v] * (double)wv[v];
      wo[v] = (T)wn;
      xo[v] = (T)((double)xv[v] + cx[v] * wn);
    }
    vstore<T, VW>(w1 + i, wo);
    vstore<T, VW>(x + i, xo);
  }
}

__global__ void k_copy_done(int B, const double* __restrict__ st, int* __restrict__ flags) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < B) flags[j] = st[MS_DONE * B + j] != 0.0 ? 1 : 0;
}


// dst[n][j] = src[n][off + j]
__global__ void k_copy_cols(int64_t N, int Bs, int off, int Bd, const double* __restrict__ src,
                            double* __restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N * Bd;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = i / Bd;
    dst[i] = src[n * Bs + off + (int)(i % Bd)];
  }
}

// dst[n][k] = src[n][idx[k]] (gather, Bd columns) or, with scatter,
// dst[n][idx[k]] = src[n][k] (Bs columns of src into Bd-wide dst).
__global__ void k_cols_by_index(int64_t N, int Bs, int Bd, int Bi, const int* __restrict__ idx,
                                const double* __restrict__ src, double* __restrict__ dst,
                                int scatter) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N * Bi;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = i / Bi;
    const int k = (int)(i - n * Bi);
    if (scatter) dst[n * Bd + idx[k]] = src[n * Bs + k];
    else dst[n * Bd + k] = src[n * Bs + idx[k]];
  }
}

__global__ void k_mul_q(int64_t nQ, int nq, const double* __restrict__ u,
                        const double* __restrict__ jxw, double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nQ;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = u[i] * jxw[i % nq];
}

// partial[blk][k] = sum_n w[n] X[n][i_k] Y[n][j_k]   (np <= 256)
__global__ __launch_bounds__(kThreads) void k_pairdot(int64_t N, int B, const double* __restrict__ X,
                                                      const double* __restrict__ Y,
                                                      const double* __restrict__ w,
                                                      const int* __restrict__ pairs, int np,
                                                      double* __restrict__ partial) {
  __shared__ double red[kThreads];
  const int tid = threadIdx.x;
  const int per = kThreads / np;
  const int k = tid % np, off = tid / np;
  double s = 0.0;
  if (off < per) {
    const int i = pairs[2 * k], j = pairs[2 * k + 1];
    for (int64_t n = (int64_t)blockIdx.x * per + off; n < N; n += (int64_t)gridDim.x * per)
      s += (w ? w[n] : 1.0) * X[n * B + i] * Y[n * B + j];
  }
  red[tid] = s;
  __syncthreads();
  if (tid < np) {
    double t = 0.0;
    for (int m = 0; m < per; ++m) t += red[m * np + tid];
    partial[(int64_t)blockIdx.x * np + tid] = t;
  }
}

// Attainable FMA rate: `CH` independent multiply-add chains per thread,
// `iters` rounds, one store per thread (out holds one value per thread).
template <typename T, int CH>
__global__ __launch_bounds__(kThreads) void k_fma_probe(T* __restrict__ out, T a, T b, int iters) {
  T acc[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) acc[c] = (T)(threadIdx.x + c);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = acc[c] * a + b;
  }
  T s = 0;
#pragma unroll
  for (int c = 0; c < CH; ++c) s += acc[c];
  out[(size_t)blockIdx.x * kThreads + threadIdx.x] = s;
}

inline int grid_for(int64_t n) {
  int64_t g = (n + kThreads - 1) / kThreads;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

// ============================================================== launchers
int64_t hamiltonian_tile_elems(int p) { return (int64_t)ntiles_of(p) * 256; }

void launch_build_hamiltonian(const Geo& g, const DevTables& t, double hh, const double* veff,
                              const double* jxw, double* h64, float* h32, hipStream_t s) {
  dim3 grid((unsigned)g.ncell), block(kThreads);
  switch (g.p) {
    case 2: k_build_h<2><<<grid, block, 0, s>>>(g, t.Lq.p, t.M1.p, t.D1.p, hh, veff, jxw, h64, h32); break;
    case 3: k_build_h<3><<<grid, block, 0, s>>>(g, t.Lq.p, t.M1.p, t.D1.p, hh, veff, jxw, h64, h32); break;
    case 4: k_build_h<4><<<grid, block, 0, s>>>(g, t.Lq.p, t.M1.p, t.D1.p, hh, veff, jxw, h64, h32); break;
    case 5: k_build_h<5><<<grid, block, 0, s>>>(g, t.Lq.p, t.M1.p, t.D1.p, hh, veff, jxw, h64, h32); break;
    default: throw Error(IDFT_EINVAL, "degree not supported by the HIP kernels (2..5)");
  }
  IDFT_HIP(hipGetLastError());
}

// Columns per work-group (CG) and per item (VC) of the sum-factorised
// kernel built into the library: CG in {2, 4, 8} (its LDS holds 684 CG
// values at p = 5, 636 CG in FP64: 40.7 KB at CG = 8), one column per item.  The kernel is
// written for VC columns per item; wider items and 16-column groups measured
// slower on every mesh tried and are not instantiated.
bool sf_cols_supported(int cols, bool) { return cols == 2 || cols == 4 || cols == 8; }
bool sf_vc_supported(int cols, int vc, bool fp64) {
  return sf_cols_supported(cols, fp64) && vc == 1;
}

static const double* sf_eo_table(int p);

template <typename T, typename TV, int P, int CG>
static void cell_sf_cg(const Geo& g, const CellOp<T>& op, const T* S, const TV* X,
                       const double* xn, const double* xcol, T* Yc, int B, double* dotp,
                       hipStream_t s) {
  check_u32(g.N * B);
  check_u32(cell_scratch_elems(g, B));
  const int64_t npair = g.ncell * ((B + CG - 1) / CG);
  check_u32(8 * ((npair + 7) / 8));
  dim3 grid((unsigned)(8 * ((npair + 7) / 8))), block(kThreads);
  const T hq = (T)op.kin_scale;
  const double* eo = sf_eo_table(P);
  // per-cell dots only with node-scaled inputs (the MINRES steps)
  if (dotp && !xn) throw Error(IDFT_EINVAL, "per-cell dots need a node-scaled input");
  // node-scaled inputs (MINRES) come as FP64 vectors; FP32 vectors never
  // carry a node scale
  if constexpr (std::is_same<TV, double>::value) {
    if (xn) {
      k_cell_sf<T, TV, P, CG, 1, true><<<grid, block, 0, s>>>(g, hq, op.W, S, X, xn, xcol, Yc, B,
                                                               dotp, eo);
      return;
    }
  } else {
    if (xn) throw Error(IDFT_EINVAL, "node-scaled input needs FP64 vectors");
  }
  k_cell_sf<T, TV, P, CG, 1, false><<<grid, block, 0, s>>>(g, hq, op.W, S, X, nullptr, xcol, Yc, B,
                                                             nullptr, eo);
}

template <typename T, typename TV, int P>
static void cell_sf_p(const Geo& g, const CellOp<T>& op, const T* S, const TV* X,
                      const double* xn, const double* xcol, T* Yc, int B, int cg, int vc,
                      double* dotp, hipStream_t s) {
  if (!op.W) throw Error(IDFT_ESTATE, "quadrature weights of the cell operator not set");
  if (!sf_vc_supported(cg, vc, std::is_same<T, double>::value))
    throw Error(IDFT_EINVAL, "shape not supported by the sum-factorised kernel");
  switch (cg) {
    case 2: cell_sf_cg<T, TV, P, 2>(g, op, S, X, xn, xcol, Yc, B, dotp, s); break;
    case 4: cell_sf_cg<T, TV, P, 4>(g, op, S, X, xn, xcol, Yc, B, dotp, s); break;
    case 8: cell_sf_cg<T, TV, P, 8>(g, op, S, X, xn, xcol, Yc, B, dotp, s); break;
    default: throw Error(IDFT_EINVAL, "columns per work-group not supported by the sum-factorised kernel");
  }
}

template <typename T, typename TV, int P>
static void cell_gemm_p(const Geo& g, const CellOp<T>& op, const T* S, const TV* X,
                        const double* xn, const double* xcol, T* Yc, int B,
                        const GemmShape& shape, hipStream_t s, double* dotp = nullptr) {
  if (shape.kernel == CK_SUM_FACTOR) {
    cell_sf_p<T, TV, P>(g, op, S, X, xn, xcol, Yc, B, shape.vectors_per_tile,
                        shape.cells_per_wg, dotp, s);
    return;
  }
  if (dotp) throw Error(IDFT_EINVAL, "per-cell dots need the sum-factorised cell product");
  const T* H = op.H;
  if (!H) throw Error(IDFT_ESTATE, "stored cell matrices not built");
  const int vt = std::min(shape.vectors_per_tile, ((B + 15) / 16) * 16);
  const int nct = std::max(1, std::min(5, vt / 16));
  const int cpw = std::max(1, shape.cells_per_wg);
  const int64_t nwg = (g.ncell + cpw - 1) / cpw;
  dim3 grid((unsigned)(8 * ((nwg + 7) / 8)), (unsigned)((B + 16 * nct - 1) / (16 * nct))),
      block(kThreads);
  if (shape.kernel == CK_TILE_PASS && nct <= (std::is_same<T, float>::value ? 2 : 1)) {
    const dim3 block2(kSymmThreads);
    if (nct == 1) k_cell_symm<T, TV, P, 1><<<grid, block2, 0, s>>>(g, H, S, X, xn, xcol, Yc, B, cpw);
    if constexpr (std::is_same<T, float>::value) {
      if (nct == 2)
        k_cell_symm<T, TV, P, 2><<<grid, block2, 0, s>>>(g, H, S, X, xn, xcol, Yc, B, cpw);
    }
    return;
  }
  switch (nct) {
    case 1: k_cell_gemm<T, TV, P, 1><<<grid, block, 0, s>>>(g, H, S, X, xn, xcol, Yc, B, cpw); break;
    case 2: k_cell_gemm<T, TV, P, 2><<<grid, block, 0, s>>>(g, H, S, X, xn, xcol, Yc, B, cpw); break;
    case 3: k_cell_gemm<T, TV, P, 3><<<grid, block, 0, s>>>(g, H, S, X, xn, xcol, Yc, B, cpw); break;
    case 4: k_cell_gemm<T, TV, P, 4><<<grid, block, 0, s>>>(g, H, S, X, xn, xcol, Yc, B, cpw); break;
    case 5: k_cell_gemm<T, TV, P, 5><<<grid, block, 0, s>>>(g, H, S, X, xn, xcol, Yc, B, cpw); break;
  }
}

// Device addresses of the even-odd tables kEo2..kEo5, per device, set by
// upload_sf_constants (once per device, before any cell product there).
static std::atomic<const double*> g_eo_table[64][4];

void upload_sf_constants() {
  static std::mutex mu;
  static bool done[64] = {};
  int dev = 0;
  IDFT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw Error(IDFT_EHIP, "device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (done[dev]) return;
  static constexpr EoTable<2> t2 = make_eo_table<2>();
  static constexpr EoTable<3> t3 = make_eo_table<3>();
  static constexpr EoTable<4> t4 = make_eo_table<4>();
  static constexpr EoTable<5> t5 = make_eo_table<5>();
  IDFT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kEo2), &t2, sizeof(t2), 0, hipMemcpyHostToDevice));
  IDFT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kEo3), &t3, sizeof(t3), 0, hipMemcpyHostToDevice));
  IDFT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kEo4), &t4, sizeof(t4), 0, hipMemcpyHostToDevice));
  IDFT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kEo5), &t5, sizeof(t5), 0, hipMemcpyHostToDevice));
  void* a[4] = {};
  IDFT_HIP(hipGetSymbolAddress(&a[0], HIP_SYMBOL(kEo2)));
  IDFT_HIP(hipGetSymbolAddress(&a[1], HIP_SYMBOL(kEo3)));
  IDFT_HIP(hipGetSymbolAddress(&a[2], HIP_SYMBOL(kEo4)));
  IDFT_HIP(hipGetSymbolAddress(&a[3], HIP_SYMBOL(kEo5)));
  for (int k = 0; k < 4; ++k) g_eo_table[dev][k].store(static_cast<const double*>(a[k]));
  done[dev] = true;
}

// The even-odd table of degree p on the current device.
static const double* sf_eo_table(int p) {
  int dev = 0;
  IDFT_HIP(hipGetDevice(&dev));
  const double* t = (dev >= 0 && dev < 64 && p >= 2 && p <= 5) ? g_eo_table[dev][p - 2].load()
                                                                : nullptr;
  if (!t) throw Error(IDFT_ESTATE, "sum-factorised tables not uploaded on this device");
  return t;
}

int64_t cell_scratch_elems(const Geo& g, int B) { return g.ncell * (int64_t)ne_of(g.p) * B; }

// Columns per gather thread: 16-byte vectors of the node-major rows where B
// allows it.
static int gather_vw(int B, int vbytes) {
  const int w = 16 / vbytes;
  if (B % w == 0) return w;
  return B % 2 == 0 ? 2 : 1;
}

// Gather grid: blocks of R chunks of kThreads / (B / VW) nodes each, with
// R = `rows_req` if positive, else chosen so that the grid has at most about
// kGatherBlocks blocks.
static void gather_grid(const Geo& g, int B, int vw, int rows_req, unsigned& gx, unsigned& gy,
                        uint32_t& R) {
  const int per = kThreads / (B / vw);
  const int64_t chunks = (g.N + per - 1) / per;
  int64_t r = rows_req > 0 ? rows_req : (chunks + kGatherBlocks - 1) / kGatherBlocks;
  if (r < 1) r = 1;
  R = (uint32_t)r;
  gx = (unsigned)((chunks + r - 1) / r);
  gy = 1;
}

int gather_blocks(const Geo& g, int B, int rows, int vbytes) {
  unsigned gx, gy;
  uint32_t R;
  gather_grid(g, B, gather_vw(B, vbytes), rows, gx, gy, R);
  return (int)(gx * gy);
}

template <typename T>
void launch_cell_gemm(const Geo& g, const CellOp<T>& op, const T* S, const T* X,
                      const double* xcol, T* Yc, int B, const GemmShape& shape, hipStream_t s) {
  if (B < 1 || B > kMaxCols) throw Error(IDFT_EINVAL, "block width out of range");
  switch (g.p) {
    case 2: cell_gemm_p<T, T, 2>(g, op, S, X, nullptr, xcol, Yc, B, shape, s); break;
    case 3: cell_gemm_p<T, T, 3>(g, op, S, X, nullptr, xcol, Yc, B, shape, s); break;
    case 4: cell_gemm_p<T, T, 4>(g, op, S, X, nullptr, xcol, Yc, B, shape, s); break;
    case 5: cell_gemm_p<T, T, 5>(g, op, S, X, nullptr, xcol, Yc, B, shape, s); break;
    default: throw Error(IDFT_EINVAL, "degree not supported by the HIP kernels (2..5)");
  }
  IDFT_HIP(hipGetLastError());
}
template void launch_cell_gemm<float>(const Geo&, const CellOp<float>&, const float*,
                                      const float*, const double*, float*, int,
                                      const GemmShape&, hipStream_t);
template void launch_cell_gemm<double>(const Geo&, const CellOp<double>&, const double*,
                                       const double*, const double*, double*, int,
                                       const GemmShape&, hipStream_t);

template <typename T, typename TV>
int launch_cell_apply(const Geo& g, const CellOp<T>& op, const T* Sg, const TV* S, const TV* X,
                      T* Yc, int B, const ApplyEpilogue<TV>& ep, const GemmShape& shape,
                      hipStream_t s) {
  if (B < 1 || B > kMaxCols) throw Error(IDFT_EINVAL, "block width out of range");
  switch (g.p) {
    case 2: cell_gemm_p<T, TV, 2>(g, op, Sg, X, ep.xn, ep.xcol, Yc, B, shape, s); break;
    case 3: cell_gemm_p<T, TV, 3>(g, op, Sg, X, ep.xn, ep.xcol, Yc, B, shape, s); break;
    case 4: cell_gemm_p<T, TV, 4>(g, op, Sg, X, ep.xn, ep.xcol, Yc, B, shape, s); break;
    case 5: cell_gemm_p<T, TV, 5>(g, op, Sg, X, ep.xn, ep.xcol, Yc, B, shape, s); break;
    default: throw Error(IDFT_EINVAL, "degree not supported by the HIP kernels (2..5)");
  }
  IDFT_HIP(hipGetLastError());
  check_u32(g.N * B);
  check_u32(cell_scratch_elems(g, B));
  const FastDiv divN1 = make_fastdiv((uint32_t)g.N1);
  unsigned gx, gy;
  uint32_t R;
  const int vw = gather_vw(B, (int)sizeof(TV));
  gather_grid(g, B, vw, shape.gather_rows, gx, gy, R);
  dim3 grid(gx, gy);
#define IDFT_GATHER(PP, VW)                                                                  \
  k_cell_gather<T, TV, PP, VW><<<grid, kThreads, 0, s>>>(                                    \
      g, divN1, (uint32_t)B, R, Yc, S, ep.Y, ep.alpha, ep.beta, ep.X, ep.a, ep.ca, ep.Z, ep.b,  \
      ep.cb,                                                                                  \
      ep.W, ep.c, ep.cc, X, ep.xn, ep.xcol, ep.V, ep.dot_partial)
#define IDFT_GATHER_P(PP)                                         \
  if (vw == 4) { if constexpr (sizeof(TV) == 4) IDFT_GATHER(PP, 4); } \
  else if (vw == 2) IDFT_GATHER(PP, 2);                           \
  else IDFT_GATHER(PP, 1);
  switch (g.p) {
    case 2: IDFT_GATHER_P(2); break;
    case 3: IDFT_GATHER_P(3); break;
    case 4: IDFT_GATHER_P(4); break;
    case 5: IDFT_GATHER_P(5); break;
  }
#undef IDFT_GATHER_P
#undef IDFT_GATHER
  IDFT_HIP(hipGetLastError());
  return (int)(gx * gy);
}
template int launch_cell_apply<float, float>(const Geo&, const CellOp<float>&, const float*,
                                             const float*, const float*, float*, int,
                                             const ApplyEpilogue<float>&, const GemmShape&,
                                             hipStream_t);
template int launch_cell_apply<double, double>(const Geo&, const CellOp<double>&,
                                               const double*, const double*, const double*,
                                               double*, int,
                                               const ApplyEpilogue<double>&, const GemmShape&,
                                               hipStream_t);
template int launch_cell_apply<float, double>(const Geo&, const CellOp<float>&, const float*,
                                              const double*, const double*, float*, int,
                                              const ApplyEpilogue<double>&, const GemmShape&,
                                              hipStream_t);

template <typename T, typename TV>
void launch_cell_dot(const Geo& g, const CellOp<T>& op, const T* Sg, const TV* X, int B,
                     const double* xn, const double* xcol, T* Yc, double* cell_dots,
                     const GemmShape& shape, hipStream_t s) {
  if (B < 1 || B > kMaxCols) throw Error(IDFT_EINVAL, "block width out of range");
  if (shape.kernel != CK_SUM_FACTOR || !cell_dots)
    throw Error(IDFT_EINVAL, "per-cell dots need the sum-factorised cell product");
  switch (g.p) {
    case 2: cell_gemm_p<T, TV, 2>(g, op, Sg, X, xn, xcol, Yc, B, shape, s, cell_dots); break;
    case 3: cell_gemm_p<T, TV, 3>(g, op, Sg, X, xn, xcol, Yc, B, shape, s, cell_dots); break;
    case 4: cell_gemm_p<T, TV, 4>(g, op, Sg, X, xn, xcol, Yc, B, shape, s, cell_dots); break;
    case 5: cell_gemm_p<T, TV, 5>(g, op, Sg, X, xn, xcol, Yc, B, shape, s, cell_dots); break;
    default: throw Error(IDFT_EINVAL, "degree not supported by the HIP kernels (2..5)");
  }
  IDFT_HIP(hipGetLastError());
}
template void launch_cell_dot<double, double>(const Geo&, const CellOp<double>&, const double*,
                                              const double*, int, const double*, const double*,
                                              double*, double*, const GemmShape&, hipStream_t);
template void launch_cell_dot<float, double>(const Geo&, const CellOp<float>&, const float*,
                                             const double*, int, const double*, const double*,
                                             float*, double*, const GemmShape&, hipStream_t);

template <typename TY>
int launch_minres_gather(const Geo& g, const TY* Yc, const double* S, const double* D,
                         const double* r2, double* r1, int B, const double* st, double* partial,
                         int gather_rows, hipStream_t s, double* w, const double* w1, double* x) {
  if (B < 1 || B > kMaxCols) throw Error(IDFT_EINVAL, "block width out of range");
  check_u32(g.N * B);
  check_u32(cell_scratch_elems(g, B));
  const FastDiv divN1 = make_fastdiv((uint32_t)g.N1);
  unsigned gx, gy;
  uint32_t R;
  const int vw = gather_vw(B, (int)sizeof(double));
  gather_grid(g, B, vw, gather_rows, gx, gy, R);
  gx = std::min<unsigned>(gx, (unsigned)kMinresGatherBlocks);
#define IDFT_MG(PP, VW)                                                                   \
  do {                                                                                     \
    if (w)                                                                                 \
      k_minres_gather<TY, double, PP, VW, true><<<dim3(gx), kThreads, 0, s>>>(             \
          g, divN1, (uint32_t)B, R, Yc, S, D, r2, r1, st, partial, w, w1, x);              \
    else                                                                                   \
      k_minres_gather<TY, double, PP, VW, false><<<dim3(gx), kThreads, 0, s>>>(            \
          g, divN1, (uint32_t)B, R, Yc, S, D, r2, r1, st, partial, nullptr, nullptr, nullptr); \
  } while (0)
#define IDFT_MG_P(PP)   \
  if (vw == 2) IDFT_MG(PP, 2); \
  else IDFT_MG(PP, 1);
  switch (g.p) {
    case 2: IDFT_MG_P(2); break;
    case 3: IDFT_MG_P(3); break;
    case 4: IDFT_MG_P(4); break;
    case 5: IDFT_MG_P(5); break;
    default: throw Error(IDFT_EINVAL, "degree not supported by the HIP kernels (2..5)");
  }
#undef IDFT_MG_P
#undef IDFT_MG
  IDFT_HIP(hipGetLastError());
  return (int)gx;
}
template int launch_minres_gather<double>(const Geo&, const double*, const double*, const double*,
                                          const double*, double*, int, const double*, double*, int,
                                          hipStream_t, double*, const double*, double*);
template int launch_minres_gather<float>(const Geo&, const float*, const double*, const double*,
                                         const double*, double*, int, const double*, double*, int,
                                         hipStream_t, double*, const double*, double*);

template <typename T>
void launch_lincomb(int64_t N, int B, T* Y, double beta, const T* X, double a, const double* ca,
                    const T* Z, double b, const double* cb, const double* ns, hipStream_t s) {
  check_u32(N * B);
  k_lincomb<T><<<grid_for(N * B), kThreads, 0, s>>>((uint32_t)(N * B), (uint32_t)B, Y, beta, X, a,
                                                    ca, Z, b, cb, nullptr, 0.0, nullptr, ns);
  IDFT_HIP(hipGetLastError());
}
template <typename T>
void launch_lincomb3(int64_t N, int B, T* Y, double beta, const T* X, double a, const double* ca,
                     const T* Z, double b, const double* cb, const T* W, double c,
                     const double* cc, const double* ns, hipStream_t s) {
  check_u32(N * B);
  k_lincomb<T><<<grid_for(N * B), kThreads, 0, s>>>((uint32_t)(N * B), (uint32_t)B, Y, beta, X, a,
                                                    ca, Z, b, cb, W, c, cc, ns);
  IDFT_HIP(hipGetLastError());
}
template void launch_lincomb<float>(int64_t, int, float*, double, const float*, double,
                                    const double*, const float*, double, const double*,
                                    const double*, hipStream_t);
template void launch_lincomb<double>(int64_t, int, double*, double, const double*, double,
                                     const double*, const double*, double, const double*,
                                     const double*, hipStream_t);
template void launch_lincomb3<float>(int64_t, int, float*, double, const float*, double,
                                     const double*, const float*, double, const double*,
                                     const float*, double, const double*, const double*,
                                     hipStream_t);
template void launch_lincomb3<double>(int64_t, int, double*, double, const double*, double,
                                      const double*, const double*, double, const double*,
                                      const double*, double, const double*, const double*,
                                      hipStream_t);

void launch_lincomb_d2f(int64_t N, int B, float* Y, const double* X, double a, const double* ca,
                        const double* Z, double b, const double* cb, hipStream_t s) {
  check_u32(N * B);
  k_lincomb_mixed<float, double, double><<<grid_for(N * B), kThreads, 0, s>>>(
      (uint32_t)(N * B), (uint32_t)B, Y, X, a, ca, Z, b, cb);
  IDFT_HIP(hipGetLastError());
}
void launch_lincomb_df2d(int64_t N, int B, double* Y, const double* X, double a,
                         const double* ca, const float* Z, double b, const double* cb,
                         hipStream_t s) {
  check_u32(N * B);
  k_lincomb_mixed<double, double, float><<<grid_for(N * B), kThreads, 0, s>>>(
      (uint32_t)(N * B), (uint32_t)B, Y, X, a, ca, Z, b, cb);
  IDFT_HIP(hipGetLastError());
}
void launch_convert_d2f(int64_t n, float* y, const double* x, hipStream_t s) {
  k_d2f<<<grid_for(n), kThreads, 0, s>>>(n, y, x);
  IDFT_HIP(hipGetLastError());
}
void launch_convert_f2d(int64_t n, double* y, const float* x, hipStream_t s) {
  k_f2d<<<grid_for(n), kThreads, 0, s>>>(n, y, x);
  IDFT_HIP(hipGetLastError());
}

int coldot_blocks(int64_t N, int B) {
  const int per = kThreads / B;
  int64_t g = (N + per - 1) / per;
  if (g > 1024) g = 1024;
  return (int)(g < 1 ? 1 : g);
}

template <typename T>
int launch_coldot_partial(int64_t N, int B, const T* X, const T* Y, const double* w,
                          double* partial, hipStream_t s) {
  if (B > kThreads) throw Error(IDFT_EINVAL, "block too wide for column reductions");
  const int nb = coldot_blocks(N, B);
  k_coldot<T><<<nb, kThreads, 0, s>>>(N, B, X, Y, w, partial);
  IDFT_HIP(hipGetLastError());
  return nb;
}
template int launch_coldot_partial<float>(int64_t, int, const float*, const float*, const double*,
                                          double*, hipStream_t);
template int launch_coldot_partial<double>(int64_t, int, const double*, const double*,
                                           const double*, double*, hipStream_t);

void launch_sum_partials(int nblk, int B, const double* partial, double* out, hipStream_t s) {
  k_sum_partials<<<B, kThreads, 0, s>>>(nblk, B, partial, out);
  IDFT_HIP(hipGetLastError());
}

// Nodes per LDS tile of the Gram kernel: 64 for narrow blocks, about 2048
// entries per tile for wide ones.
static int tall_tile_rows(int B) { return std::max(8, std::min(64, 2048 / std::max(B, 1))); }

// Gram partial rows: enough work-groups to fill the CUs (up to 2048, 8 per
// CU), at most 2^21 partial values.
int gram_partial_blocks(int64_t N, int B) {
  const int rows = tall_tile_rows(B);
  const int64_t cap = std::max<int64_t>(64, std::min<int64_t>(2048, (int64_t(1) << 21) / ((int64_t)B * B)));
  const int64_t g = std::min<int64_t>((N + rows - 1) / rows, cap);
  return (int)(g < 1 ? 1 : g);
}

void launch_gram(int64_t N, int B, const double* X, const double* Y, double* partial, double* G,
                 hipStream_t s) {
  if (B < 1 || B > kMaxCols) throw Error(IDFT_EINVAL, "block width out of range for the Gram kernel");
  const int nb = gram_partial_blocks(N, B), rows = tall_tile_rows(B);
  dim3 grid(nb, (B * B + kThreads - 1) / kThreads);
  k_gram<<<grid, kThreads, 2 * rows * B * sizeof(double), s>>>(N, B, rows, X, Y, partial);
  IDFT_HIP(hipGetLastError());
  k_sum_rows<<<B * B, kRedThreads, 0, s>>>(nb, partial, G);
  IDFT_HIP(hipGetLastError());
}

void launch_right_mult(int64_t N, int B, const double* X, const double* Q, double* Y,
                       const double* ns, double beta, hipStream_t s) {
  if (B < 1 || B > kMaxCols) throw Error(IDFT_EINVAL, "block width out of range");
  check_u32(N * B);
  // Q and the tile within 64 KB of LDS (8 nodes per tile at the widest block)
  const int rows = std::max(8, std::min(64, (8192 - B * B) / B));
  const int64_t tiles = (N + rows - 1) / rows;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, 4096));
  k_right_mult<<<grid, kThreads, ((size_t)B * B + (size_t)rows * B) * sizeof(double), s>>>(
      (uint32_t)N, (uint32_t)B, (uint32_t)rows, X, Q, Y, ns, beta);
  IDFT_HIP(hipGetLastError());
}

int launch_residual_partial(int64_t N, int B, const double* HX, const double* X,
                            const double* ev, double* partial, hipStream_t s) {
  const int nb = coldot_blocks(N, B);
  k_residual<double><<<nb, kThreads, 0, s>>>(N, B, HX, X, ev, partial);
  IDFT_HIP(hipGetLastError());
  return nb;
}

void launch_vxc_to_quad(const Geo& g, const DevTables& t, const double* pot,
                        const double* potBase, double* veff, hipStream_t s) {
  dim3 grid((unsigned)g.ncell), block(kThreads);
  switch (g.p) {
    case 2: k_vxc_to_quad<2><<<grid, block, 0, s>>>(g, t.Wq.p, pot, potBase, veff); break;
    case 3: k_vxc_to_quad<3><<<grid, block, 0, s>>>(g, t.Wq.p, pot, potBase, veff); break;
    case 4: k_vxc_to_quad<4><<<grid, block, 0, s>>>(g, t.Wq.p, pot, potBase, veff); break;
    case 5: k_vxc_to_quad<5><<<grid, block, 0, s>>>(g, t.Wq.p, pot, potBase, veff); break;
    default: throw Error(IDFT_EINVAL, "degree not supported");
  }
  IDFT_HIP(hipGetLastError());
}

void launch_density(const Geo& g, const DevTables& t, const double* X, const double* S, int B,
                    const double* occ, double* rho, hipStream_t s) {
  dim3 grid((unsigned)g.ncell), block(kThreads);
  switch (g.p) {
    case 2: k_density<2, 8><<<grid, block, 0, s>>>(g, t.Lq.p, X, S, B, occ, rho); break;
    case 3: k_density<3, 8><<<grid, block, 0, s>>>(g, t.Lq.p, X, S, B, occ, rho); break;
    case 4: k_density<4, 8><<<grid, block, 0, s>>>(g, t.Lq.p, X, S, B, occ, rho); break;
    case 5: k_density<5, 8><<<grid, block, 0, s>>>(g, t.Lq.p, X, S, B, occ, rho); break;
    default: throw Error(IDFT_EINVAL, "degree not supported");
  }
  IDFT_HIP(hipGetLastError());
}

void launch_loss(int64_t nQ, int nq, const double* rho, const double* rhoT, const double* w,
                 const double* veff, const double* jxw, double* u, double* partial,
                 hipStream_t s) {
  k_loss<<<256, kThreads, 0, s>>>(nQ, nq, rho, rhoT, w, veff, jxw, u, partial);
  IDFT_HIP(hipGetLastError());
}

void launch_pair_quad(const Geo& g, const DevTables& t, const double* A, const double* Bv, int B,
                      const int* pairs, int np, const double* wq, int64_t wstride, double* out,
                      hipStream_t s) {
  if (np == 0) return;
  dim3 grid((unsigned)g.ncell), block(kThreads);
  switch (g.p) {
    case 2: k_pair_quad<2, 8><<<grid, block, 0, s>>>(g, t.Lq.p, A, Bv, B, pairs, np, wq, wstride, out); break;
    case 3: k_pair_quad<3, 8><<<grid, block, 0, s>>>(g, t.Lq.p, A, Bv, B, pairs, np, wq, wstride, out); break;
    case 4: k_pair_quad<4, 8><<<grid, block, 0, s>>>(g, t.Lq.p, A, Bv, B, pairs, np, wq, wstride, out); break;
    case 5: k_pair_quad<5, 8><<<grid, block, 0, s>>>(g, t.Lq.p, A, Bv, B, pairs, np, wq, wstride, out); break;
    default: throw Error(IDFT_EINVAL, "degree not supported");
  }
  IDFT_HIP(hipGetLastError());
}

void launch_quad_operator(const Geo& g, const DevTables& t, const double* X, int B,
                          const double* wq, double* Y, int colour, hipStream_t s) {
  const int64_t nw = colour_count(g, colour);
  if (nw == 0) return;
  dim3 grid(colour_grid(nw)), block(kThreads);
  switch (g.p) {
    case 2: k_quad_operator<2, 8><<<grid, block, 0, s>>>(g, t.Lq.p, X, B, wq, Y, colour, nw); break;
    case 3: k_quad_operator<3, 8><<<grid, block, 0, s>>>(g, t.Lq.p, X, B, wq, Y, colour, nw); break;
    case 4: k_quad_operator<4, 8><<<grid, block, 0, s>>>(g, t.Lq.p, X, B, wq, Y, colour, nw); break;
    case 5: k_quad_operator<5, 8><<<grid, block, 0, s>>>(g, t.Lq.p, X, B, wq, Y, colour, nw); break;
    default: throw Error(IDFT_EINVAL, "degree not supported");
  }
  IDFT_HIP(hipGetLastError());
}

template <int P>
static void force_p(const Geo& g, const DevTables& t, const double* Psi, const double* Pv, int B,
                    int mode, double jxwChild, const double* jxwParent, double* force,
                    int colour, int64_t nw, hipStream_t s) {
  dim3 grid(colour_grid(nw)), block(kThreads);
  if (mode == 0)
    k_force<P, 4, 8><<<grid, block, 0, s>>>(g, t.Lc.p, t.Wc.p, Psi, Pv, B, jxwChild, nullptr,
                                            force, colour, nw);
  else
    k_force<P, P + 3, 8><<<grid, block, 0, s>>>(g, t.Lq.p, t.Wq.p, Psi, Pv, B, 0.0, jxwParent,
                                                force, colour, nw);
}

void launch_force(const Geo& g, const DevTables& t, const double* Psi, const double* P, int B,
                  int mode, double jxwChild, const double* jxwParent, double* force, int colour,
                  hipStream_t s) {
  const int64_t nw = colour_count(g, colour);
  if (nw == 0) return;
  switch (g.p) {
    case 2: force_p<2>(g, t, Psi, P, B, mode, jxwChild, jxwParent, force, colour, nw, s); break;
    case 3: force_p<3>(g, t, Psi, P, B, mode, jxwChild, jxwParent, force, colour, nw, s); break;
    case 4: force_p<4>(g, t, Psi, P, B, mode, jxwChild, jxwParent, force, colour, nw, s); break;
    case 5: force_p<5>(g, t, Psi, P, B, mode, jxwChild, jxwParent, force, colour, nw, s); break;
    default: throw Error(IDFT_EINVAL, "degree not supported");
  }
  IDFT_HIP(hipGetLastError());
}

void launch_zero_rows(int64_t N, int B, double* X, const unsigned char* mask, hipStream_t s) {
  check_u32(N * B);
  k_zero_rows<<<grid_for(N * B), kThreads, 0, s>>>((uint32_t)(N * B), (uint32_t)B, X, mask);
  IDFT_HIP(hipGetLastError());
}


void launch_minres_start(int B, int nblk, const double* partial, const double* negEig, double* st,
                         hipStream_t s, int nq) {
  k_minres_start<<<B, kThreads, 0, s>>>(B, nblk, partial, negEig, st, nq);
  IDFT_HIP(hipGetLastError());
}
int launch_minres_start_dots(int64_t N, int B, const double* r, const double* D, double* partial,
                             hipStream_t s) {
  if (B > kThreads) throw Error(IDFT_EINVAL, "block too wide for column reductions");
  const int nb = coldot_blocks(N, B);
  k_minres_start_dots<<<nb, kThreads, 0, s>>>(N, B, r, D, partial);
  IDFT_HIP(hipGetLastError());
  return nb;
}
void launch_minres_alpha_cell(int B, int ncell, const double* cell_dots, double* st, hipStream_t s) {
  k_minres_alpha_cell<<<B, kRedThreads, 0, s>>>(B, ncell, cell_dots, st);
  IDFT_HIP(hipGetLastError());
}
void launch_minres_alpha(int B, int nblk, const double* partial, double* st, hipStream_t s) {
  k_minres_alpha<<<B, kThreads, 0, s>>>(B, nblk, partial, st);
  IDFT_HIP(hipGetLastError());
}
template <typename T>
int launch_minres_lanczos(int64_t N, int B, const T* y, const T* r2, T* rnew, const double* D,
                          const double* st, double* partial, hipStream_t s) {
  const int nb = coldot_blocks(N, B);
  k_minres_lanczos<T><<<nb, kThreads, 0, s>>>(N, B, y, r2, rnew, D, st, partial);
  IDFT_HIP(hipGetLastError());
  return nb;
}
template int launch_minres_lanczos<float>(int64_t, int, const float*, const float*, float*,
                                          const double*, const double*, double*, hipStream_t);
template int launch_minres_lanczos<double>(int64_t, int, const double*, const double*, double*,
                                           const double*, const double*, double*, hipStream_t);
void launch_minres_givens(int B, int nblk, const double* partial, const double* negEig, double* st,
                          hipStream_t s, int nq) {
  k_minres_givens<<<B, kRedThreads, 0, s>>>(B, nblk, partial, negEig, st, nq);
  IDFT_HIP(hipGetLastError());
}
template <typename T>
void launch_minres_update(int64_t N, int B, const T* r, const double* D, T* w1, const T* w, T* x,
                          const double* st, hipStream_t s) {
  check_u32(N * B);
  if (B % 2 == 0)
    k_minres_update<T, 2><<<grid_for(N * B / 2), kThreads, 0, s>>>((uint32_t)(N * B), (uint32_t)B,
                                                                  r, D, w1, w, x, st);
  else
    k_minres_update<T, 1><<<grid_for(N * B), kThreads, 0, s>>>((uint32_t)(N * B), (uint32_t)B, r,
                                                               D, w1, w, x, st);
  IDFT_HIP(hipGetLastError());
}
template void launch_minres_update<float>(int64_t, int, const float*, const double*, float*,
                                          const float*, float*, const double*, hipStream_t);
template void launch_minres_update<double>(int64_t, int, const double*, const double*, double*,
                                           const double*, double*, const double*, hipStream_t);
void launch_copy_done(int B, const double* st, int* flags, hipStream_t s) {
  k_copy_done<<<(B + 63) / 64, 64, 0, s>>>(B, st, flags);
  IDFT_HIP(hipGetLastError());
}


void launch_copy_cols(int64_t N, int Bs, int off, int Bd, const double* src, double* dst,
                      hipStream_t s) {
  k_copy_cols<<<grid_for(N * Bd), kThreads, 0, s>>>(N, Bs, off, Bd, src, dst);
  IDFT_HIP(hipGetLastError());
}
void launch_gather_cols(int64_t N, int Bs, const double* src, int Bd, const int* idx, double* dst,
                        hipStream_t s) {
  k_cols_by_index<<<grid_for(N * Bd), kThreads, 0, s>>>(N, Bs, Bd, Bd, idx, src, dst, 0);
  IDFT_HIP(hipGetLastError());
}
void launch_scatter_cols(int64_t N, int Bs, const double* src, int Bd, const int* idx, double* dst,
                         hipStream_t s) {
  k_cols_by_index<<<grid_for(N * Bs), kThreads, 0, s>>>(N, Bs, Bd, Bs, idx, src, dst, 1);
  IDFT_HIP(hipGetLastError());
}
void launch_mul_q(int64_t nQ, int nq, const double* u, const double* jxw, double* out,
                  hipStream_t s) {
  k_mul_q<<<grid_for(nQ), kThreads, 0, s>>>(nQ, nq, u, jxw, out);
  IDFT_HIP(hipGetLastError());
}
int launch_pairdot_partial(int64_t N, int B, const double* X, const double* Y, const double* w,
                           const int* pairs, int np, double* partial, hipStream_t s) {
  if (np > kThreads) throw Error(IDFT_EINVAL, "too many degenerate pairs");
  const int nb = coldot_blocks(N, np);
  k_pairdot<<<nb, kThreads, 0, s>>>(N, B, X, Y, w, pairs, np, partial);
  IDFT_HIP(hipGetLastError());
  return nb;
}


// ------------------------------------------------------ FMA rate probe
template <typename T, int CH>
static double fma_probe_t(hipStream_t s) {
  int dev = 0, ncu = 0;
  IDFT_HIP(hipGetDevice(&dev));
  IDFT_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = std::max(1, ncu) * 32, iters = 2048;
  T* out = nullptr;
  IDFT_HIP(hipMalloc(&out, sizeof(T) * (size_t)blocks * kThreads));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.0f;
  hipError_t err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  if (err == hipSuccess) {
    k_fma_probe<T, CH><<<blocks, kThreads, 0, s>>>(out, (T)0.999, (T)0.001, iters);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipEventRecord(e0, s);
  if (err == hipSuccess) {
    k_fma_probe<T, CH><<<blocks, kThreads, 0, s>>>(out, (T)0.999, (T)0.001, iters);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipEventRecord(e1, s);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  IDFT_HIP(err);
  const double flops = 2.0 * CH * (double)iters * blocks * kThreads;
  return flops / ((double)ms * 1e-3) / 1e12;
}
double launch_fma_probe(bool fp64, int chains, hipStream_t s) {
  switch (chains) {
    case 4: return fp64 ? fma_probe_t<double, 4>(s) : fma_probe_t<float, 4>(s);
    case 8: return fp64 ? fma_probe_t<double, 8>(s) : fma_probe_t<float, 8>(s);
    case 16: return fp64 ? fma_probe_t<double, 16>(s) : fma_probe_t<float, 16>(s);
    default: throw Error(IDFT_EINVAL, "chains must be 4, 8 or 16");
  }
}

}  // namespace idft
