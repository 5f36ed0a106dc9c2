// Prints the compile-time 1D tables of the sum-factorised cell product
// (invdft_amd/csrc/sf_tables.h) for degrees 2..5, one table per line:
//   <p> L|M|D <values...>
// Built and run by tests/test_sf_tables.py (host compiler only).
#include <cstdio>

#include "sf_tables.h"

template <int P>
static void dump() {
  constexpr auto& t = idft::SfC<P>::v;
  constexpr int N1 = P + 1, Q = P + 3;
  std::printf("%d L", P);
  for (int i = 0; i < Q * N1; ++i) std::printf(" %.17g", t.L[i]);
  std::printf("\n%d M", P);
  for (int i = 0; i < N1 * N1; ++i) std::printf(" %.17g", t.M[i]);
  std::printf("\n%d D", P);
  for (int i = 0; i < N1 * N1; ++i) std::printf(" %.17g", t.D[i]);
  std::printf("\n");
}

int main() {
  dump<2>();
  dump<3>();
  dump<4>();
  dump<5>();
  return 0;
}
