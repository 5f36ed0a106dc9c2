// Fixed device scratch for the allocation-free entry point
// (idft_adj_vmult_noalloc): the cell-product scratch of one operator
// application lives in a file-scope device array instead of a buffer the
// library allocates.  Each GPU the library runs on holds its own copy of the
// array, at its own address, until the process ends.
//
// On the host the array's name is only a stand-in for the device symbol, so
// host code never passes it to a kernel or a copy: the device address is
// looked up once per device (hipGetSymbolAddress, with that device current)
// and kept here, and every use goes through noalloc_scratch().
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.h"

namespace {

constexpr int kMaxDevices = 64;

// kNoallocScratchElems doubles (64 MiB), 256-byte aligned.
__device__ __align__(256) double scratch[idft::kNoallocScratchElems];

std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

}  // namespace

namespace idft {

// Resolves (once) and returns the device address of `scratch` on the current
// device.  Throws idft::Error on a HIP failure.
double* noalloc_scratch() {
  int dev = 0;
  IDFT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error(IDFT_EHIP, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    IDFT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(IDFT_EHIP, "hipGetSymbolAddress returned no address");
    g_ptr[dev] = static_cast<double*>(p);
  }
  return g_ptr[dev];
}

}  // namespace idft
