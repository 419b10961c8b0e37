// Device scratch memory of the allocation-free entry point (mpsxx_dotc_noalloc): one statically sized array that
// every device holds from module load to process end.  The host resolves its address once per device with
// hipGetSymbolAddress (the array's name is only a host-side stand-in) and hands that address to the kernels.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "noalloc.hpp"

namespace mpx {

namespace {

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

void noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(MPSXX_ERR_ARG, "noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> g(g_mu);
  if (g_ptr[device]) return;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess || cur != device)
    throw Error(MPSXX_ERR_HIP, "noalloc: scratch set-up needs its device current");
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw Error(MPSXX_ERR_HIP, std::string("noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_ptr[device] = p;
}

void* noalloc_scratch() {
  int dev = -1;
  const hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) throw Error(MPSXX_ERR_HIP, std::string("noalloc: hipGetDevice: ") + hipGetErrorString(e));
  noalloc_init(dev);
  std::lock_guard<std::mutex> g(g_mu);
  return g_ptr[dev];
}

}  // namespace mpx
