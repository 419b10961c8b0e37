// Scratch memory of the allocation-free entry point: a file-scope device array of fixed size in noalloc.hip.
#pragma once

#include <cstddef>

#include "layout.hpp"
#include "mpsxx_amd.h"

namespace mpx {

// Size of the per-device scratch array (16 MiB): segment list, chunk list, partial sums and result of one call.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;

// Resolve the scratch array's device address for `device`, which must be the current device (once per device; later
// calls return at once).  Throws Error(MPSXX_ERR_HIP) if hipGetSymbolAddress fails.
void noalloc_init(int device);
// Device address of the current device's scratch array (resolved on first use if noalloc_init was not called).
void* noalloc_scratch();

}  // namespace mpx
