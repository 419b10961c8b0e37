// FP64 instantiation of the 64x64 grouped GEMM (gemm_tiles.hpp).  Built with -mllvm -amdgpu-mfma-vgpr-form: the
// FP64 tile loop otherwise keeps its accumulators in VGPRs and copies them into AGPRs around every k panel (measured
// on the 24-orbital M = 1024 merged sigma: 42.4 -> 38.3 ms).  The FP32 and split-BF16 kernels are faster without
// the option, so they stay in kernels.hip.
#include "gemm_tiles.hpp"

namespace mpx {

template void launch_gemm<double, double>(const BigTile*, const int (&)[4], const RedTile*, int, void*, const GemmTask*,
                                          const GemmPair*, Operands, double, int, hipStream_t, int);

}  // namespace mpx
