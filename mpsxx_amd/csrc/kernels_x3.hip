// Instantiation of the split-BF16 (x3) grouped GEMM tiles (gemm_x3.hpp), 64x64 and 128x128, FP32 and FP64 C.
#include "gemm_x3.hpp"

namespace mpx {

#define MPX_X3_INST(FN, TC)                                                                                        \
  template void FN<TC>(const BigTile*, const int (&)[4], const RedTile*, int, void*, const GemmTask*,               \
                       const GemmPair*, Operands, double, int, hipStream_t, int);
MPX_X3_INST(launch_gemm_x3, float)
MPX_X3_INST(launch_gemm_x3, double)
MPX_X3_INST(launch_gemm_x3_128, float)
MPX_X3_INST(launch_gemm_x3_128, double)
#undef MPX_X3_INST

}  // namespace mpx
