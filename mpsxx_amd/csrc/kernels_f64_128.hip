// FP64 instantiation of the 128x128 grouped GEMM (gemm128.hpp).  Its 4x4 MFMA tiles per wave keep 128 accumulator
// registers, which only fit beside the panel prefetch in the accumulator (AGPR) file: this unit is built without the
// VGPR-form option of kernels_f64.hip.
#include "gemm128.hpp"

namespace mpx {

template void launch_gemm128<double, double>(const BigTile*, const int (&)[4], const RedTile*, int, void*, const GemmTask*,
                                             const GemmPair*, Operands, double, int, hipStream_t, int);

}  // namespace mpx
