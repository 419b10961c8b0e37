// Instantiation of the 128x128 split-BF16 (x3) grouped GEMM (gemm_x3_128.hpp), FP32 and FP64 C.
#include "gemm_x3_128.hpp"

namespace mpx {

template void launch_gemm_x3_128<float>(const BigTile*, const int (&)[4], const RedTile*, int, void*, const GemmTask*,
                                        const GemmPair*, Operands, double, int, hipStream_t, int);
template void launch_gemm_x3_128<double>(const BigTile*, const int (&)[4], const RedTile*, int, void*, const GemmTask*,
                                         const GemmPair*, Operands, double, int, hipStream_t, int);

}  // namespace mpx
