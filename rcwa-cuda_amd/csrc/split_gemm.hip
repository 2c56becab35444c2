// Split-precision complex GEMM (FP32 matrix cores).  Not built yet: every
// call reports hipErrorNotSupported so that no caller silently gets FP64.
#include "common.hpp"
#include "runtime.hpp"

int zgemm_split_batched(hipStream_t, int, int, int, zc, ZCMat, ZCMat, zc, ZCMat, ZMat, int) {
  return (int)hipErrorNotSupported;
}
