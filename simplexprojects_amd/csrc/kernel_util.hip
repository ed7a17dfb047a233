// kernel_util.hip -- small helper kernels.  Counters and flags are cleared with a fill
// kernel rather than hipMemsetAsync so that the same sequence can be captured into a
// hipGraph and replayed.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace blpg {

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

}  // namespace blpg
