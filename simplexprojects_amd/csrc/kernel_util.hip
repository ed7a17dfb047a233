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

// Expands a packed range of FP64 values: element i is vals[pref[w] + (set bits of mask[w]
// below bit i%64)] when bit i%64 of mask[w] (w = i/64) is set, else +0.0.  pref[w] counts
// the set bits of the words before w.
__global__ void k_expand_f64(const uint64_t* __restrict__ mask, const int32_t* __restrict__ pref,
                             const double* __restrict__ vals, int64_t n, double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t w = i >> 6;
    const int b = (int)(i & 63);
    const uint64_t m = mask[w];
    out[i] = ((m >> b) & 1ull) ? vals[pref[w] + __popcll(m & ((1ull << b) - 1ull))] : 0.0;
  }
}

hipError_t launch_expand_f64(const uint64_t* mask, const int32_t* pref, const double* vals,
                             int64_t n, double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_expand_f64, dim3((unsigned)blocks), dim3(256), 0, st, mask, pref, vals, n,
                     out);
  return hipGetLastError();
}

}  // namespace blpg
