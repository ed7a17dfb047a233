// noalloc.cpp -- blpg_solve_shared_A_noalloc: the shared-A batch solve for callers that
// cannot allocate (code under hipGraph capture, tight serving loops, harnesses that forbid
// hipMalloc).  No counterpart in the reference.
//
// Inputs and outputs are the caller's device buffers.  Everything else the solve needs --
// the LP descriptors, the work-queue counter, the staged A | b | C pool and the per-LP
// outputs the caller does not take -- lives in `scratch`, a file-scope device array of
// kScratchBytes that each GPU holds until the process ends.  Its device address is looked
// up once per device (hipGetSymbolAddress) and only that address is handed to kernels and
// copies.  Every call on a GPU shares that GPU's array: calls that overlap on one GPU race
// (serialise them on one stream).  The call enqueues on the caller's stream and returns;
// nothing in it allocates, synchronises or uses hipMemsetAsync, so it can be captured.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "../../include/blpg/blpg.h"
#include "blpg_internal.hpp"
#include "ctx.hpp"
#include "kernels.hpp"
#include "tuner.hpp"

namespace blpg {

constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each GPU

// The scratch array's address on the current device (looked up once per device).
hipError_t scratch_on(int dev, char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = (char*)p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Scratch layout of one call (byte offsets; every section 256-byte aligned).
struct Layout {
  size_t queue, desc, pool, spool, p1, unb, wfin, amax, pivots, total;
  int64_t a_len, b_off, c_off;  // pool elements: A at 0, b at b_off, C at c_off
};

Layout layout(int m, int n, int64_t count, bool own_pivots) {
  Layout L{};
  L.a_len = (int64_t)m * n;
  L.b_off = (L.a_len + 7) & ~7ll;
  L.c_off = L.b_off + ((m + 7) & ~7);
  size_t off = 0;
  L.queue = off; off += 256;
  L.desc = off; off += al256(sizeof(LpDesc) * (size_t)count);
  L.pool = off; off += al256(sizeof(double) * (size_t)(L.c_off + count * (int64_t)n));
  L.spool = off; off += al256((size_t)m);
  L.p1 = off; off += al256(sizeof(int32_t) * (size_t)count);
  L.unb = off; off += al256(sizeof(int32_t) * (size_t)count);
  L.wfin = off; off += al256(sizeof(double) * (size_t)count);
  L.amax = off; off += al256(sizeof(double) * (size_t)count);
  L.pivots = off;
  if (own_pivots) off += al256(sizeof(int32_t) * (size_t)count);
  L.total = off;
  return L;
}

// Descriptors of the shared-A batch: LP k reads A and b at the pool's front and row k of C;
// its x and basis go to rows k of the caller's count x n and count x m arrays.
__global__ void k_shared_desc(LpDesc* desc, int32_t m, int32_t n, int64_t count, int64_t b_off,
                              int64_t c_off, int has_sense) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x) {
    LpDesc d;
    d.m = m;
    d.n = n;
    d.a = 0;
    d.b = b_off;
    d.c = c_off + k * n;
    d.s = has_sense ? 0 : -1;
    d.x = k * n;
    d.basis = k * m;
    d.t = 0;
    desc[k] = d;
  }
}

}  // namespace

int64_t noalloc_max_count(int32_t m, int32_t n) {
  if (m < 1 || n < 1) return 0;
  const Layout one = layout(m, n, 1, true);
  if (one.total > kScratchBytes) return 0;
  // per LP: descriptor, one row of C, p1, unb, wfin, amax, pivots (plus alignment slack)
  const int64_t per = (int64_t)sizeof(LpDesc) + 8 * (int64_t)n + 4 + 4 + 8 + 8 + 4;
  int64_t k = (int64_t)(kScratchBytes - one.total) / per + 1;
  while (k > 0 && layout(m, n, k, true).total > kScratchBytes) --k;
  return k;
}

}  // namespace blpg

using namespace blpg;

extern "C" {

int blpg_solve_shared_A_noalloc(int32_t m, int32_t n, const double* A, const double* b,
                                const int8_t* sense, const double* C, int64_t count,
                                int32_t* status, double* objective, double* x, int32_t* basis,
                                int32_t* pivots, void* stream) {
  if (count < 0 || m < 1 || n < 1 ||
      (count > 0 && (!A || !b || !C || !status || !objective || !x || !basis)))
    return set_free_error(BLPG_E_INVALID, "blpg_solve_shared_A_noalloc: invalid arguments");
  if (count == 0) return BLPG_OK;
  const int v = Tuner::best_tile(m, n, BLPG_FP64);
  if (v < 0)
    return set_free_error(BLPG_E_INVALID, "blpg_solve_shared_A_noalloc: %dx%d exceeds the "
                          "register-tile kernels", (int)m, (int)n);
  const int64_t maxc = noalloc_max_count(m, n);
  if (count > maxc)
    return set_free_error(BLPG_E_BUDGET, "blpg_solve_shared_A_noalloc: %lld LPs of %dx%d need "
                          "more than the %zu bytes of scratch (at most %lld LPs)",
                          (long long)count, (int)m, (int)n, kScratchBytes, (long long)maxc);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return set_free_error(BLPG_E_NODEVICE, "blpg_solve_shared_A_noalloc: no current device");
  char* base = nullptr;
  hipError_t e = scratch_on(dev, &base);
  if (e != hipSuccess)
    return set_free_error(BLPG_E_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                          hipGetErrorString(e));
  hipStream_t st = (hipStream_t)stream;
  const Layout L = layout(m, n, count, pivots == nullptr);
  double* pool = (double*)(base + L.pool);
  int8_t* spool = (int8_t*)(base + L.spool);
  LpDesc* desc = (LpDesc*)(base + L.desc);
  int32_t* queue = (int32_t*)(base + L.queue);
  e = hipMemcpyAsync(pool, A, sizeof(double) * (size_t)L.a_len, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pool + L.b_off, b, sizeof(double) * (size_t)m, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pool + L.c_off, C, sizeof(double) * (size_t)(count * n),
                       hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess && sense)
    e = hipMemcpyAsync(spool, sense, (size_t)m, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess)
    return set_free_error(BLPG_E_HIP, "staging copy failed: %s", hipGetErrorString(e));
  const int64_t blocks = std::min<int64_t>((count + 255) / 256, 1024);
  hipLaunchKernelGGL(k_shared_desc, dim3((unsigned)blocks), dim3(256), 0, st, desc, m, n, count,
                     L.b_off, L.c_off, sense ? 1 : 0);
  e = hipGetLastError();
  if (e == hipSuccess) e = launch_fill_i32(queue, 0, 1, st);
  if (e != hipSuccess)
    return set_free_error(BLPG_E_HIP, "set-up kernels failed: %s", hipGetErrorString(e));
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return set_free_error(BLPG_E_HIP, "hipDeviceGetAttribute failed");
  const int per_cu = tile_max_blocks_per_cu(BLPG_FP64, v);
  if (per_cu <= 0) return set_free_error(BLPG_E_NODEVICE, "tile variant %d cannot be resident", v);
  TileLaunch T{};
  T.precision = BLPG_FP64;
  T.variant = v;
  T.desc = desc;
  T.pool = pool;
  T.spool = spool;
  T.idx = nullptr;
  T.count = (int32_t)count;
  T.queue = queue;
  T.out.status = status;
  T.out.pivots = pivots ? pivots : (int32_t*)(base + L.pivots);
  T.out.p1 = (int32_t*)(base + L.p1);
  T.out.unb = (int32_t*)(base + L.unb);
  T.out.obj = objective;
  T.out.wfin = (double*)(base + L.wfin);
  T.out.amax = (double*)(base + L.amax);
  T.out.x = x;
  T.out.basis = basis;
  T.out.tab = nullptr;
  T.eps_d = 1e-9;
  T.eps_p = 1e-9;
  T.tol = 1e-7;
  T.max_iter = 0;
  T.bland_after = 50;
  T.grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ncu * per_cu, count));
  T.warm = nullptr;
  e = launch_tile(T, st);
  if (e != hipSuccess) return set_free_error(BLPG_E_HIP, "solve launch failed: %s", hipGetErrorString(e));
  return BLPG_OK;
}

int64_t blpg_noalloc_max_count(int32_t m, int32_t n) { return noalloc_max_count(m, n); }

}  // extern "C"
