// kernel_stream.hip -- the size-generic simplex kernel.
//
// One workgroup per LP, persistent over a work queue (an atomic LP counter), for any
// (m, n).  The condensed tableau ((m+2) x (n+2), row-major) lives either in LDS, when it
// fits the 160 KB of a CU, or in an HBM workspace slab owned by the workgroup, streamed
// with coalesced row accesses on every pivot.  Per pivot:
//   pricing   one pass over the objective row -> workgroup argmin
//   ratio     one pass over the pivot column (kept in LDS as `colbuf`) -> argmin
//   pivot     pivot row scaled into LDS (`rowbuf`), then the rank-1 update
//             T[i][j] = fma(-colbuf[i], rowbuf[j], T[i][j]); rows with colbuf[i] == 0
//             are skipped (their entries are unchanged up to the sign of zero)
// This replaces the reference's mykernel (simplex.cu:35-385), which keeps the full
// tableau (slack identity included) in global memory, reduces through global scratch
// and updates with one thread per row looping over columns.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "simplex_device.hpp"

namespace blpg {

template <typename T, int NT>
struct StreamTab {
  static constexpr int NW = NT / 64;
  static constexpr bool kBlockCrash = false;  // crash one target per pivot
  T* tab;
  int ld, R, C, m, n;
  T* colbuf;
  T* rowbuf;
  int* bs;
  int* nb;
  uint8_t* dead;
  uint8_t* redund;
  uint8_t* target;  // target[var] = 1: var belongs to the warm-start basis
  double* s_am;     // per-wave max |A_ij| (NW entries)
  Reducer<T, NW> red;

  __device__ __forceinline__ T& at(int i, int j) { return tab[(int64_t)i * ld + j]; }
  __device__ __forceinline__ void loop_fence() {}

  __device__ __forceinline__ void init(const double* A, const double* b, const double* c, const int8_t* sense) {
    const int tid = threadIdx.x;
    for (int e = tid; e < R * C; e += NT) {
      const int i = e / C, j = e - (e / C) * C;
      at(i, j) = init_entry<T>(i, j, m, n, A, b, c, sense);
    }
    // max |A_ij| in FP64 (the check scales its tolerances by it)
    double am = 0.0;
    for (int e = tid; e < m * n; e += NT) am = fmax(am, fabs(A[e]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) am = fmax(am, __shfl_xor(am, off));
    if ((tid & 63) == 0) s_am[tid >> 6] = am;
    for (int i = tid; i < R; i += NT) { bs[i] = i >= 2 ? n + i - 2 : -1; redund[i] = 0; }
    for (int j = tid; j < C; j += NT) {
      nb[j] = j == 0 ? -1 : (j <= n ? j - 1 : n + m);
      dead[j] = (j == 0 || j == C - 1) ? 1 : 0;
    }
    __syncthreads();
  }

  // marks the warm-start targets; false when the list holds no valid variable id
  __device__ __forceinline__ bool set_target(const int32_t* warm) {
    for (int v = threadIdx.x; v < n + m; v += NT) target[v] = 0;
    __syncthreads();
    int any = 0;
    if (warm)
      for (int i = threadIdx.x; i < m; i += NT) {
        const int v = warm[i];
        if (v >= 0 && v < n + m) { target[v] = 1; any = 1; }
      }
    return __syncthreads_or(any) != 0;
  }

  // dual simplex leaving row: most negative rhs below -tol (Bland: smallest basic id)
  __device__ __forceinline__ int dual_row(bool bland, T tol) {
    __syncthreads();
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    for (int i = 2 + (int)threadIdx.x; i < R; i += NT) {
      if (redund[i]) continue;
      const T beta = at(i, 0);
      if (!(beta < -tol)) continue;
      const T key = bland ? T(bs[i]) : beta;
      if (better(key, i, bv, bk)) { bv = key; bk = i; bp = i; }
    }
    red.argmin(bv, bk, bp);
    return bp;
  }

  // dual simplex entering column for leaving row r
  __device__ __forceinline__ int dual_ratio(int r, bool bland, T eps_p, T& theta) {
    __syncthreads();
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    for (int j = 1 + (int)threadIdx.x; j < C; j += NT) {
      if (dead[j]) continue;
      const T a = at(r, j);
      if (!(a < -eps_p)) continue;
      const T d = at(0, j);
      const T th = (d > T(0) ? d : T(0)) / -a;
      const T key = th > T(0) ? th : a;
      const int kk = bland ? nb[j] : j;
      if (better(key, kk, bv, bk)) { bv = key; bk = kk; bp = j; }
    }
    red.argmin(bv, bk, bp);
    theta = bv > T(0) ? bv : T(0);
    return bp;
  }

  // row of largest |T[i][s]| among constraint rows whose basic variable is not a target
  __device__ __forceinline__ int absmax_row(int s, T eps_p) {
    __syncthreads();
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    for (int i = 2 + (int)threadIdx.x; i < R; i += NT) {
      const int v = bs[i];
      if (redund[i] || v < 0 || target[v]) continue;
      const T key = -tabs(at(i, s));
      if (better(key, i, bv, bk)) { bv = key; bk = i; bp = i; }
    }
    red.argmin(bv, bk, bp);
    if (bp < 0 || !(-bv > eps_p)) return -1;
    return bp;
  }

  __device__ __forceinline__ int bs_of(int i) {
    __syncthreads();
    return uniform(bs[i]);
  }
  __device__ __forceinline__ int bs_now(int i) { return bs_of(i); }
  __device__ __forceinline__ int warm_at(const int32_t* warm, int ci) const { return warm[ci]; }

  __device__ __forceinline__ int pricing(int orow, bool bland, T eps_d) {
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    for (int j = 1 + (int)threadIdx.x; j < C; j += NT) {
      if (dead[j]) continue;
      const T d = at(orow, j);
      if (!(d < -eps_d)) continue;
      const T v = bland ? T(nb[j]) : d;
      if (better(v, j, bv, bk)) { bv = v; bk = j; bp = j; }
    }
    red.argmin(bv, bk, bp);
    return bp;
  }

  __device__ __forceinline__ int ratio(int s, bool bland, T eps_p, T& theta) {
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    for (int i = threadIdx.x; i < R; i += NT) {
      const T a = at(i, s);
      colbuf[i] = a;
      if (i >= 2 && a > eps_p) {
        const T beta = at(i, 0);
        const T th = (beta > T(0) ? beta : T(0)) / a;
        const T v = th > T(0) ? th : -a;  // degenerate rows: largest pivot first
        const int k = bland ? bs[i] : i;
        if (better(v, k, bv, bk)) { bv = v; bk = k; bp = i; }
      }
    }
    red.argmin(bv, bk, bp);
    theta = bv > T(0) ? bv : T(0);
    return bp;
  }

  __device__ __forceinline__ void pivot(int r, int s) {
    const int tid = threadIdx.x;
    // every row's pivot-column entry (the ratio pass filled colbuf; refresh it here so
    // phase-I pivots chosen without a ratio pass see it too)
    __syncthreads();
    for (int i = tid; i < R; i += NT) colbuf[i] = at(i, s);
    __syncthreads();
    const T p = uniform(colbuf[r]);
    const T rp = T(1) / p;
    for (int j = tid; j < C; j += NT) rowbuf[j] = (j == s) ? rp : at(r, j) * rp;
    if (tid == 0) { const int t = bs[r]; bs[r] = nb[s]; nb[s] = t; }
    __syncthreads();
    const int lane = tid & 63, w = tid >> 6;
    for (int i = w; i < R; i += NW) {
      const T ci = -colbuf[i];
      if (i == r) {
        for (int j = lane; j < C; j += 64) at(r, j) = rowbuf[j];
      } else if (ci != T(0)) {
        T* row = &at(i, 0);
        for (int j = lane; j < C; j += 64) {
          const T old = (j == s) ? T(0) : row[j];
          row[j] = fma_t(ci, rowbuf[j], old);
        }
      }
    }
    __syncthreads();
  }

  __device__ __forceinline__ int absmax_col(int r, T eps_p) {
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    for (int j = 1 + (int)threadIdx.x; j < C; j += NT) {
      if (dead[j]) continue;
      const T v = -tabs(at(r, j));
      if (better(v, j, bv, bk)) { bv = v; bk = j; bp = j; }
    }
    red.argmin(bv, bk, bp);
    if (bp < 0 || !(-bv > eps_p)) return -1;
    return bp;
  }

  // A dead column never enters again and influences no other entry (an update of (i, j)
  // reads only column s and row r), so it keeps its values: B^-1 columns of fixed slacks
  // stay available to the FP64 check.
  __device__ __forceinline__ void kill_col(int j) {
    __syncthreads();
    if (threadIdx.x == 0) dead[j] = 1;
    __syncthreads();
  }

  __device__ __forceinline__ T get(int i, int j) {
    __syncthreads();
    return uniform(at(i, j));
  }

  __device__ __forceinline__ void set(int i, int j, T v) {
    __syncthreads();
    if (threadIdx.x == 0) at(i, j) = v;
    __syncthreads();
  }

  __device__ __forceinline__ void mark_redund(int r) {
    __syncthreads();
    if (threadIdx.x == 0) { redund[r] = 1; at(r, 0) = T(0); }
    __syncthreads();
  }

  __device__ __forceinline__ int min_rhs_row() {
    __syncthreads();
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    for (int i = 2 + (int)threadIdx.x; i < R; i += NT) {
      if (redund[i]) continue;
      const T v = at(i, 0);
      if (better(v, i, bv, bk)) { bv = v; bk = i; bp = i; }
    }
    red.argmin(bv, bk, bp);
    return bp;
  }

  __device__ __forceinline__ void setup_x0() {
    __syncthreads();
    const int x0 = C - 1;
    if (threadIdx.x == 0) dead[x0] = 0;
    for (int i = threadIdx.x; i < R; i += NT)
      at(i, x0) = (i >= 2) ? (redund[i] ? T(0) : T(-1)) : (i == 1 ? T(1) : T(0));
    for (int j = threadIdx.x; j < C - 1; j += NT) at(1, j) = T(0);
    __syncthreads();
  }

  __device__ __forceinline__ int find_basic(int var) {
    __syncthreads();
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    for (int i = 2 + (int)threadIdx.x; i < R; i += NT)
      if (bs[i] == var) { bv = T(0); bk = i; bp = i; }
    red.argmin(bv, bk, bp);
    return bp;
  }

  __device__ __forceinline__ int find_nonbasic(int var) {
    __syncthreads();
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    for (int j = 1 + (int)threadIdx.x; j < C; j += NT)
      if (nb[j] == var) { bv = T(0); bk = j; bp = j; }
    red.argmin(bv, bk, bp);
    return bp;
  }

  __device__ __forceinline__ int nb_of(int j) {
    __syncthreads();
    return uniform(nb[j]);
  }

  __device__ __forceinline__ void write_outputs(int lp, const LpDesc& d, const Out& out, int status, int it,
                                int p1, int unb, T wfin, int last) {
    __syncthreads();
    const int tid = threadIdx.x;
    if (tid == 0) {
      double am = s_am[0];
      for (int w = 1; w < NW; ++w) am = fmax(am, s_am[w]);
      out.amax[lp] = am;
      out.status[lp] = status;
      out.pivots[lp] = it;
      out.p1[lp] = p1;
      out.unb[lp] = unb;
      if (out.lastpiv) out.lastpiv[lp] = last;
      out.obj[lp] = (double)at(0, 0);
      out.wfin[lp] = (double)wfin;
    }
    double* x = out.x + d.x;
    int32_t* basis = out.basis + d.basis;
    for (int j = tid; j < n; j += NT) x[j] = 0.0;
    for (int i = tid; i < m; i += NT) basis[i] = bs[2 + i];
    __syncthreads();
    for (int i = 2 + tid; i < R; i += NT) {
      const int v = bs[i];
      if (v >= 0 && v < n) x[v] = (double)at(i, 0);
    }
    // final tableau and nonbasic ids for the FP64 check (B^-1 is in the slack columns)
    if (out.tab) {
      T* tp = (T*)(out.tab + 8 * d.t);
      for (int e = tid; e < R * C; e += NT) tp[e] = tab[e];
      int32_t* np = (int32_t*)(out.tab + 8 * (d.t + (int64_t)R * C));
      for (int j = tid; j < C; j += NT) np[j] = nb[j];
    }
  }
};

// LDS carve-up shared by host and device:
// [tab?][colbuf][rowbuf][bs][nb][dead][redund][target][slots]
template <typename T>
struct StreamLayout {
  size_t tab, colbuf, rowbuf, bs, nb, dead, redund, target, slots, total;
  __host__ __device__ StreamLayout(bool lds_tab, int nw, int rcap, int ccap) {
    auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
    size_t o = 0;
    tab = o; if (lds_tab) o += al(sizeof(T) * (size_t)rcap * ccap);
    colbuf = o; o += al(sizeof(T) * rcap);
    rowbuf = o; o += al(sizeof(T) * ccap);
    bs = o; o += al(sizeof(int) * rcap);
    nb = o; o += al(sizeof(int) * ccap);
    dead = o; o += al(ccap);
    redund = o; o += al(rcap);
    target = o; o += al(rcap + ccap);
    slots = o; o += al(sizeof(RedSlot<T>) * 2 * nw);
    total = o;
  }
};

template <typename T, int NT, bool LDS_TAB>
__global__ __launch_bounds__(NT) void k_stream(const LpDesc* __restrict__ desc,
                                               const double* __restrict__ pool,
                                               const int8_t* __restrict__ spool,
                                               const int32_t* __restrict__ idx, int32_t count,
                                               const int32_t* count_dev, int32_t* queue,
                                               Out out, Opt<T> o, T* ws,
                                               int64_t ws_stride, int rcap, int ccap,
                                               const int32_t* __restrict__ warm_pool) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_q;
  __shared__ double s_am[NT / 64];
  const StreamLayout<T> L(LDS_TAB, NT / 64, rcap, ccap);
  StreamTab<T, NT> t;
  t.s_am = s_am;
  t.colbuf = (T*)(smem + L.colbuf);
  t.rowbuf = (T*)(smem + L.rowbuf);
  t.bs = (int*)(smem + L.bs);
  t.nb = (int*)(smem + L.nb);
  t.dead = smem + L.dead;
  t.redund = smem + L.redund;
  t.target = smem + L.target;
  t.red.init((RedSlot<T>*)(smem + L.slots));
  t.tab = LDS_TAB ? (T*)(smem + L.tab) : ws + (int64_t)blockIdx.x * ws_stride;
  if (count_dev) count = *count_dev;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
    __syncthreads();
    const int q = uniform(s_q);
    if (q >= count) break;
    const int lp = idx ? idx[q] : q;
    const LpDesc d = desc[lp];
    t.m = d.m; t.n = d.n; t.R = d.m + 2; t.C = d.n + 2; t.ld = t.C;
    const int8_t* sense = d.s >= 0 ? spool + d.s : nullptr;
    t.init(pool + d.a, pool + d.b, pool + d.c, sense);
    const int32_t* warm = warm_pool ? warm_pool + d.basis : nullptr;
    if (!t.set_target(warm)) warm = nullptr;
    int status, it, p1, unb, last;
    T wfin;
    run_simplex<T>(t, d.m, d.n, sense, o, status, it, p1, unb, wfin, last, warm);
    t.write_outputs(lp, d, out, status, it, p1, unb, wfin, last);
  }
}

// ---------------------------------------------------------------------------------
// host side

size_t stream_lds_bytes(int precision, bool lds_tab, int nt, int rcap, int ccap) {
  if (precision == 0) return StreamLayout<double>(lds_tab, nt / 64, rcap, ccap).total;
  return StreamLayout<float>(lds_tab, nt / 64, rcap, ccap).total;
}

template <typename T, int NT, bool LDS>
static const void* stream_fn() {
  return reinterpret_cast<const void*>(&k_stream<T, NT, LDS>);
}

const void* stream_kernel_ptr(int precision, bool lds_tab, int nt) {
#define BLPG_PICK(T)                                              \
  switch (nt) {                                                   \
    case 64: return lds_tab ? stream_fn<T, 64, true>() : stream_fn<T, 64, false>();    \
    case 128: return lds_tab ? stream_fn<T, 128, true>() : stream_fn<T, 128, false>(); \
    case 256: return lds_tab ? stream_fn<T, 256, true>() : stream_fn<T, 256, false>(); \
    case 512: return lds_tab ? stream_fn<T, 512, true>() : stream_fn<T, 512, false>(); \
    default: return nullptr;                                      \
  }
  if (precision == 0) { BLPG_PICK(double) }
  BLPG_PICK(float)
#undef BLPG_PICK
}

hipError_t launch_stream(const StreamLaunch& L, hipStream_t st) {
  const void* fn = stream_kernel_ptr(L.precision, L.lds_tab, L.nt);
  if (!fn) return hipErrorInvalidValue;
  const size_t lds = stream_lds_bytes(L.precision, L.lds_tab, L.nt, L.rcap, L.ccap);
  if (L.precision == 0) {
    Opt<double> o{L.eps_d, L.eps_p, L.tol, L.max_iter, L.bland_after};
    void* args[] = {(void*)&L.desc, (void*)&L.pool, (void*)&L.spool, (void*)&L.idx,
                    (void*)&L.count, (void*)&L.count_dev, (void*)&L.queue, (void*)&L.out, (void*)&o,
                    (void*)&L.ws, (void*)&L.ws_stride, (void*)&L.rcap, (void*)&L.ccap, (void*)&L.warm};
    return hipLaunchKernel(fn, dim3(L.grid), dim3(L.nt), args, lds, st);
  }
  Opt<float> o{(float)L.eps_d, (float)L.eps_p, (float)L.tol, L.max_iter, L.bland_after};
  void* args[] = {(void*)&L.desc, (void*)&L.pool, (void*)&L.spool, (void*)&L.idx,
                  (void*)&L.count, (void*)&L.count_dev, (void*)&L.queue, (void*)&L.out, (void*)&o,
                  (void*)&L.ws, (void*)&L.ws_stride, (void*)&L.rcap, (void*)&L.ccap, (void*)&L.warm};
  return hipLaunchKernel(fn, dim3(L.grid), dim3(L.nt), args, lds, st);
}

int stream_max_blocks_per_cu(int precision, bool lds_tab, int nt, int rcap, int ccap) {
  const void* fn = stream_kernel_ptr(precision, lds_tab, nt);
  if (!fn) return 0;
  int nb = 0;
  const size_t lds = stream_lds_bytes(precision, lds_tab, nt, rcap, ccap);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, nt, lds) != hipSuccess) return 0;
  return nb;
}

}  // namespace blpg
