// kernel_wave.hip -- column-per-lane register simplex kernel for gfx950.
//
// One workgroup of NW wavefronts per LP, persistent over a work queue.  Thread t owns
// whole tableau columns: column 1 + t + NT*q (q < CQ, NT = 64*NW), every row, in VGPRs
// (`v[q][i]`, every index static so the array lives in registers).  The
// right-hand side (column 0) is kept instead as a row-distributed vector in every
// wavefront (`rhs[p]` holds row lane + 64p), which is what the ratio test consumes.
//
// Per pivot:
//   pricing   each lane's CQ row-0 candidates -> DPP wavefront argmin (-> LDS combine
//             across wavefronts when NW > 1: one barrier)
//   column s  its owner writes the R entries to LDS (the only published vector;
//             one barrier when NW > 1)
//   ratio     every wavefront reads the column row-distributed and runs the same DPP
//             argmin over rows, so all of them hold the pivot row without a broadcast
//   update    each lane takes its own pivot-row entry out of its registers (a uniform
//             branch picks the 16-row group, a select chain the row), then R*CQ FMAs
//             with the column read back from LDS as wave-uniform broadcasts
// With NW = 1 no barrier is left at all; the wavefront's LDS accesses are in program
// order.  Arithmetic is that of the oracle and of the other kernels, operation for
// operation, so results are bit-identical across kernels.  Replaces the reference's
// mykernel (simplex.cu:35-385), whose tableau lives in global memory and whose pivot
// takes ~4 log2(blockDim) barriers.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "simplex_device.hpp"

namespace blpg {

// 16-byte LDS vector of the tableau type (one ds_read_b128)
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  static constexpr int N = 4;
  typedef float V __attribute__((ext_vector_type(4)));
};
template <>
struct Vec16<double> {
  static constexpr int N = 2;
  typedef double V __attribute__((ext_vector_type(2)));
};

// Minimum over the 64 lanes of a wavefront (all lanes active): four in-row DPP steps,
// then the row-15 / row-31 broadcasts fold the rows into lane 63, read back uniform.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_mv(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_mv(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp((int)(b & 0xffffffffll), (int)(b & 0xffffffffll), CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, ROWS, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// A register move the optimiser cannot see through: a switch whose cases are such moves
// is not folded back into a dynamically indexed array (which would live in scratch).
__device__ __forceinline__ float rmov(float x) {
  float y;
  asm("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
  return y;
}
__device__ __forceinline__ double rmov(double x) {
  double y;
  asm("v_mov_b64 %0, %1" : "=v"(y) : "v"(x));
  return y;
}
__device__ __forceinline__ float tmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double tmin(double a, double b) { return __builtin_fmin(a, b); }

template <typename T>
__device__ __forceinline__ T wave_fmin(T x);

__device__ __forceinline__ float wlane_read(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double wlane_read(double x, int l) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <typename T>
__device__ __forceinline__ T wave_fmin(T x) {
  if constexpr (sizeof(T) == 4) {
    // FP32: each step is one v_min_f32 with a DPP source (the same IEEE minimum as fminf
    // on the finite / infinite keys that reach it, without a separate move and the NaN
    // canonicalisation fminf adds); rows outside a step's row mask keep their value.
    // s_nop 1: the two wait states a DPP read of a VGPR needs after the VALU write.
    asm volatile(
        "s_nop 1\n\t"
        "v_min_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_min_f32_dpp %0, %0, %0 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_min_f32_dpp %0, %0, %0 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_min_f32_dpp %0, %0, %0 row_mirror row_mask:0xf bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_min_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "s_nop 1\n\t"
        "v_min_f32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        "s_nop 1"
        : "+v"(x));
    return wlane_read(x, 63);
  }
  x = tmin(x, dpp_mv<kDppXor1, 0xF>(x));
  x = tmin(x, dpp_mv<kDppXor2, 0xF>(x));
  x = tmin(x, dpp_mv<kDppHalfMirror, 0xF>(x));
  x = tmin(x, dpp_mv<kDppMirror, 0xF>(x));
  x = tmin(x, dpp_mv<kDppBcast15, 0xA>(x));
  x = tmin(x, dpp_mv<kDppBcast31, 0xC>(x));
  return wlane_read(x, 63);
}

template <typename T, int NW, int CQ, int RCAP>
struct WaveTab {
  using KO = KeyOps<T>;
  using K = typename KO::K;
  using V4 = typename Vec16<T>::V;
  static constexpr int N4 = Vec16<T>::N;
  static constexpr int CS = 16;                  // rows per group of the pivot-row pick
  static constexpr int NCH = (RCAP + CS - 1) / CS;
  static_assert(RCAP % 16 == 0, "row capacity is a multiple of 16");
  static constexpr int NT = NW * 64;
  static constexpr int kRcap = RCAP;
  static constexpr int RP = (RCAP + 63) / 64;    // row slots per lane (row-distributed)
  static constexpr int CCAP = 1 + NT * CQ;       // columns incl. the rhs
  static_assert(CQ <= 32, "dead-column mask is 32 bits");
  static constexpr bool kBlockCrash = false;
  static constexpr int KB = 1;

  T v[CQ][RCAP];  // v[q][i] = T[i][1 + t + NT*q]
  T rhs[RP];      // T[lane + 64p][0] (every wavefront holds all rows)
  T acol[RP];     // the published column, row-distributed (set by publish_col)
  int bsr[RP];    // basic variable of row lane + 64p (every wavefront's own copy)
  int nbv[CQ];    // nonbasic variable of each owned column
  uint32_t dead;  // bit q: column 1 + t + NT*q cannot enter
  int tid, lane, wv;
  int R, C, m, n;
  int col_ready;  // column whose entries are in colbuf (-1: none)
  int phase;
  T* colbuf;      // RCAP: published column
  T* scal;
  int* bs;        // RCAP: basic variable of each row (LDS mirror of bsr, for rare paths)
  int* nb;        // CCAP: nonbasic variable of each column (LDS mirror of nbv)
  int* colvar;    // the published column's nonbasic variable
  uint8_t* redund;
  uint8_t* target;
  int* wl;
  int* colof;
  double* s_am;
  K* kslots;      // 2*NW keys for cross-wavefront combines
  int blk;        // (block crash is not used by this kernel)

  // Called once per state-machine step: the per-thread index and the LP's sizes become
  // opaque values, so the compiler does not hoist every row's predicates (i < R, ...) out
  // of the loop and hold them in registers for the whole kernel.
  __device__ __forceinline__ void loop_fence() {
    asm volatile("" : "+v"(tid), "+v"(lane), "+s"(R), "+s"(C), "+s"(m), "+s"(n));
  }

  __device__ __forceinline__ int col_of(int q) const { return 1 + tid + NT * q; }

  // T[r][col_of(q)] for wave-uniform r, and its overwrite: a scalar branch tree on r
  // (switches on the group of CS rows, then the row) ends in one register move, so the
  // register indices stay static
  template <int C0>
  __device__ __forceinline__ T vget_g(int q, int k) const {
    switch (k) {
#define BLPG_VG(u) case u: return rmov(v[q][C0 * CS + u]);
      BLPG_VG(0) BLPG_VG(1) BLPG_VG(2) BLPG_VG(3) BLPG_VG(4) BLPG_VG(5) BLPG_VG(6) BLPG_VG(7)
      BLPG_VG(8) BLPG_VG(9) BLPG_VG(10) BLPG_VG(11) BLPG_VG(12) BLPG_VG(13) BLPG_VG(14)
      default: return rmov(v[q][C0 * CS + 15]);
#undef BLPG_VG
    }
  }
  template <int C0>
  __device__ __forceinline__ void vset_g(int q, int k, T x) {
    switch (k) {
#define BLPG_VS(u) case u: v[q][C0 * CS + u] = rmov(x); break;
      BLPG_VS(0) BLPG_VS(1) BLPG_VS(2) BLPG_VS(3) BLPG_VS(4) BLPG_VS(5) BLPG_VS(6) BLPG_VS(7)
      BLPG_VS(8) BLPG_VS(9) BLPG_VS(10) BLPG_VS(11) BLPG_VS(12) BLPG_VS(13) BLPG_VS(14)
      default: v[q][C0 * CS + 15] = rmov(x); break;
#undef BLPG_VS
    }
  }
  template <int C0>
  __device__ __forceinline__ T vget_c(int q, int ch, int k) const {
    if constexpr (C0 + 1 >= NCH) return vget_g<C0>(q, k);
    else return ch == C0 ? vget_g<C0>(q, k) : vget_c<(C0 + 1 < NCH ? C0 + 1 : C0)>(q, ch, k);
  }
  template <int C0>
  __device__ __forceinline__ void vset_c(int q, int ch, int k, T x) {
    if constexpr (C0 + 1 >= NCH) vset_g<C0>(q, k, x);
    else if (ch == C0) vset_g<C0>(q, k, x);
    else vset_c<(C0 + 1 < NCH ? C0 + 1 : C0)>(q, ch, k, x);
  }
  __device__ __forceinline__ T vget(int q, int r) const {
    const int ch = r / CS;
    return vget_c<0>(q, ch, r - ch * CS);
  }
  __device__ __forceinline__ void vset(int q, int r, T x) {
    const int ch = r / CS;
    vset_c<0>(q, ch, r - ch * CS, x);
  }
  // basic variable of row i (wave-uniform i)
  __device__ __forceinline__ int bsr_get(int i) const {
    const int ps = i >> 6;
    int x = bsr[0];
#pragma unroll
    for (int p = 1; p < RP; ++p) x = (p == ps) ? bsr[p] : x;
    return __builtin_amdgcn_readlane(x, i & 63);
  }
  // row i of the row-distributed rhs (wave-uniform i)
  __device__ __forceinline__ T rhs_get(int i) const {
    const int ps = i >> 6;
    T x = rhs[0];
#pragma unroll
    for (int p = 1; p < RP; ++p) x = (p == ps) ? rhs[p] : x;
    return wlane_read(x, i & 63);
  }
  __device__ __forceinline__ void rhs_set(int i, T x) {
#pragma unroll
    for (int p = 0; p < RP; ++p) rhs[p] = (lane + 64 * p == i) ? x : rhs[p];
  }

  // combine the wavefronts' minimum keys (column reductions); uniform result
  __device__ __forceinline__ K combine(K a) {
    if constexpr (NW > 1) {
      K* s = kslots + phase * NW;
      if (lane == 0) s[wv] = a;
      __syncthreads();
      a = s[0];
#pragma unroll
      for (int w = 1; w < NW; ++w) {
        const K b = s[w];
        a = KO::sel(KO::lt(b, a), b, a);
      }
      phase ^= 1;
    }
    return a;
  }
  __device__ __forceinline__ void decode(K a, T& bv, int& bk, int& bp) const {
    if (KO::kp(a) == ~0u) { bv = inf_of<T>(); bk = INT_MAX; bp = -1; }
    else key_decode<T>(a, bv, bk, bp);
    bv = uniform(bv);
    bk = uniform(bk);
    bp = uniform(bp);
  }
  // minimum over the columns of the workgroup
  __device__ __forceinline__ void col_min(const ArgMin<T>& am, T& bv, int& bk, int& bp) {
    decode(combine(wave_min_key<T>(am.key)), bv, bk, bp);
  }
  // minimum over the rows (every wavefront covers all rows: no combine)
  __device__ __forceinline__ void row_min(const ArgMin<T>& am, T& bv, int& bk, int& bp) const {
    decode(wave_min_key<T>(am.key), bv, bk, bp);
  }

  __device__ __forceinline__ void init(const double* A, const double* b, const double* c, const int8_t* sense) {
    double am = 0.0;
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const int j = col_of(q);
      const bool strc = j >= 1 && j <= n;
      const double* Aj = A + (strc ? j - 1 : 0);
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        // one group of rows at a time: its loads go out together, the next group's after
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < CS; ++k) {
          const int i = ch * CS + k;
          const bool crow = i >= 2 && i < R;
          const int r = crow ? i - 2 : 0;
          // row sign: wave-uniform, one scalar load per row
          const bool neg = crow && sense != nullptr && uniform((int)sense[r]) == 2;
          const bool isA = crow && strc, isC = i == 0 && strc;
          const double* src = isA ? Aj + (int64_t)r * n : isC ? c + (j - 1) : b;
          const double x = *src;
          const bool ng = isC || (isA && neg);
          v[q][i] = (isA || isC) ? T(ng ? -x : x) : T(0);
          if (isA) am = fmax(am, fabs(x));
        }
      }
    }
#pragma unroll
    for (int p = 0; p < RP; ++p) {
      const int i = lane + 64 * p;
      const bool crow = i >= 2 && i < R;
      const int r = crow ? i - 2 : 0;
      const double x = b[r];
      const bool ng = crow && sense && sense[r] == 2;
      rhs[p] = crow ? T(ng ? -x : x) : T(0);
      bsr[p] = crow ? n + i - 2 : -1;
    }
    dead = 0;
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const int j = col_of(q);
      nbv[q] = j >= C ? -1 : (j <= n ? j - 1 : n + m);
    }
#pragma unroll
    for (int q = 0; q < CQ; ++q)
      if (col_of(q) >= C - 1) dead |= 1u << q;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) am = fmax(am, __shfl_xor(am, off));
    if (lane == 0) s_am[wv] = am;
    for (int i = tid; i < RCAP; i += NT) { bs[i] = (i >= 2 && i < R) ? n + i - 2 : -1; redund[i] = 0; }
    for (int j = tid; j < CCAP; j += NT) nb[j] = (j == 0 || j >= C) ? -1 : (j <= n ? j - 1 : n + m);
    for (int x = tid; x < RCAP + CCAP; x += NT)
      colof[x] = x < n ? x + 1 : (x == n + m ? C - 1 : -1);
    col_ready = -1;
    __syncthreads();
  }

  __device__ __forceinline__ bool set_target(const int32_t* warm) {
    for (int x = tid; x < RCAP + CCAP; x += NT) target[x] = 0;
    __syncthreads();
    int any = 0;
    if (warm)
      for (int i = tid; i < m; i += NT) {
        const int x = warm[i];
        wl[i] = x;
        if (x >= 0 && x < n + m) { target[x] = 1; any = 1; }
      }
    return __syncthreads_or(any) != 0;
  }

  __device__ __forceinline__ int warm_at(const int32_t*, int ci) const { return uniform(wl[ci]); }
  __device__ __forceinline__ int bs_now(int i) const { return uniform(bs[i]); }
  __device__ __forceinline__ int bs_of(int i) {
    __syncthreads();
    return uniform(bs[i]);
  }
  __device__ __forceinline__ int nb_of(int j) {
    __syncthreads();
    return uniform(nb[j]);
  }
  // (the last pivot's map update is behind a barrier only for the wavefront that made it)
  __device__ __forceinline__ int find_nonbasic(int var) {
    if constexpr (NW > 1) __syncthreads();
    return uniform(colof[var]);
  }

  // the owner of column s writes it to colbuf; then every wavefront reads it
  // row-distributed into acol.  With `sync` a barrier first keeps the write behind every
  // wavefront's reads of the previous column.
  __device__ __forceinline__ void publish_col(int s, bool sync) {
    if (sync) __syncthreads();
    const int qs = (s - 1) / NT, ts = (s - 1) - qs * NT;
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      if (q == qs && tid == ts) {
#pragma unroll
        for (int i = 0; i < RCAP; i += N4) {
          V4 x;
#pragma unroll
          for (int u = 0; u < N4; ++u) x[u] = v[q][i + u];
          *(V4*)(colbuf + i) = x;
        }
        *colvar = nbv[q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < RP; ++p) {
      const int i = lane + 64 * p;
      acol[p] = i < RCAP ? colbuf[i] : T(0);
    }
    col_ready = s;
  }

  // row of largest |T[i][s]| among constraint rows whose basic variable is not a target
  __device__ __forceinline__ int absmax_row(int s, T eps_p) {
    publish_col(s, true);
    ArgMin<T> am;
#pragma unroll
    for (int p = 0; p < RP; ++p) {
      const int i = lane + 64 * p;
      const bool in = i >= 2 && i < R;
      const int x = in ? bsr[p] : -1;
      const bool ok = in && !redund[i] && x >= 0 && !target[x < 0 ? 0 : x];
      am.add(ok, -tabs(acol[p]), i, i);
    }
    T bv;
    int bk, bp;
    row_min(am, bv, bk, bp);
    if (bp < 0 || !(-bv > eps_p)) return -1;
    return bp;
  }

  // Dantzig pricing on objective row orow (0 or 1); Bland: smallest nonbasic id.
  // Dantzig ties go to the lowest column: the float minimum first, then the lowest lane
  // (slot) holding it -- the lexicographic (value, column) minimum of the key reduction.
  __device__ __forceinline__ int pricing(int orow, bool bland, T eps_d) {
    if (!bland) {
      T lv = inf_of<T>();
      int lq = -1;
#pragma unroll
      for (int q = 0; q < CQ; ++q) {
        const T d = orow == 0 ? v[q][0] : v[q][1];
        const bool ok = !((dead >> q) & 1u) && d < -eps_d;
        const bool take = ok && (lq < 0 || d < lv);
        lv = take ? d : lv;
        lq = take ? q : lq;
      }
      const T wmin = wave_fmin<T>(lv);
      int best = INT_MAX;
#pragma unroll
      for (int q = 0; q < CQ; ++q) {  // lowest column: slot first, then lane
        const unsigned long long b = __ballot(lq == q && lv == wmin);
        if (best == INT_MAX && b != 0ull) best = 1 + (__ffsll((long long)b) - 1) + 64 * wv + NT * q;
      }
      best = uniform(best);
      if constexpr (NW > 1) {
        K key = best == INT_MAX ? KO::none() : KO::make(wmin, best, best);
        decode(combine(key), lv, best, lq);
        return lq;
      }
      return best == INT_MAX ? -1 : best;
    }
    ArgMin<T> am;
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const int j = col_of(q);
      const T d = orow == 0 ? v[q][0] : v[q][1];
      const bool ok = !((dead >> q) & 1u) && d < -eps_d;
      am.add(ok, bland ? T(nbv[q]) : d, j, j);
    }
    T bv;
    int bk, bp;
    col_min(am, bv, bk, bp);
    return bp;
  }

  __device__ __forceinline__ int ratio(int s, bool bland, T eps_p, T& theta) {
    // pricing's combine (NW > 1) already ordered every wavefront's last reads of colbuf
    publish_col(s, false);
    if (!bland) {
      // each lane's rows lane + 64p: its smallest key, lowest row on ties
      T lv = inf_of<T>();
      int lp = -1;
#pragma unroll
      for (int p = 0; p < RP; ++p) {
        const int i = lane + 64 * p;
        const T a = acol[p];
        const T beta = rhs[p];
        const T th = (beta > T(0) ? beta : T(0)) / (a > eps_p ? a : T(1));
        const T key = th > T(0) ? th : -a;
        const bool ok = i >= 2 && i < R && a > eps_p;
        const bool take = ok && (lp < 0 || key < lv);
        lv = take ? key : lv;
        lp = take ? p : lp;
      }
      const T wmin = wave_fmin<T>(lv);
      int best = -1;
#pragma unroll
      for (int p = 0; p < RP; ++p) {  // lowest row: slot first, then lane
        const unsigned long long b = __ballot(lp == p && lv == wmin);
        if (best < 0 && b != 0ull) best = (__ffsll((long long)b) - 1) + 64 * p;
      }
      best = uniform(best);
      theta = (best >= 0 && wmin > T(0)) ? wmin : T(0);
      return best;
    }
    ArgMin<T> am;
#pragma unroll
    for (int p = 0; p < RP; ++p) {
      const int i = lane + 64 * p;
      const T a = acol[p];
      const T beta = rhs[p];
      const T th = (beta > T(0) ? beta : T(0)) / (a > eps_p ? a : T(1));
      const T key = th > T(0) ? th : -a;
      const bool ok = i >= 2 && i < R && a > eps_p;
      am.add(ok, key, bland ? bsr[p] : i, i);
    }
    T bv;
    int bk, bp;
    row_min(am, bv, bk, bp);
    theta = bv > T(0) ? bv : T(0);
    return bp;
  }

  __device__ __forceinline__ void pivot(int r, int s) {
    if (col_ready != s) publish_col(s, true);
    const int qs = (s - 1) / NT, ts = (s - 1) - qs * NT;
    const T rp = T(1) / uniform(colbuf[r]);
    T prow[CQ];
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const T x = vget(q, r);
      prow[q] = (q == qs && tid == ts) ? rp : x * rp;
    }
    const T prow0 = rhs_get(r) * rp;
    const int e = uniform(*colvar), lv = bsr_get(r);  // entering, leaving variable
    if (tid == 0) {
      bs[r] = e;
      nb[s] = lv;
      colof[e] = -1;
      colof[lv] = s;
    }
#pragma unroll
    for (int q = 0; q < CQ; ++q)
      if (q == qs && tid == ts) nbv[q] = lv;
#pragma unroll
    for (int p = 0; p < RP; ++p) bsr[p] = (lane + 64 * p == r) ? e : bsr[p];
    // the entering column's entries become fma(-T[i][s], rp, 0)
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      if (q == qs && wv == ts / 64) {
        if (tid == ts) {
#pragma unroll
          for (int i = 0; i < RCAP; ++i) v[q][i] = T(0);
        }
      }
    }
    // fma(-T[i][s], prow, x) computed as fma(T[i][s], -prow, x): the same exactly rounded
    // value (the product's sign is symmetric), and the accumulating form updates each
    // register in place
    T nprow[CQ];
#pragma unroll
    for (int q = 0; q < CQ; ++q) nprow[q] = -prow[q];
    // The column is read back from LDS in groups of G rows, one group ahead of the FMAs
    // that consume it; the scheduling barriers keep the compiler from hoisting every
    // group's reads to the top (one temporary register per row).
    constexpr int G = 32, NG = (RCAP + G - 1) / G;
    V4 cur[G / N4], nxt[G / N4];
#pragma unroll
    for (int u = 0; u < G / N4; ++u) cur[u] = *(const V4*)(colbuf + (u * N4 < RCAP ? u * N4 : 0));
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      if (g + 1 < NG) {
#pragma unroll
        for (int u = 0; u < G / N4; ++u) {
          const int i = (g + 1) * G + u * N4;
          nxt[u] = *(const V4*)(colbuf + (i < RCAP ? i : 0));
        }
      }
#pragma unroll
      for (int u = 0; u < G / N4; ++u) {
#pragma unroll
        for (int e = 0; e < N4; ++e) {
          const int i = g * G + u * N4 + e;
          if (i < RCAP) {
#pragma unroll
            for (int q = 0; q < CQ; ++q) v[q][i] = fma_t(cur[u][e], nprow[q], v[q][i]);
          }
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < G / N4; ++u) cur[u] = nxt[u];
    }
#pragma unroll
    for (int q = 0; q < CQ; ++q) vset(q, r, prow[q]);
#pragma unroll
    for (int p = 0; p < RP; ++p) {
      const int i = lane + 64 * p;
      rhs[p] = (i == r) ? prow0 : fma_t(acol[p], -prow0, rhs[p]);
    }
    col_ready = -1;
  }

  // column of largest |T[r][j]| among live columns; -1 if all <= eps_p
  __device__ __forceinline__ int absmax_col(int r, T eps_p) {
    __syncthreads();
    ArgMin<T> am;
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const int j = col_of(q);
      am.add(!((dead >> q) & 1u), -tabs(vget(q, r)), j, j);
    }
    T bv;
    int bk, bp;
    col_min(am, bv, bk, bp);
    if (bp < 0 || !(-bv > eps_p)) return -1;
    return bp;
  }

  __device__ __forceinline__ int dual_row(bool bland, T tol) {
    __syncthreads();
    ArgMin<T> am;
#pragma unroll
    for (int p = 0; p < RP; ++p) {
      const int i = lane + 64 * p;
      const bool in = i >= 2 && i < R;
      const T beta = rhs[p];
      am.add(in && !redund[i] && beta < -tol, bland ? T(bsr[p]) : beta, i, i);
    }
    T bv;
    int bk, bp;
    row_min(am, bv, bk, bp);
    return bp;
  }

  // dual simplex entering column for leaving row r: smallest max(T[0][j],0)/-T[r][j] over
  // T[r][j] < -eps_p (dual-degenerate columns by the largest |T[r][j]|)
  __device__ __forceinline__ int dual_ratio(int r, bool bland, T eps_p, T& theta) {
    col_ready = -1;
    ArgMin<T> am;
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const int j = col_of(q);
      const T a = vget(q, r);
      const T d = v[q][0];
      const bool ok = !((dead >> q) & 1u) && a < -eps_p;
      const T th = (d > T(0) ? d : T(0)) / (ok ? -a : T(1));
      const T key = th > T(0) ? th : a;
      am.add(ok, key, bland ? nbv[q] : j, j);
    }
    T bv;
    int bk, bp;
    col_min(am, bv, bk, bp);
    theta = bv > T(0) ? bv : T(0);
    return bp;
  }

  __device__ __forceinline__ void kill_col(int j) {
    const int qj = (j - 1) / NT, tj = (j - 1) - qj * NT;
#pragma unroll
    for (int q = 0; q < CQ; ++q)
      if (q == qj && tid == tj) dead |= 1u << q;
  }

  // only the right-hand side (column 0) is read this way by the driver
  __device__ __forceinline__ T get(int i, int j) {
    (void)j;
    return rhs_get(i);
  }
  __device__ __forceinline__ void set(int i, int j, T val) {
    (void)j;
    rhs_set(i, val);
  }

  __device__ __forceinline__ void mark_redund(int r) {
    __syncthreads();
    if (tid == 0) redund[r] = 1;
    rhs_set(r, T(0));
    __syncthreads();
  }

  __device__ __forceinline__ int min_rhs_row() {
    __syncthreads();
    ArgMin<T> am;
#pragma unroll
    for (int p = 0; p < RP; ++p) {
      const int i = lane + 64 * p;
      const bool in = i >= 2 && i < R;
      am.add(in && !redund[i], rhs[p], i, i);
    }
    T bv;
    int bk, bp;
    row_min(am, bv, bk, bp);
    return bp;
  }

  // auxiliary column x0 (column C-1): -1 on live constraint rows, 1 on the phase-I row;
  // the phase-I row is 0 elsewhere (its rhs included)
  __device__ __forceinline__ void setup_x0() {
    __syncthreads();
    const int x0 = C - 1, qx = (x0 - 1) / NT, tx = (x0 - 1) - qx * NT;
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      if (q == qx && tid == tx) {
#pragma unroll
        for (int i = 0; i < RCAP; ++i)
          v[q][i] = (i >= 2 && i < R) ? (redund[i] ? T(0) : T(-1)) : (i == 1 ? T(1) : T(0));
        dead &= ~(1u << q);
      }
    }
#pragma unroll
    for (int q = 0; q < CQ; ++q)
      if (col_of(q) != x0) v[q][1] = T(0);
    rhs_set(1, T(0));
    __syncthreads();
  }

  __device__ __forceinline__ int find_basic(int var) {
    __syncthreads();
    ArgMin<T> am;
#pragma unroll
    for (int p = 0; p < RP; ++p) {
      const int i = lane + 64 * p;
      const bool in = i >= 2 && i < R;
      am.add(in && bsr[p] == var, T(0), i, i);
    }
    T bv;
    int bk, bp;
    row_min(am, bv, bk, bp);
    return bp;
  }

  // block crash entry points (unused: kBlockCrash is false)
  __device__ __forceinline__ int blk_word(int) const { return 0; }
  __device__ __forceinline__ int crash_prepare(int&, const int8_t*, T) { return -1; }

  __device__ __forceinline__ void write_outputs(int lp, const LpDesc& d, const Out& out, int status, int it,
                                                int p1, int unb, T wfin, int last) {
    __syncthreads();
    const T obj = rhs_get(0);
    if (tid == 0) {
      double am = s_am[0];
      for (int w = 1; w < NW; ++w) am = fmax(am, s_am[w]);
      out.amax[lp] = am;
      out.status[lp] = status;
      out.pivots[lp] = it;
      out.p1[lp] = p1;
      out.unb[lp] = unb;
      if (out.lastpiv) out.lastpiv[lp] = last;
      out.obj[lp] = (double)obj;
      out.wfin[lp] = (double)wfin;
    }
    double* x = out.x + d.x;
    int32_t* basis = out.basis + d.basis;
    for (int j = tid; j < n; j += NT) x[j] = 0.0;
    for (int i = tid; i < m; i += NT) basis[i] = bs[2 + i];
    __syncthreads();
    if (wv == 0) {
#pragma unroll
      for (int p = 0; p < RP; ++p) {
        const int i = lane + 64 * p;
        if (i >= 2 && i < R) {
          const int var = bs[i];
          if (var >= 0 && var < n) x[var] = (double)rhs[p];
        }
      }
    }
    // final tableau and nonbasic ids for the FP64 check (B^-1 is in the slack columns)
    if (out.tab) {
      T* tp = (T*)(out.tab + 8 * d.t);
      if (wv == 0) {
#pragma unroll
        for (int p = 0; p < RP; ++p) {
          const int i = lane + 64 * p;
          if (i < R) tp[(int64_t)i * C] = rhs[p];
        }
      }
#pragma unroll
      for (int q = 0; q < CQ; ++q) {
        const int j = col_of(q);
        if (j < C) {
#pragma unroll
          for (int i = 0; i < RCAP; ++i)
            if (i < R) tp[(int64_t)i * C + j] = v[q][i];
        }
      }
      int32_t* np = (int32_t*)(out.tab + 8 * (d.t + (int64_t)R * C));
      for (int j = tid; j < C; j += NT) np[j] = nb[j];
    }
  }
};

// WPE: minimum waves per SIMD the register allocator must allow (1: no constraint)
template <typename T, int NW, int CQ, int RCAP, int WPE>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(WPE))) void k_wave(
    const LpDesc* __restrict__ desc, const double* __restrict__ pool, const int8_t* __restrict__ spool,
    const int32_t* __restrict__ idx, int32_t count, const int32_t* count_dev, int32_t* queue, Out out,
    Opt<T> o, const int32_t* __restrict__ warm_pool) {
  using Tab = WaveTab<T, NW, CQ, RCAP>;
  __shared__ __align__(64) T colbuf[Tab::kRcap];
  __shared__ T scal[2];
  __shared__ int bs[Tab::kRcap], nb[Tab::CCAP];
  __shared__ uint8_t redund[Tab::kRcap], target[Tab::kRcap + Tab::CCAP];
  __shared__ int wl[Tab::kRcap], colof[Tab::kRcap + Tab::CCAP];
  __shared__ typename KeyOps<T>::K kslots[2 * NW];
  __shared__ double s_am[NW];
  __shared__ int s_q, s_colvar;
  Tab t;
  t.s_am = s_am;
  t.tid = threadIdx.x;
  t.lane = threadIdx.x & 63;
  t.wv = uniform((int)(threadIdx.x >> 6));
  t.phase = 0;
  t.colbuf = colbuf;
  t.scal = scal;
  t.bs = bs; t.nb = nb; t.redund = redund; t.target = target;
  t.wl = wl; t.colof = colof;
  t.kslots = kslots;
  t.colvar = &s_colvar;
  t.blk = -1;
  if (count_dev) count = *count_dev;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
    __syncthreads();
    const int q = uniform(s_q);
    if (q >= count) break;
    const int lp = idx ? idx[q] : q;
    const LpDesc d = desc[lp];
    t.m = d.m; t.n = d.n; t.R = d.m + 2; t.C = d.n + 2;
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
// variant table: (waves per LP, column slots per thread, rows); capacity rows x columns
// 1 + 64*NW*CQ.  A variant exists in a precision (e32 / e64) where its tile leaves room
// for the ~80 VGPRs the state machine needs besides it; w32 / w64: waves-per-SIMD floors
// for the register allocator.
//   id nw cq rows e32 e64 w32 w64
#define BLPG_WAVE_VARIANTS(X)    \
  X(0, 1, 1, 32, 1, 1, 5, 3)     \
  X(1, 1, 1, 64, 1, 0, 3, 2)     \
  X(2, 2, 1, 64, 1, 0, 3, 2)     \
  X(3, 2, 1, 80, 1, 0, 2, 1)     \
  X(4, 2, 1, 112, 1, 0, 2, 1)

constexpr int kWaveVariants = 5;

int wave_variant_count() { return kWaveVariants; }

WaveVariant wave_variant(int precision, int v) {
  switch (v) {
#define X(id, nw, cq, rows, e32, e64, w32, w64) \
  case id: return WaveVariant{nw, cq, rows, (precision == 0 ? e64 : e32) != 0};
    BLPG_WAVE_VARIANTS(X)
#undef X
    default: return WaveVariant{1, 1, 0, false};
  }
}

template <typename T>
static const void* wave_fn(int v) {
  switch (v) {
#define X(id, nw, cq, rows, e32, e64, w32, w64)                                               \
  case id:                                                                                     \
    if constexpr ((sizeof(T) == 4 ? e32 : e64) != 0)                                           \
      return reinterpret_cast<const void*>(                                                    \
          &k_wave<T, nw, cq, rows, sizeof(T) == 4 ? w32 : w64>);                               \
    return nullptr;
    BLPG_WAVE_VARIANTS(X)
#undef X
    default: return nullptr;
  }
}

int wave_max_blocks_per_cu(int precision, int v) {
  const void* fn = precision == 0 ? wave_fn<double>(v) : wave_fn<float>(v);
  if (!fn) return 0;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, wave_variant(precision, v).nw * 64, 0) !=
      hipSuccess)
    return 0;
  return nb;
}

hipError_t launch_wave(const TileLaunch& L, hipStream_t st) {
  const void* fn = L.precision == 0 ? wave_fn<double>(L.variant) : wave_fn<float>(L.variant);
  if (!fn) return hipErrorInvalidValue;
  const int nt = wave_variant(L.precision, L.variant).nw * 64;
  if (L.precision == 0) {
    Opt<double> o{L.eps_d, L.eps_p, L.tol, L.max_iter, L.bland_after};
    void* args[] = {(void*)&L.desc, (void*)&L.pool, (void*)&L.spool, (void*)&L.idx,
                    (void*)&L.count, (void*)&L.count_dev, (void*)&L.queue, (void*)&L.out,
                    (void*)&o, (void*)&L.warm};
    return hipLaunchKernel(fn, dim3(L.grid), dim3(nt), args, 0, st);
  }
  Opt<float> o{(float)L.eps_d, (float)L.eps_p, (float)L.tol, L.max_iter, L.bland_after};
  void* args[] = {(void*)&L.desc, (void*)&L.pool, (void*)&L.spool, (void*)&L.idx,
                  (void*)&L.count, (void*)&L.count_dev, (void*)&L.queue, (void*)&L.out,
                  (void*)&o, (void*)&L.warm};
  return hipLaunchKernel(fn, dim3(L.grid), dim3(nt), args, 0, st);
}

}  // namespace blpg
