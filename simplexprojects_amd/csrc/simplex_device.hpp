// simplex_device.hpp -- device-side pieces shared by every simplex kernel:
// deterministic (value, index) reductions and the phase-I / phase-II driver.
//
// The driver is written once against a tableau "backend" (HBM/LDS-streamed or
// register-tiled).  Every backend performs the same floating-point operations in the
// same order per tableau entry, so all kernels, and the CPU oracle, produce identical
// pivot sequences and bit-identical tableaux.
//
// Rule (the build's own; the reference's is simplex.cu:93-376):
//   pricing  s = argmin (T[orow][j], j) over live columns with T[orow][j] < -eps_d
//            (reference: most negative reduced cost incl. 0, simplex.cu:123-143)
//   ratio    r = argmin (key_i, i) over rows with T[i][s] > eps_p, key_i = theta_i =
//            max(T[i][0],0) / T[i][s] when positive, else -T[i][s] (degenerate rows rank
//            by the largest pivot element)
//            (reference: simplex.cu:186-236, tolerance 1e-10, ">=" ties)
//   no ratio candidate in phase II = unbounded (the reference exits silently, :268)
//   Bland's rule after `bland_after` consecutive degenerate pivots (the reference has
//   no anti-cycling and no iteration cap, :93)
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "blpg_internal.hpp"

namespace blpg {

// lexicographic (value, index) order, evaluated without short-circuit branches
template <typename T>
__device__ __forceinline__ bool better(T v, int k, T bv, int bk) {
  return (v < bv) | ((v == bv) & (k < bk));
}

template <typename T>
__device__ __forceinline__ T inf_of() {
  return __builtin_huge_val();
}
template <>
__device__ __forceinline__ float inf_of<float>() {
  return __builtin_huge_valf();
}

// Wave-uniform copies (SGPR) of values every lane already holds: lets the compiler turn
// comparisons against pivot indices into scalar branches.
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uniform(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double uniform(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ---- (value, index, payload) reductions on packed keys ----
// A candidate (v, k, p) becomes one integer key whose unsigned order is the lexicographic
// order of (v, k): v is mapped to an order-preserving integer (sign-magnitude folded;
// -0 is first turned into +0 so that equal values give equal keys, as `better` compares
// them), followed by k and p in 16 bits each (indices and variable ids are < 8193; 0xFFFF
// encodes "none").  FP32: one 64-bit key; FP64: a 64-bit value word plus a 32-bit index
// word.  Minimising keys is branch-free: a DPP move per word, a compare and selects.
template <typename T>
struct KeyOps;

template <>
struct KeyOps<float> {
  typedef unsigned long long K;
  static __device__ __forceinline__ K none() { return ~0ull; }
  static __device__ __forceinline__ K make(float v, int k, int p) {
    const uint32_t u = __float_as_uint(v + 0.0f);
    const uint32_t o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ((K)o << 32) | ((uint32_t)(k & 0xFFFF) << 16) | (uint32_t)(p & 0xFFFF);
  }
  static __device__ __forceinline__ bool lt(K a, K b) { return a < b; }
  static __device__ __forceinline__ K sel(bool c, K a, K b) { return c ? a : b; }
  template <int CTRL, int ROWS>
  static __device__ __forceinline__ K dpp(K a) {
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)a, (int)(uint32_t)a, CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(a >> 32), (int)(uint32_t)(a >> 32), CTRL,
                                               ROWS, 0xF, false);
    return ((K)(uint32_t)hi << 32) | (uint32_t)lo;
  }
  static __device__ __forceinline__ K lane(K a, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)a, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(a >> 32), l);
    return ((K)hi << 32) | lo;
  }
  static __device__ __forceinline__ float value(K a) {
    const uint32_t o = (uint32_t)(a >> 32);
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
  }
  static __device__ __forceinline__ uint32_t kp(K a) { return (uint32_t)a; }
};

template <>
struct KeyOps<double> {
  struct K {
    unsigned long long v;
    uint32_t kp;
  };
  static __device__ __forceinline__ K none() { return K{~0ull, ~0u}; }
  static __device__ __forceinline__ K make(double v, int k, int p) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v + 0.0);
    const unsigned long long o = (u >> 63) ? ~u : (u | (1ull << 63));
    return K{o, ((uint32_t)(k & 0xFFFF) << 16) | (uint32_t)(p & 0xFFFF)};
  }
  static __device__ __forceinline__ bool lt(const K& a, const K& b) {
    return (a.v < b.v) | ((a.v == b.v) & (a.kp < b.kp));
  }
  static __device__ __forceinline__ K sel(bool c, const K& a, const K& b) {
    return K{c ? a.v : b.v, c ? a.kp : b.kp};
  }
  template <int CTRL, int ROWS>
  static __device__ __forceinline__ K dpp(const K& a) {
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)a.v, (int)(uint32_t)a.v, CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(a.v >> 32), (int)(uint32_t)(a.v >> 32), CTRL,
                                               ROWS, 0xF, false);
    const int q = __builtin_amdgcn_update_dpp((int)a.kp, (int)a.kp, CTRL, ROWS, 0xF, false);
    return K{((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo, (uint32_t)q};
  }
  static __device__ __forceinline__ K lane(const K& a, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)a.v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(a.v >> 32), l);
    const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)a.kp, l);
    return K{((unsigned long long)hi << 32) | lo, q};
  }
  static __device__ __forceinline__ double value(const K& a) {
    const unsigned long long o = a.v;
    return __longlong_as_double((long long)((o >> 63) ? (o & ~(1ull << 63)) : ~o));
  }
  static __device__ __forceinline__ uint32_t kp(const K& a) { return a.kp; }
};

// DPP controls (gfx9): quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror, row_mirror
// (once quads / half-rows hold equal keys the mirrors act as xor-4 / xor-8), and
// row_bcast:15 / row_bcast:31 (lane 15 of each row, lane 31, into the following rows).
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140,
              kDppBcast15 = 0x142, kDppBcast31 = 0x143;

// Minimum key over the 64 lanes of a wavefront (all lanes active): four steps leave each
// 16-lane row with its minimum, two row broadcasts fold the rows into lane 63, whose key
// is returned to every lane (wave-uniform, in scalar registers).
template <typename T>
__device__ __forceinline__ typename KeyOps<T>::K wave_min_key(typename KeyOps<T>::K a) {
  using KO = KeyOps<T>;
  typename KO::K b;
  b = KO::template dpp<kDppXor1, 0xF>(a);       a = KO::sel(KO::lt(b, a), b, a);
  b = KO::template dpp<kDppXor2, 0xF>(a);       a = KO::sel(KO::lt(b, a), b, a);
  b = KO::template dpp<kDppHalfMirror, 0xF>(a); a = KO::sel(KO::lt(b, a), b, a);
  b = KO::template dpp<kDppMirror, 0xF>(a);     a = KO::sel(KO::lt(b, a), b, a);
  b = KO::template dpp<kDppBcast15, 0xA>(a);    a = KO::sel(KO::lt(b, a), b, a);
  b = KO::template dpp<kDppBcast31, 0xC>(a);    a = KO::sel(KO::lt(b, a), b, a);
  return KO::lane(a, 63);
}

template <typename T>
__device__ __forceinline__ void key_decode(const typename KeyOps<T>::K& a, T& v, int& k, int& p) {
  using KO = KeyOps<T>;
  const uint32_t q = KO::kp(a);
  v = KO::value(a);
  k = (q >> 16) == 0xFFFFu ? INT_MAX : (int)(q >> 16);
  p = (q & 0xFFFFu) == 0xFFFFu ? -1 : (int)(q & 0xFFFFu);
}

// Running minimum of one lane's candidates: add(v, k, p) when `take`.
template <typename T>
struct ArgMin {
  using KO = KeyOps<T>;
  typename KO::K key;
  __device__ __forceinline__ ArgMin() : key(KO::none()) {}
  __device__ __forceinline__ void add(bool take, T v, int k, int p) {
    const typename KO::K c = KO::make(v, k, p);
    key = KO::sel(take & KO::lt(c, key), c, key);
  }
};

// Wavefront minimum of every lane's running minimum, decoded (all lanes active).
template <typename T>
__device__ __forceinline__ void wave_finish(const ArgMin<T>& a, T& v, int& k, int& p) {
  using KO = KeyOps<T>;
  const typename KO::K m = wave_min_key<T>(a.key);
  if (KO::kp(m) == ~0u) { v = inf_of<T>(); k = INT_MAX; p = -1; return; }
  key_decode<T>(m, v, k, p);
}

// Argmin over the 64 lanes of a wavefront (all lanes active) of (v, k) in lexicographic
// order, carrying payload p; every lane ends with the same (wave-uniform) result.  No
// candidate: v = +inf, k = INT_MAX, p = -1.
template <typename T>
__device__ __forceinline__ void wave_argmin(T& v, int& k, int& p) {
  using KO = KeyOps<T>;
  const bool real = k != INT_MAX;
  typename KO::K a = KO::sel(real, KO::make(v, k, p), KO::none());
  a = wave_min_key<T>(a);
  if (KO::kp(a) == ~0u) { v = inf_of<T>(); k = INT_MAX; p = -1; return; }
  key_decode<T>(a, v, k, p);
}

template <typename T>
struct RedSlot {
  typename KeyOps<T>::K key;
};

// Workgroup argmin: wavefront minimum, then the NW per-wave minima through LDS.
// Slots are double-buffered so one barrier per reduction suffices.
template <typename T, int NW>
struct Reducer {
  RedSlot<T>* slots;  // 2*NW entries in LDS
  int phase;
  __device__ void init(RedSlot<T>* s) { slots = s; phase = 0; }
  __device__ __forceinline__ void argmin(T& v, int& k, int& p) {
    using KO = KeyOps<T>;
    const bool real = k != INT_MAX;
    typename KO::K a = KO::sel(real, KO::make(v, k, p), KO::none());
    a = wave_min_key<T>(a);
    if constexpr (NW > 1) {
      RedSlot<T>* s = slots + phase * NW;
      const int w = threadIdx.x >> 6;
      if ((threadIdx.x & 63) == 0) s[w].key = a;
      __syncthreads();
      a = s[0].key;
#pragma unroll
      for (int q = 1; q < NW; ++q) {
        const typename KO::K b = s[q].key;
        a = KO::sel(KO::lt(b, a), b, a);
      }
      phase ^= 1;
    }
    if (KO::kp(a) == ~0u) { v = inf_of<T>(); k = INT_MAX; p = -1; }
    else key_decode<T>(a, v, k, p);
    v = uniform(v);
    k = uniform(k);
    p = uniform(p);
  }
};

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T>
__device__ __forceinline__ T tabs(T a) {
  return a < T(0) ? -a : a;
}

// Initial tableau entry (i, j) of an LP; identical to the oracle's construction.
//   row 0: -c (the reference's Zj - Cj row, simplex.cu:886); row 1: phase-I row (0);
//   row 2+r: constraint r ('>=' rows negated), column 0: b.
template <typename T>
__device__ __forceinline__ T init_entry(int i, int j, int m, int n, const double* A,
                                        const double* b, const double* c, const int8_t* sense) {
  if (i == 0) return (j >= 1 && j <= n) ? T(-c[j - 1]) : T(0);
  if (i == 1 || i >= m + 2) return T(0);
  const int r = i - 2;
  const bool neg = sense && sense[r] == 2;
  if (j == 0) return T(neg ? -b[r] : b[r]);
  if (j <= n) {
    const double a = A[(int64_t)r * n + (j - 1)];
    return T(neg ? -a : a);
  }
  return T(0);
}

// Phase I (equality pivots, then one auxiliary column) and phase II, written as a state
// machine so that every backend's pivot / pricing / ratio code is inlined exactly once.
// Statement for statement the same decisions, in the same order, as the oracle's
// solve_impl (oracle/simplex_core.inc), which is what makes the results bit-identical.
enum : int {
  kStCrash = 0, kStEq = 1, kStAux = 2, kStP1 = 3, kStP1End = 4, kStP2 = 5, kStDualChk = 6,
  kStDual = 7
};

// `warm` (optional, m entries): a target basis (variable ids, -1 = none).  Before phase I
// its structurals are pivoted in one by one (crash), each on the row of largest |T[i][s]|
// among rows whose basic variable is not itself a target; a structural with no eligible
// nonzero is skipped (its row keeps the slack).  Repair: when the crashed basis is dual
// feasible (no live column with T[0][j] < -eps_d), dual simplex pivots (leaving row: most
// negative rhs below -tol; entering column: smallest max(T[0][j],0) / -T[r][j] over
// T[r][j] < -eps_p, dual-degenerate columns by the largest |T[r][j]|) restore primal
// feasibility; then the auxiliary phase I (if still needed) and phase II finish.
// `last`: the final pivot (row | column << 16), so that a caller can undo it: the variable
// that left at that row is the final nonbasic variable of that column.
template <typename T, class Tab>
__device__ __forceinline__ void run_simplex(Tab& t, int m, int n, const int8_t* sense,
                                            const Opt<T>& o, int& status, int& it, int& p1,
                                            int& unb, T& wfin, int& last,
                                            const int32_t* warm = nullptr) {
  const int maxit = o.max_iter > 0 ? o.max_iter : 50 * (m + n) + 1000;
  const int C = n + 2;
  it = 0; p1 = 0; status = kStatusOptimal; unb = -1; wfin = T(0); last = -1;
  // the last pivot is kept as row | column << 16 where both fit
  const bool keep_last = m + 2 < 65536 && C < 32768;
  int stage = warm ? kStCrash : (sense ? kStEq : kStAux);
  int ei = 0, deg = 0, ci = 0;
  int bt = 0, bnp = 0;  // block crash: next pivot of the prepared block, its pivots
  for (;;) {
    t.loop_fence();
    int pr = -1, ps = -1;
    bool kill_after = false, count_p1 = false;
    if (stage == kStCrash) {
      bool blocked = false;
      if constexpr (Tab::kBlockCrash) {
        // up to KB targets prepared at once, then one pivot per turn from the prepared
        // block through the shared pivot site: the same pivots, in the same order, as the
        // one-at-a-time crash below.  Prepared only when none can reach the iteration limit.
        if (bt < bnp) {
          const int wd = t.blk_word(bt++);
          pr = wd & 0xFFF;
          ps = (wd >> 12) & 0xFFF;
          t.blk = (wd >> 24) & 0x3F;
          kill_after = ((wd >> 30) & 1) != 0;
          count_p1 = true;
          blocked = true;
        } else if (it + Tab::KB <= maxit) {
          bt = 0;
          bnp = t.crash_prepare(ci, sense, o.eps_p);
          if (bnp < 0) {
            bnp = 0;
            stage = sense ? kStEq : kStDualChk;
          }
          continue;
        }
      }
      if (!blocked) {
        int v = -1;
        while (ci < m && v < 0) {
          const int w = t.warm_at(warm, ci);
          ++ci;
          if (w >= 0 && w < n) v = w;
        }
        if (v < 0) { stage = sense ? kStEq : kStDualChk; continue; }
        const int s = t.find_nonbasic(v);
        if (s < 0) continue;
        if (it >= maxit) { status = kStatusIterLimit; break; }
        const int r = t.absmax_row(s, o.eps_p);
        if (r < 0) continue;
        const int lv = t.bs_now(r);  // absmax_row ended at a barrier, no pivot since
        pr = r; ps = s; count_p1 = true;
        kill_after = lv >= n && lv < n + m && sense && sense[lv - n] == 1;
      }
    } else if (stage == kStEq) {
      while (ei < m && sense[ei] != 1) ++ei;
      if (ei >= m) { stage = warm ? kStDualChk : kStAux; continue; }
      const int r = 2 + ei;
      ++ei;
      if (t.bs_of(r) != n + r - 2) continue;  // its slack already left (crash)
      if (it >= maxit) { status = kStatusIterLimit; break; }
      const int j = t.absmax_col(r, o.eps_p);
      if (j < 0) {
        const T bb = t.get(r, 0);
        if (tabs(bb) <= o.tol) { t.mark_redund(r); continue; }
        status = kStatusInfeasible;
        break;
      }
      pr = r; ps = j; kill_after = true; count_p1 = true;
    } else if (stage == kStDualChk) {
      deg = 0;
      stage = t.pricing(0, false, o.eps_d) < 0 ? kStDual : kStAux;
      continue;
    } else if (stage == kStDual) {
      const bool bland = deg >= o.bland_after;
      const int r = t.dual_row(bland, o.tol);
      if (r < 0) { stage = kStAux; continue; }
      if (it >= maxit) { status = kStatusIterLimit; break; }
      T th;
      const int s = t.dual_ratio(r, bland, o.eps_p, th);
      if (s < 0) { stage = kStAux; continue; }
      pr = r; ps = s; count_p1 = true;
      deg = (th == T(0)) ? deg + 1 : 0;
    } else if (stage == kStAux) {
      const int rmin = t.min_rhs_row();
      deg = 0;
      if (rmin >= 0 && t.get(rmin, 0) < -o.tol) {
        t.setup_x0();
        if (it >= maxit) { status = kStatusIterLimit; break; }
        pr = rmin; ps = C - 1; count_p1 = true;
        stage = kStP1;
      } else {
        stage = kStP2;
        continue;
      }
    } else if (stage == kStP1 || stage == kStP2) {
      if (it >= maxit) { status = kStatusIterLimit; break; }
      const bool bland = deg >= o.bland_after;
      const int s = t.pricing(stage == kStP1 ? 1 : 0, bland, o.eps_d);
      if (s < 0) {
        if (stage == kStP1) { stage = kStP1End; continue; }
        status = kStatusOptimal;
        break;
      }
      T th;
      const int r = t.ratio(s, bland, o.eps_p, th);
      if (r < 0) {
        if (stage == kStP2) { status = kStatusUnbounded; unb = t.nb_of(s); }
        else status = kStatusNumeric;
        break;
      }
      pr = r; ps = s; count_p1 = stage == kStP1;
      deg = (th == T(0)) ? deg + 1 : 0;
    } else {  // kStP1End
      wfin = t.get(1, 0);
      if (wfin < -o.tol) { status = kStatusInfeasible; break; }
      const int rx = t.find_basic(n + m);
      stage = kStP2;
      deg = 0;
      if (rx >= 0) {
        const int j = t.absmax_col(rx, o.eps_p);
        t.set(rx, 0, T(0));
        if (j < 0) { t.mark_redund(rx); continue; }
        if (it >= maxit) { status = kStatusIterLimit; break; }
        pr = rx; ps = j; kill_after = true; count_p1 = true;
      } else {
        // x0 left the basis during phase I: its column goes (through the shared kill site
        // below -- a second inlined kill costs the whole kernel registers)
        const int j = t.find_nonbasic(n + m);
        if (j < 0) continue;
        ps = j;
        kill_after = true;
      }
    }
    if (pr >= 0) {
      t.pivot(pr, ps);
      ++it;
      if (count_p1) ++p1;
      last = keep_last ? pr | (ps << 16) : -1;
    }
    if (kill_after) {
      t.kill_col(ps);
      last = -1;  // a killed column's pivot is not undone
    }
  }
}

}  // namespace blpg
