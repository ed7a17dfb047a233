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

template <typename T>
__device__ __forceinline__ bool better(T v, int k, T bv, int bk) {
  return v < bv || (v == bv && k < bk);
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

// DPP lane movement inside 16-lane rows (gfx9 encodings): quad_perm [1,0,3,2] and
// [2,3,0,1], row_half_mirror (lane i -> 7-i in each 8-lane half), row_mirror (i -> 15-i).
// Once the lanes of a quad / half-row hold equal keys, the mirrors act as xor-4 / xor-8.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

template <int CTRL>
__device__ __forceinline__ int dpp(int x) {
  return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float x) {
  return __int_as_float(dpp<CTRL>(__float_as_int(x)));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp<CTRL>((int)(b & 0xffffffffll));
  const int hi = dpp<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int CTRL, typename T>
__device__ __forceinline__ void dpp_step(T& v, int& k, int& p) {
  const T ov = dpp<CTRL>(v);
  const int ok = dpp<CTRL>(k);
  const int op = dpp<CTRL>(p);
  const bool b = better(ov, ok, v, k);
  v = b ? ov : v;
  k = b ? ok : k;
  p = b ? op : p;
}

__device__ __forceinline__ int lane_read(int x, int lane) { return __builtin_amdgcn_readlane(x, lane); }
__device__ __forceinline__ float lane_read(float x, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}
__device__ __forceinline__ double lane_read(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Argmin over the 64 lanes of a wavefront (all lanes active).  Four DPP steps leave every
// lane of each 16-lane row with the row's winner; the four row winners are read into
// scalar registers and combined.  Every lane ends with the same (wave-uniform) result.
// (value, key) is a strict total order, so the result does not depend on the order of
// combination; `p` is a payload carried with the winning key.
template <typename T>
__device__ __forceinline__ void wave_argmin(T& v, int& k, int& p) {
  dpp_step<kDppXor1>(v, k, p);
  dpp_step<kDppXor2>(v, k, p);
  dpp_step<kDppHalfMirror>(v, k, p);
  dpp_step<kDppMirror>(v, k, p);
  T bv = lane_read(v, 0);
  int bk = lane_read(k, 0), bp = lane_read(p, 0);
#pragma unroll
  for (int row = 1; row < 4; ++row) {
    const T rv = lane_read(v, 16 * row);
    const int rk = lane_read(k, 16 * row), rp = lane_read(p, 16 * row);
    const bool b = better(rv, rk, bv, bk);
    bv = b ? rv : bv;
    bk = b ? rk : bk;
    bp = b ? rp : bp;
  }
  v = bv;
  k = bk;
  p = bp;
}

template <typename T>
struct RedSlot {
  T v;
  int k, p;
};

// Workgroup argmin: wavefront butterfly, then the NW per-wave winners through LDS.
// Slots are double-buffered so one barrier per reduction suffices.
template <typename T, int NW>
struct Reducer {
  RedSlot<T>* slots;  // 2*NW entries in LDS
  int phase;
  __device__ void init(RedSlot<T>* s) { slots = s; phase = 0; }
  __device__ __forceinline__ void argmin(T& v, int& k, int& p) {
    wave_argmin(v, k, p);
    if constexpr (NW > 1) {
      RedSlot<T>* s = slots + phase * NW;
      const int w = threadIdx.x >> 6;
      if ((threadIdx.x & 63) == 0) { s[w].v = v; s[w].k = k; s[w].p = p; }
      __syncthreads();
      v = s[0].v; k = s[0].k; p = s[0].p;
#pragma unroll
      for (int q = 1; q < NW; ++q) {
        const T qv = s[q].v;
        const int qk = s[q].k, qp = s[q].p;
        const bool b = better(qv, qk, v, k);
        v = b ? qv : v;
        k = b ? qk : k;
        p = b ? qp : p;
      }
      phase ^= 1;
    }
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
template <typename T, class Tab>
__device__ __forceinline__ void run_simplex(Tab& t, int m, int n, const int8_t* sense,
                                            const Opt<T>& o, int& status, int& it, int& p1,
                                            int& unb, T& wfin, const int32_t* warm = nullptr) {
  const int maxit = o.max_iter > 0 ? o.max_iter : 50 * (m + n) + 1000;
  const int C = n + 2;
  it = 0; p1 = 0; status = kStatusOptimal; unb = -1; wfin = T(0);
  int stage = warm ? kStCrash : (sense ? kStEq : kStAux);
  int ei = 0, deg = 0, ci = 0;
  for (;;) {
    int pr = -1, ps = -1;
    bool kill_after = false, count_p1 = false;
    if (stage == kStCrash) {
      int v = -1;
      while (ci < m && v < 0) {
        const int w = warm[ci];
        ++ci;
        if (w >= 0 && w < n) v = w;
      }
      if (v < 0) { stage = sense ? kStEq : kStDualChk; continue; }
      const int s = t.find_nonbasic(v);
      if (s < 0) continue;
      if (it >= maxit) { status = kStatusIterLimit; break; }
      const int r = t.absmax_row(s, o.eps_p);
      if (r < 0) continue;
      const int lv = t.bs_of(r);
      pr = r; ps = s; count_p1 = true;
      kill_after = lv >= n && lv < n + m && sense && sense[lv - n] == 1;
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
        const int j = t.find_nonbasic(n + m);
        if (j >= 0) t.kill_col(j);
        continue;
      }
    }
    t.pivot(pr, ps);
    ++it;
    if (count_p1) ++p1;
    if (kill_after) t.kill_col(ps);
  }
}

}  // namespace blpg
