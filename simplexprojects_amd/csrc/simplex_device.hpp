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
//   ratio    r = argmin (max(T[i][0],0) / T[i][s], i) over rows with T[i][s] > eps_p
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

// Butterfly argmin over the 64 lanes of a wavefront; every lane ends with the result.
// (value, key) is a strict total order, so the result does not depend on the order of
// combination; `p` is a payload carried with the winning key.
template <typename T>
__device__ __forceinline__ void wave_argmin(T& v, int& k, int& p) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const T ov = __shfl_xor(v, off);
    const int ok = __shfl_xor(k, off);
    const int op = __shfl_xor(p, off);
    if (better(ov, ok, v, k)) { v = ov; k = ok; p = op; }
  }
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
      for (int q = 1; q < NW; ++q)
        if (better(s[q].v, s[q].k, v, k)) { v = s[q].v; k = s[q].k; p = s[q].p; }
      phase ^= 1;
    }
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

// simplex iterations on objective row `orow`:
// 0 optimal, 1 unbounded (s_out = entering column), 2 iteration cap, 3 numeric trouble
template <typename T, class Tab>
__device__ int simplex_loop(Tab& t, int orow, const Opt<T>& o, int maxit, int& it, int* p1,
                            int& s_out) {
  int deg = 0;
  for (;;) {
    if (it >= maxit) return 2;
    const bool bland = deg >= o.bland_after;
    const int s = t.pricing(orow, bland, o.eps_d);
    if (s < 0) return 0;
    T th;
    const int r = t.ratio(s, bland, o.eps_p, th);
    if (r < 0) { s_out = s; return orow == 0 ? 1 : 3; }
    t.pivot(r, s);
    ++it;
    if (p1) ++*p1;
    deg = (th == T(0)) ? deg + 1 : 0;
  }
}

// Phase I (equality pivots, then one auxiliary column) and phase II.  Mirrors the
// oracle's solve_impl statement by statement.
template <typename T, class Tab>
__device__ void run_simplex(Tab& t, int m, int n, const int8_t* sense, const Opt<T>& o,
                            int& status, int& it, int& p1, int& unb, T& wfin) {
  const int maxit = o.max_iter > 0 ? o.max_iter : 50 * (m + n) + 1000;
  const int C = n + 2;
  it = 0; p1 = 0; status = kStatusOptimal; unb = -1; wfin = T(0);

  if (sense) {
    for (int i = 0; i < m && status == kStatusOptimal; ++i) {
      if (sense[i] != 1) continue;
      const int r = 2 + i;
      if (it >= maxit) { status = kStatusIterLimit; break; }
      const int j = t.absmax_col(r, o.eps_p);
      if (j < 0) {
        const T bb = t.get(r, 0);
        if (tabs(bb) <= o.tol) t.mark_redund(r);
        else status = kStatusInfeasible;
        continue;
      }
      t.pivot(r, j);
      t.kill_col(j);
      ++it; ++p1;
    }
  }

  if (status == kStatusOptimal) {
    const int rmin = t.min_rhs_row();
    if (rmin >= 0 && t.get(rmin, 0) < -o.tol) {
      t.setup_x0();
      if (it >= maxit) {
        status = kStatusIterLimit;
      } else {
        t.pivot(rmin, C - 1);
        ++it; ++p1;
        int sd = -1;
        const int rc = simplex_loop(t, 1, o, maxit, it, &p1, sd);
        if (rc == 2) status = kStatusIterLimit;
        else if (rc == 3) status = kStatusNumeric;
        else {
          wfin = t.get(1, 0);
          if (wfin < -o.tol) {
            status = kStatusInfeasible;
          } else {
            const int rx = t.find_basic(n + m);
            if (rx >= 0) {
              const int j = t.absmax_col(rx, o.eps_p);
              t.set(rx, 0, T(0));
              if (j >= 0) {
                if (it >= maxit) status = kStatusIterLimit;
                else { t.pivot(rx, j); t.kill_col(j); ++it; ++p1; }
              } else {
                t.mark_redund(rx);
              }
            } else {
              const int j = t.find_nonbasic(n + m);
              if (j >= 0) t.kill_col(j);
            }
          }
        }
      }
    }
  }

  if (status == kStatusOptimal) {
    int s_unb = -1;
    const int rc = simplex_loop(t, 0, o, maxit, it, (int*)nullptr, s_unb);
    if (rc == 1) { status = kStatusUnbounded; unb = t.nb_of(s_unb); }
    else if (rc == 2) status = kStatusIterLimit;
    else if (rc == 3) status = kStatusNumeric;
  }
}

}  // namespace blpg
