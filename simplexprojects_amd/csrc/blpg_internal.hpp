// blpg_internal.hpp -- structures shared by the host library and the HIP kernels.
#pragma once
#include <cstdint>

namespace blpg {

// One LP inside a packed device batch.  Offsets index the batch's FP64 pool
// (A row-major m*n, b m, c n), its int8 sense pool (-1: all '<='), and the output
// pools for x (n doubles) and the basis (m int32).
struct LpDesc {
  int32_t m, n;
  int64_t a, b, c, s;
  int64_t x, basis;
  int64_t t;  // final-tableau dump: 8-byte units into Out::tab ((m+2)(n+2) entries, then
              // the n+2 nonbasic variable ids as int32)
};

// Per-LP outputs of a solve (device arrays, indexed by LP id within the batch).
struct Out {
  int32_t* status;
  int32_t* pivots;
  int32_t* p1;
  int32_t* unb;     // variable id of the entering column when unbounded, else -1
  double* obj;      // tableau objective
  double* wfin;     // phase-I objective at the end of phase I
  double* x;        // pool
  int32_t* basis;   // pool
  double* amax;     // max |A_ij| of the LP's input (written by the solve kernels at set-up)
  int32_t* lastpiv; // the last pivot, row | column << 16 (tableau indices), -1 when there is
                    // none or it cannot be undone from the final state; nullptr: not kept
  char* tab;        // final tableaux (row-major, solve precision) + nonbasic ids, for the
                    // FP64 check; nullptr: not dumped
};

template <typename T>
struct Opt {
  T eps_d, eps_p, tol;
  int32_t max_iter;  // <= 0: 50*(m+n)+1000
  int32_t bland_after;
};

// Result of the FP64 check of a final basis (a verified result's x, the FP64 basic
// solution, overwrites Out::x).
struct Check {
  double* obj;      // FP64 objective c^T x
  double* kkt_p;    // primal residual (relative)
  double* kkt_d;    // dual residual (relative)
  int32_t* ok;      // 1 verified, 0 failed (needs fallback), 2 not checkable (status kept)
  int32_t* fb;      // set to 1 when an FP32 result failed and went to the FP64 fallback
};

constexpr int kStatusUndecided = 1;
constexpr int kStatusInfeasible = 4;
constexpr int kStatusOptimal = 5;
constexpr int kStatusUnbounded = 6;
constexpr int kStatusIterLimit = 100;
constexpr int kStatusNumeric = 101;

}  // namespace blpg
