/*
 * ref_rule.c -- restatement of the reference kernel's pivot rule (TEST INFRASTRUCTURE
 * ONLY; see oracle.h).  Follows BLPG-Solver/src/simplex.cu:35-385 (mykernel) on the
 * reference's own tableau (simplex.cu:485-534, 880-890), phase II only.
 *
 * Tableau: column-major, M = m+2 rows, N = n+m+3 columns, element (r,c) at r + c*M.
 *   column 0 basic variable index (1-based, as a double), column 1 c_B,
 *   columns 2..n+1 A, n+2..n+m+1 slack identity, N-1 b;
 *   row M-2 reduced costs "Zj-Cj" (initially -c), row M-1 copy of c.
 * The two reductions are simulated exactly, including the reference's block size
 * (32*floor(N/32)+32, simplex.cu:585-587), the "<= 0" admission of zero reduced costs
 * (simplex.cu:127), the ">=" right-hand tie rule of the ratio reduction (:225) and the
 * odd-stride step (:139-141, :232-234).  The products in the elimination are written
 * as fma, the contraction nvcc applies by default to `S - col1*S_r` (simplex.cu:310-312).
 */
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define REF(r, c) T[(size_t)(r) + (size_t)(c) * (size_t)M]

int orc_ref_rule_solve(int32_t m, int32_t n, const double *A, const double *b, const double *c,
                       int32_t max_iter, double *obj, int32_t *pivots) {
  if (m < 1 || n < 1) return -1;
  for (int i = 0; i < m; ++i)
    if (b[i] < 0) return -1; /* two-phase branch (simplex.cu:602-875) not restated */
  const int M = m + 2, N = n + m + 3;
  double *T = (double *)calloc((size_t)M * N, sizeof(double));
  int tpb = 32 * (N / 32) + 32; /* simplex.cu:585 */
  if (tpb > 1024) tpb = 1024;   /* maxThreadsPerBlock, simplex.cu:586-587 */
  double *Rd = (double *)malloc(sizeof(double) * tpb);
  int *Ri = (int *)malloc(sizeof(int) * tpb);
  double *col1 = (double *)malloc(sizeof(double) * M);

  /* simplex.cu:506-534 */
  for (int i = 0; i < m; ++i) {
    REF(i, 0) = (double)(1 + n + i);
    for (int j = 0; j < n; ++j) REF(i, 2 + j) = A[(size_t)i * n + j];
    REF(i, n + 2 + i) = 1.0;
    REF(i, N - 1) = b[i];
  }
  /* simplex.cu:880-890 */
  for (int j = 0; j < n; ++j) {
    REF(M - 2, 2 + j) = -c[j];
    REF(M - 1, 2 + j) = c[j];
  }

  const double IMAX = (double)INT_MAX;
  int done = 0, piv = 0;
  while (!done && piv < max_iter) {
    done = 1; /* c = true, simplex.cu:105 */
    /* ---- first reduction: entering column (simplex.cu:108-167) ---- */
    for (int t = 0; t < tpb; ++t) {
      Rd[t] = (t >= 2 && t < N - 1) ? REF(M - 2, t) : IMAX;
      Ri[t] = t;
    }
    int notEntered = 1;
    for (int i = tpb / 2; i > 0;) {
      for (int tid = 0; tid < i; ++tid) {
        if (Rd[tid] > Rd[tid + i] && Rd[tid + i] <= 0) {
          Rd[tid] = Rd[tid + i];
          Ri[tid] = Ri[tid + i];
          notEntered = 0;
        }
      }
      i >>= 1;
      if (i != 1 && (i % 2) != 0) i = i + 1;
    }
    const int p = notEntered ? -1 : Ri[0];
    int row = -1;
    if (p >= 0) {
      /* ---- second reduction: leaving row (simplex.cu:170-257) ---- */
      int notEntered2 = 1;
      for (int t = 0; t < tpb; ++t) {
        Ri[t] = t;
        Rd[t] = IMAX;
        if (t < M - 2) {
          const double a = REF(t, p), bb = REF(t, N - 1);
          if (a > 1e-10 && bb >= 0) { /* simplex.cu:193 */
            Rd[t] = bb / a;
            notEntered2 = 0;
          }
        }
      }
      for (int s = tpb / 2; s > 0;) {
        for (int tid = 0; tid < s; ++tid) {
          if (Rd[tid] >= Rd[tid + s]) { /* simplex.cu:225 */
            Rd[tid] = Rd[tid + s];
            Ri[tid] = Ri[tid + s];
          }
        }
        s >>= 1;
        if (s != 1 && (s % 2) != 0) s = s + 1;
      }
      if (!notEntered2 && Rd[0] != IMAX) row = Ri[0];
    }
    if (row > -1 && p != -1) {
      /* ---- pivot (simplex.cu:268-318) ---- */
      for (int t = 0; t < M; ++t) col1[t] = REF(t, p);
      REF(row, 1) = REF(M - 1, p);
      REF(row, 0) = (double)(p - 1);
      for (int j = 2; j < N; ++j) REF(row, j) = REF(row, j) / col1[row];
      for (int t = 0; t < M - 1; ++t) {
        if (t == row) continue;
        for (int j = 2; j < N; ++j) REF(t, j) = fma(-col1[t], REF(row, j), REF(t, j));
      }
      ++piv;
      /* ---- termination scan (simplex.cu:334-341) ---- */
      for (int j = 2; j < N - 1; ++j)
        if (REF(M - 2, j) < 0.0) { done = 0; break; }
    }
  }
  *obj = REF(M - 2, N - 1); /* simplex.cu:380 */
  *pivots = piv;
  free(T); free(Rd); free(Ri); free(col1);
  return 0;
}
