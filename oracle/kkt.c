/*
 * kkt.c -- FP64 basis re-solve and KKT residuals (TEST INFRASTRUCTURE ONLY; see oracle.h).
 * No reference counterpart: the reference returns only T[M-2][N-1] (simplex.cu:380).
 *
 * For max c^T x, A'x + s = b', x >= 0, s_i >= 0 (<= rows) / s_i = 0 (= rows), where
 * A', b' are A, b with '>=' rows negated, and a basis of m variable ids:
 *   B x_B = b'   (B: columns of [A' | I] for the basic ids), B^T y = c_B,
 *   primal residual = max( -x_j, -s_i (<= rows), |s_i| (= rows), |b' - A'x - s|_inf )
 *                     / (1 + |b|_inf + |A|_max |x_B|_inf),
 *   dual residual   = max over nonbasic, non-fixed variables of (c_j - y^T a'_j)_+
 *                     / (1 + |c|_inf + |A|_max |y|_inf), with c = 0, a' = e_i for slacks.
 * LU with partial pivoting, dense, O(m^3).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static int lu_factor(double *B, int m, int *piv) {
  double amax = 0;
  for (int k = 0; k < m * m; ++k) amax = fmax(amax, fabs(B[k]));
  for (int k = 0; k < m; ++k) {
    int p = k;
    double best = fabs(B[(size_t)k * m + k]);
    for (int i = k + 1; i < m; ++i) {
      const double v = fabs(B[(size_t)i * m + k]);
      if (v > best) { best = v; p = i; }
    }
    if (!(best > 1e-13 * (amax > 0 ? amax : 1))) return -1;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < m; ++j) {
        const double t = B[(size_t)k * m + j];
        B[(size_t)k * m + j] = B[(size_t)p * m + j];
        B[(size_t)p * m + j] = t;
      }
    const double d = B[(size_t)k * m + k];
    for (int i = k + 1; i < m; ++i) {
      const double l = B[(size_t)i * m + k] / d;
      B[(size_t)i * m + k] = l;
      for (int j = k + 1; j < m; ++j) B[(size_t)i * m + j] -= l * B[(size_t)k * m + j];
    }
  }
  return 0;
}

/* solve LU x = rhs (row-permuted) in place */
static void lu_solve(const double *LU, int m, const int *piv, double *x) {
  for (int k = 0; k < m; ++k)
    if (piv[k] != k) { const double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < i; ++j) x[i] -= LU[(size_t)i * m + j] * x[j];
  for (int i = m - 1; i >= 0; --i) {
    for (int j = i + 1; j < m; ++j) x[i] -= LU[(size_t)i * m + j] * x[j];
    x[i] /= LU[(size_t)i * m + i];
  }
}

/* solve (LU)^T y = rhs with the row permutation: B = P^T L U  =>  B^T = U^T L^T P */
static void lu_solve_t(const double *LU, int m, const int *piv, double *y) {
  for (int i = 0; i < m; ++i) { /* U^T z = rhs */
    for (int j = 0; j < i; ++j) y[i] -= LU[(size_t)j * m + i] * y[j];
    y[i] /= LU[(size_t)i * m + i];
  }
  for (int i = m - 1; i >= 0; --i) /* L^T w = z */
    for (int j = i + 1; j < m; ++j) y[i] -= LU[(size_t)j * m + i] * y[j];
  for (int k = m - 1; k >= 0; --k) /* y = P^T w */
    if (piv[k] != k) { const double t = y[k]; y[k] = y[piv[k]]; y[piv[k]] = t; }
}

int orc_kkt_check(const orc_lp *lp, const int32_t *basis, double *x, double *y, double *obj,
                  double *res_primal, double *res_dual) {
  const int m = lp->m, n = lp->n;
  double *B = (double *)calloc((size_t)m * m, sizeof(double));
  double *xb = (double *)malloc(sizeof(double) * m);
  double *s = (double *)calloc(m, sizeof(double));
  int *piv = (int *)malloc(sizeof(int) * m);
  unsigned char *inb = (unsigned char *)calloc(n + m, 1);
  int rc = 0;
  for (int k = 0; k < m; ++k) {
    const int v = basis[k];
    if (v < 0 || v >= n + m || inb[v]) { rc = -1; goto out; }
    inb[v] = 1;
    for (int i = 0; i < m; ++i) {
      const double sg = (lp->sense && lp->sense[i] == 2) ? -1.0 : 1.0;
      B[(size_t)i * m + k] = v < n ? sg * lp->A[(size_t)i * n + v] : (v - n == i ? 1.0 : 0.0);
    }
  }
  if (lu_factor(B, m, piv) != 0) { rc = -1; goto out; }
  for (int i = 0; i < m; ++i) xb[i] = (lp->sense && lp->sense[i] == 2) ? -lp->b[i] : lp->b[i];
  lu_solve(B, m, piv, xb);
  for (int j = 0; j < n; ++j) x[j] = 0;
  for (int k = 0; k < m; ++k) {
    const int v = basis[k];
    if (v < n) x[v] = xb[k]; else s[v - n] = xb[k];
  }
  double rp = 0;
  for (int j = 0; j < n; ++j) rp = fmax(rp, -x[j]);
  for (int i = 0; i < m; ++i) {
    const int eq = lp->sense && lp->sense[i] == 1;
    rp = fmax(rp, eq ? fabs(s[i]) : -s[i]);
    const double sg = (lp->sense && lp->sense[i] == 2) ? -1.0 : 1.0;
    double r = sg * lp->b[i] - s[i];
    for (int j = 0; j < n; ++j) r -= sg * lp->A[(size_t)i * n + j] * x[j];
    rp = fmax(rp, fabs(r));
  }
  for (int k = 0; k < m; ++k) y[k] = basis[k] < n ? lp->c[basis[k]] : 0.0;
  lu_solve_t(B, m, piv, y);
  double rd = 0;
  for (int j = 0; j < n; ++j) {
    if (inb[j]) continue;
    double r = lp->c[j];
    for (int i = 0; i < m; ++i) {
      const double sg = (lp->sense && lp->sense[i] == 2) ? -1.0 : 1.0;
      r -= y[i] * sg * lp->A[(size_t)i * n + j];
    }
    rd = fmax(rd, r);
  }
  for (int i = 0; i < m; ++i) {
    if (inb[n + i] || (lp->sense && lp->sense[i] == 1)) continue;
    rd = fmax(rd, -y[i]);
  }
  /* scale-aware normalisation, as in the GPU check */
  double bmax = 0, cmax = 0, amax = 0, xm = 0, ym = 0;
  for (int i = 0; i < m; ++i) bmax = fmax(bmax, fabs(lp->b[i]));
  for (int j = 0; j < n; ++j) cmax = fmax(cmax, fabs(lp->c[j]));
  for (int e = 0; e < m * n; ++e) amax = fmax(amax, fabs(lp->A[e]));
  for (int k = 0; k < m; ++k) { xm = fmax(xm, fabs(xb[k])); ym = fmax(ym, fabs(y[k])); }
  rp /= 1.0 + bmax + amax * xm;
  rd /= 1.0 + cmax + amax * ym;
  double z = 0;
  for (int j = 0; j < n; ++j) z += lp->c[j] * x[j];
  *obj = z;
  *res_primal = rp;
  *res_dual = rd;
out:
  free(B); free(xb); free(s); free(piv); free(inb);
  return rc;
}
