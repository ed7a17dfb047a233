/*
 * oracle.h -- CPU oracle for the batched dense simplex (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the checker, never the product: only tests/, the smoke test and
 * the bench's cpu_baseline leg load it.  It restates, in plain C, the algorithm the
 * HIP kernels run, so that GPU results can be compared bit for bit.
 *
 *   orc_solve_d / orc_solve_f   the build's pivot rule (condensed tableau, Dantzig
 *                               pricing, lowest-index ties, Bland after a run of
 *                               degenerate pivots, two-phase with one auxiliary
 *                               column) in FP64 / FP32.  The LP form and the
 *                               maximisation convention follow the reference's
 *                               Simplex class (BLPG-Solver/src/simplex.cu:463-550,
 *                               880-890; min->max flip GPU_Tester.cu:119-127) and
 *                               its GLPK wrapper (glpk_lp_solver.cpp:102-145:
 *                               rows GLP_UP, columns GLP_LO 0, GLP_MAX).
 *   orc_ref_rule_solve          restatement of the reference kernel's own rule on
 *                               its own tableau layout (simplex.cu:35-385, phase II
 *                               only), including its tree-reduction tie-breaks.
 *   orc_kkt_check               FP64 re-solve of B x_B = b, B^T y = c_B from a basis
 *                               and primal/dual residuals (no reference counterpart).
 *
 * Parity pins (see tests/): HiGHS golden fixtures, netlib known answers quoted at
 * BLPG-Solver/src/parseBenchmark.cpp:17-25, hand LPs (parseBenchmark.cpp:128-319),
 * vertex enumeration, exact-rational replay, Beale's cycling example.
 */
#ifndef BLPG_ORACLE_H
#define BLPG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: GLPK's (glpk.h GLP_OPT=5, GLP_NOFEAS=4, GLP_UNBND=6) + two of ours */
#define ORC_OPTIMAL 5
#define ORC_INFEASIBLE 4
#define ORC_UNBOUNDED 6
#define ORC_ITER_LIMIT 100
#define ORC_NUMERIC 101

typedef struct {
  int32_t m, n;
  const double *A;     /* m*n row-major */
  const double *b;     /* m */
  const double *c;     /* n, maximise c^T x */
  const int8_t *sense; /* m: 0 '<=', 1 '=', 2 '>='; NULL = all '<=' */
} orc_lp;

typedef struct {
  double eps_d;    /* entering threshold: d_j < -eps_d */
  double eps_p;    /* pivot threshold: T[i][s] > eps_p */
  double tol_feas; /* phase-I decisions */
  int32_t max_iter;
  int32_t bland_after;
} orc_opts;

typedef struct {
  int32_t status, pivots, phase1_pivots, unb_var;
  double objective; /* tableau objective value */
  double w_final;   /* phase-I objective at the end of phase I (0 if none) */
  double *x;        /* n, caller-allocated (may be NULL) */
  int32_t *basis;   /* m, caller-allocated (may be NULL) */
} orc_res;

/* returns 0, or -1 on invalid input */
int orc_solve_d(const orc_lp *lp, const orc_opts *o, orc_res *res);
int orc_solve_f(const orc_lp *lp, const orc_opts *o, orc_res *res);
/* same, but also copies the final (m+2)x(n+2) tableau (row-major) into tab_out and the
 * variable id of every column (n+2; -1 for the rhs column) into nb_out */
int orc_solve_tab_d(const orc_lp *lp, const orc_opts *o, orc_res *res, double *tab_out,
                    int32_t *nb_out);
int orc_solve_tab_f(const orc_lp *lp, const orc_opts *o, orc_res *res, float *tab_out,
                    int32_t *nb_out);

/* Warm start: `warm` (m variable ids, -1 = none) is a target basis whose structurals are
 * pivoted in first (crash: largest |T[i][s]| among rows whose basic variable is not a
 * target; skipped when none exceeds eps_p), then phase I / II repair and finish. */
int orc_solve_warm_d(const orc_lp *lp, const orc_opts *o, const int32_t *warm, orc_res *res);
int orc_solve_warm_f(const orc_lp *lp, const orc_opts *o, const int32_t *warm, orc_res *res);

/* Batch driver for the CPU baseline: `count` LPs laid out as in the packed pool
 * (offsets in doubles / int8 entries), solved with `threads` OpenMP threads. */
int orc_solve_batch_d(int64_t count, const int32_t *m, const int32_t *n, const int64_t *a_off,
                      const int64_t *b_off, const int64_t *c_off, const int64_t *s_off,
                      const double *pool, const int8_t *spool, const orc_opts *o,
                      int32_t *status, int32_t *pivots, double *obj, int threads);

/* Reference kernel rule (simplex.cu:35-385), phase II only (b >= 0 required).
 * Returns 0 on success; *obj = T[M-2][N-1]; *pivots = number of pivots. */
int orc_ref_rule_solve(int32_t m, int32_t n, const double *A, const double *b, const double *c,
                       int32_t max_iter, double *obj, int32_t *pivots);

/* KKT check from a basis (m variable ids; id<n structural, n<=id<n+m slack of row id-n).
 * Outputs x (n), y (m), objective, primal / dual residuals.  Returns 0, or -1 when the
 * basis matrix is singular / ids invalid. */
int orc_kkt_check(const orc_lp *lp, const int32_t *basis, double *x, double *y, double *obj,
                  double *res_primal, double *res_dual);

#ifdef __cplusplus
}
#endif
#endif
