/*
 * simplex_oracle.c -- CPU oracle entry points (TEST INFRASTRUCTURE ONLY; see oracle.h).
 * Built with -ffp-contract=off so that every fused multiply-add is the explicit
 * fma()/fmaf() of simplex_core.inc, matching the HIP kernels operation for operation.
 */
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

#define REAL double
#define SFX(x) x##_d
#define FMA fma
#include "simplex_core.inc"
#undef REAL
#undef SFX
#undef FMA

#define REAL float
#define SFX(x) x##_f
#define FMA fmaf
#include "simplex_core.inc"
#undef REAL
#undef SFX
#undef FMA

int orc_solve_d(const orc_lp *lp, const orc_opts *o, orc_res *res) { return solve_impl_d(lp, o, res, NULL, NULL, NULL); }
int orc_solve_f(const orc_lp *lp, const orc_opts *o, orc_res *res) { return solve_impl_f(lp, o, res, NULL, NULL, NULL); }
int orc_solve_tab_d(const orc_lp *lp, const orc_opts *o, orc_res *res, double *tab,
                    int32_t *nb) {
  return solve_impl_d(lp, o, res, tab, nb, NULL);
}
int orc_solve_tab_f(const orc_lp *lp, const orc_opts *o, orc_res *res, float *tab,
                    int32_t *nb) {
  return solve_impl_f(lp, o, res, tab, nb, NULL);
}
int orc_solve_warm_d(const orc_lp *lp, const orc_opts *o, const int32_t *warm, orc_res *res) {
  return solve_impl_d(lp, o, res, NULL, NULL, warm);
}
int orc_solve_warm_f(const orc_lp *lp, const orc_opts *o, const int32_t *warm, orc_res *res) {
  return solve_impl_f(lp, o, res, NULL, NULL, warm);
}

/* Batch driver: the CPU baseline of the benchmark.  Mirrors the reference's GLPK loop
 * (GPU_Tester.cu:151-173: one independent solve per LP, setup included) but spreads
 * the LPs over `threads` OpenMP threads. */
int orc_solve_batch_d(int64_t count, const int32_t *m, const int32_t *n, const int64_t *a_off,
                      const int64_t *b_off, const int64_t *c_off, const int64_t *s_off,
                      const double *pool, const int8_t *spool, const orc_opts *o,
                      int32_t *status, int32_t *pivots, double *obj, int threads) {
  int bad = 0;
#ifdef _OPENMP
  if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4) reduction(| : bad)
#endif
  for (int64_t k = 0; k < count; ++k) {
    orc_lp lp;
    lp.m = m[k];
    lp.n = n[k];
    lp.A = pool + a_off[k];
    lp.b = pool + b_off[k];
    lp.c = pool + c_off[k];
    lp.sense = (s_off && s_off[k] >= 0) ? spool + s_off[k] : NULL;
    orc_res r;
    memset(&r, 0, sizeof r);
    if (solve_impl_d(&lp, o, &r, NULL, NULL, NULL) != 0) { bad = 1; continue; }
    status[k] = r.status;
    pivots[k] = r.pivots;
    obj[k] = r.objective;
  }
  return bad ? -1 : 0;
}
