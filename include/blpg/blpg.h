/*
 * blpg.h -- C ABI of the MI355X batched simplex solver (libblpg.so).
 *
 * The reference exposes its batched solver as the C++ class `Simplex`
 * (BLPG-Solver/src/simplex.cuh:37-83).  This header is the C boundary under the
 * C++ drop-in (include/blpg/simplex.hpp): plain pointers and sizes, no C++ or
 * torch types.  Conventions:
 *   - every LP is   maximise c^T x  s.t.  A x (<=|=|>=) b,  x >= 0
 *     (the reference's form: rows GLP_UP, columns GLP_LO 0, GLP_MAX,
 *      glpk_lp_solver.cpp:102-145; simplex.cu:463-550, 880-890);
 *   - functions return BLPG_OK (0) or a negative BLPG_E_* code; the message is
 *     in blpg_last_error(ctx);
 *   - per-LP status uses GLPK's codes (GLP_OPT 5, GLP_NOFEAS 4, GLP_UNBND 6)
 *     plus BLPG_ITER_LIMIT / BLPG_NUMERIC;
 *   - inputs are borrowed for the duration of a call, outputs are caller-owned;
 *   - one context per host thread; the library never resets the device
 *     (the reference calls cudaDeviceReset per batch, simplex.cu:974).
 */
#ifndef BLPG_H
#define BLPG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---- */
#define BLPG_OK 0
#define BLPG_E_INVALID -1  /* bad argument (sizes, NaN/Inf input, NULL pointer) */
#define BLPG_E_HIP -2      /* a HIP runtime call failed */
#define BLPG_E_NOMEM -3    /* device or host allocation failed */
#define BLPG_E_NODEVICE -4 /* no HIP device / the library's kernels cannot run */
#define BLPG_E_BUDGET -5   /* one LP does not fit the HBM budget */

/* ---- per-LP status (GLPK codes, glpk_lp_solver.h:18-26, + ours) ---- */
#define BLPG_UNDECIDED 1   /* not solved (yet) */
#define BLPG_INFEASIBLE 4  /* GLP_NOFEAS */
#define BLPG_OPTIMAL 5     /* GLP_OPT */
#define BLPG_UNBOUNDED 6   /* GLP_UNBND */
#define BLPG_ITER_LIMIT 100
#define BLPG_NUMERIC 101   /* the FP64 check failed and no fallback could fix it */

/* ---- precision ---- */
#define BLPG_FP64 0 /* FP64 tableau (the reference's arithmetic, simplex.cu:35) */
#define BLPG_FP32 1 /* FP32 tableau + FP64 re-solve / KKT check + FP64 fallback */

/* ---- kernel selection (blpg_options.kernel) ---- */
#define BLPG_KERNEL_AUTO 0   /* autotuner picks per size bin */
#define BLPG_KERNEL_STREAM 1 /* tableau streamed from HBM (any size) */
#define BLPG_KERNEL_LDS 2    /* tableau resident in LDS (when it fits) */
#define BLPG_KERNEL_TILE 3   /* tableau resident in VGPRs (when it fits) */

typedef struct blpg_ctx blpg_ctx;
typedef struct blpg_batch blpg_batch;

/* One LP.  Replaces the pair (math::matrix<double> A, std::vector<double> b) given to
 * Simplex::setConstratint (simplex.cuh:75) plus one row of the objective matrix C given
 * to Simplex::ComputeLP (simplex.cuh:77). */
typedef struct {
  int32_t m, n;
  const double *A;     /* m*n, row-major */
  const double *b;     /* m */
  const double *c;     /* n, maximised */
  const int8_t *sense; /* m: 0 '<=', 1 '=', 2 '>='; NULL = all '<=' */
  int32_t family;      /* parametric family id for the learned warm start; -1 none */
  const int32_t *warm_basis; /* m variable ids to crash in before phase I (id < n structural,
                                n <= id < n+m slack of row id-n; -1 skip), or NULL (cold) */
} blpg_lp;

/* Per-LP result.  Replaces Simplex::getResultAll / getStatusAll (simplex.cu:424-461),
 * which return only the objective and a status the kernel never sets. */
typedef struct {
  int32_t status;        /* BLPG_OPTIMAL, BLPG_INFEASIBLE, ... */
  int32_t pivots;        /* total pivots (phase I + phase II) of the run that produced the result */
  int32_t phase1_pivots;
  int32_t fallback_used; /* 1 if the FP32 result failed the FP64 check and FP64 re-solved it */
  double objective;      /* c^T x (FP64 re-solve when verified, else the tableau value) */
  double kkt_primal;     /* FP64 primal residual of the final basis (-1 if not computed) */
  double kkt_dual;       /* FP64 dual residual (-1 if not computed) */
  double *x;             /* caller-allocated n entries, or NULL */
  int32_t *basis;        /* caller-allocated m entries (variable ids), or NULL */
} blpg_result;

typedef struct {
  int32_t precision;   /* BLPG_FP64 (default) or BLPG_FP32 */
  int32_t max_iter;    /* pivot cap per LP; 0 = 50*(m+n)+1000 */
  int32_t bland_after; /* consecutive degenerate pivots before Bland's rule; 0 = 50 */
  int32_t verify;      /* FP64 KKT re-solve of the final basis: 1 on (default), 0 off (FP64 only) */
  double eps_d;        /* entering tolerance; 0 = default (1e-9 FP64, 1e-6 FP32) */
  double eps_p;        /* pivot tolerance;    0 = default (1e-9 FP64, 1e-6 FP32) */
  double tol_feas;     /* phase-I tolerance;  0 = default (1e-7 FP64, 1e-4 FP32) */
  double kkt_tol;      /* relative KKT tolerance; 0 = 1e-9 */
  int64_t hbm_budget;  /* device bytes the library may hold for one batch; 0 = no limit */
  int32_t kernel;      /* BLPG_KERNEL_* */
  int32_t device;      /* HIP device ordinal */
  int32_t use_graph;   /* 1 = replay the batch through a captured hipGraph (default 1) */
  int32_t reserved;
} blpg_options;

void blpg_default_options(blpg_options *opt);

/* Context: device, stream, workspaces, tuning table.  No counterpart in the reference
 * (its state lives in the Simplex object, simplex.cuh:38-58). */
int blpg_create(blpg_ctx **ctx, const blpg_options *opt);
void blpg_destroy(blpg_ctx *ctx);
const char *blpg_last_error(const blpg_ctx *ctx);
int blpg_set_options(blpg_ctx *ctx, const blpg_options *opt);

/* Blocking batch solve of `count` independent LPs of any sizes.  Replaces
 * Simplex(N_S,..) + setConstratint(A,b) + ComputeLP(C, streams) + getResultAll()
 * (simplex.cu:387-978, called at GPU_Tester.cu:244-259). */
int blpg_solve_batch(blpg_ctx *ctx, const blpg_lp *lps, int64_t count, blpg_result *out);

/* The reference's batch shape: one A, b (and senses) shared by `count` objectives C
 * (count x n, row-major), GPU_Tester.cu:119-127 + simplex.cu:463-550, 880-890. */
int blpg_solve_shared_A(blpg_ctx *ctx, int32_t m, int32_t n, const double *A, const double *b,
                        const int8_t *sense, const double *C, int64_t count, blpg_result *out);

/* Device-resident batches (inputs uploaded once, solved many times; used by the
 * serving loop and the benchmark).  No reference counterpart: the reference copies
 * the whole tableau host->device on every ComputeLP (simplex.cu:944-953). */
int blpg_batch_upload(blpg_ctx *ctx, const blpg_lp *lps, int64_t count, blpg_batch **batch);
int blpg_batch_solve(blpg_ctx *ctx, blpg_batch *batch); /* enqueue on the context stream */
int blpg_batch_sync(blpg_ctx *ctx);
int blpg_batch_download(blpg_ctx *ctx, blpg_batch *batch, blpg_result *out);
void blpg_batch_free(blpg_ctx *ctx, blpg_batch *batch);

/* Timing and routing of the last blpg_batch_solve, for the benchmark:
 * kernel_ms[k] = HIP-event time of the solve kernels of bin k (on the stream they run on). */
typedef struct {
  int32_t nbins;
  int32_t kernel[16];   /* BLPG_KERNEL_* chosen per bin */
  int32_t rcap[16], ccap[16];
  int64_t count[16];
  double kernel_ms[16];
  double verify_ms;     /* longest FP64 check of one bin */
  double fallback_ms;   /* longest FP64 fallback (+ re-check) of one bin */
  int64_t fallback_count;
  int64_t total_pivots;
  double total_ms;      /* event time from first to last GPU operation of the solve */
} blpg_batch_info;
int blpg_batch_info_get(blpg_ctx *ctx, const blpg_batch *batch, blpg_batch_info *info);

/* Seeded synthetic LP generators (counter-based; bit-identical on every host).
 * config 1: A ~ U[0,1) 30% zeroed, b ~ U[1,10), c ~ U[0,1);
 * config 2: A ~ U[-1,1) 40% zeroed, b ~ U[0,10) with floor(0.1m) zero rows, c ~ N(0,1),
 *           floor(0.05 batch) LPs with one unbounded column;
 * config 3: A ~ U[-1,1) 50% zeroed, b ~ N(0,5), c ~ N(0,1), round(0.3m) '=' rows,
 *           floor(0.03 batch) LPs with a contradictory row pair;
 * config 4: parametric family `shape_idx` (A_f fixed; b, c scaled by 1 + 0.05 N(0,1));
 * config 5: 27x32 as config 1, other shapes as config 2 without planted columns.
 * Writes `count` LPs [first, first+count) of a
 * batch of `batch_count` LPs with shape (m, n): A (count*m*n), b (count*m),
 * c (count*n), sense (count*m, may be NULL for configs without senses),
 * flags (count; bit0 planted unbounded, bit1 planted infeasible). */
int blpg_gen_config(int32_t config, uint64_t seed, int32_t shape_idx, int64_t first,
                    int64_t count, int64_t batch_count, int32_t m, int32_t n, double *A,
                    double *b, double *c, int8_t *sense, int32_t *flags);

/* Library version / build string. */
const char *blpg_version(void);

#ifdef __cplusplus
}
#endif
#endif /* BLPG_H */
