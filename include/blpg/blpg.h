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
#define BLPG_KERNEL_TILE 3   /* tableau resident in VGPRs, 2-D blocks per thread (when it fits) */
#define BLPG_KERNEL_WAVE 4   /* tableau resident in VGPRs, whole columns per lane (when it fits) */

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
  double kkt_primal;     /* FP64 primal residual of the final basis (-1 if not computed);
                            FP32 INFEASIBLE: the Farkas certificate's relative violation */
  double kkt_dual;       /* FP64 dual residual (-1 if not computed); FP32 INFEASIBLE: the
                            certificate's margin -y^T b' / ((1 + |b|_inf) |y|_inf) */
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
  int32_t sibling_warm; /* blpg_solve_shared_A: 1 = crash LPs 1.. from LP 0's final basis
                           when LP 0 ends OPTIMAL or UNBOUNDED (default 0: every LP cold) */
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
 * (simplex.cu:387-978, called at GPU_Tester.cu:244-259).  A batch larger than the HBM
 * budget (or the device's free memory) is solved in consecutive parts that fit, with no
 * device reset (the reference's multi-batch mode, GPU_Tester.cu:194-239, resets the
 * device between parts). */
int blpg_solve_batch(blpg_ctx *ctx, const blpg_lp *lps, int64_t count, blpg_result *out);

/* The reference's batch shape: one A, b (and senses) shared by `count` objectives C
 * (count x n, row-major), GPU_Tester.cu:119-127 + simplex.cu:463-550, 880-890.
 * With opt.sibling_warm, LP 0 is solved first; when it ends OPTIMAL or UNBOUNDED its final
 * basis is primal feasible for every LP of the batch (same A, b, senses), so LPs 1.. are
 * crashed into it (warm_basis) and skip phase I: their results equal a solve with that
 * warm list (status, objective; pivots count the crash and the rest from there). */
int blpg_solve_shared_A(blpg_ctx *ctx, int32_t m, int32_t n, const double *A, const double *b,
                        const int8_t *sense, const double *C, int64_t count, blpg_result *out);

/* Shared-A batch solve that allocates nothing, for callers that cannot allocate (code
 * under hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same LPs and results as blpg_solve_shared_A with
 * precision BLPG_FP64 and verify = 0 (status, objective = the tableau's, x, basis, pivots),
 * on shapes the register-tile kernels hold.  Every pointer is a device pointer:
 * A (m*n, row-major), b (m), sense (m, or NULL for all '<='), C (count*n); outputs status,
 * objective (count), x (count*n), basis (count*m), pivots (count, or NULL).  The call
 * enqueues on `stream` (a hipStream_t; NULL = the null stream) and returns; it does not
 * synchronise.  Its descriptors, queue counter, staged A | b | C and internal outputs live
 * in a static device array of 32 MiB per GPU (held until the process ends):
 * blpg_noalloc_max_count(m, n) LPs at most, larger counts fail with BLPG_E_BUDGET.  Every
 * call on a GPU shares that array, so calls that overlap on one GPU race: serialise them
 * on one stream.  Errors: blpg_last_error(NULL), per host thread. */
int blpg_solve_shared_A_noalloc(int32_t m, int32_t n, const double *A, const double *b,
                                const int8_t *sense, const double *C, int64_t count,
                                int32_t *status, double *objective, double *x, int32_t *basis,
                                int32_t *pivots, void *stream);
int64_t blpg_noalloc_max_count(int32_t m, int32_t n);

/* Device-resident batches (inputs uploaded once, solved many times; used by the
 * serving loop and the benchmark).  No reference counterpart: the reference copies
 * the whole tableau host->device on every ComputeLP (simplex.cu:944-953). */
int blpg_batch_upload(blpg_ctx *ctx, const blpg_lp *lps, int64_t count, blpg_batch **batch);
int blpg_batch_solve(blpg_ctx *ctx, blpg_batch *batch); /* enqueue on the context stream */
int blpg_batch_sync(blpg_ctx *ctx);
int blpg_batch_download(blpg_ctx *ctx, blpg_batch *batch, blpg_result *out);
void blpg_batch_free(blpg_ctx *ctx, blpg_batch *batch);
/* New LPs of the same shapes (and senses / warm lists present where they were) into an
 * uploaded batch: host work only, into the batch's pinned staging, like the reference's
 * setConstratint building its pinned tableau (simplex.cu:463-550).  The next
 * blpg_batch_solve starts with the host-to-device copy of the inputs, as ComputeLP does
 * (simplex.cu:944-953): each size bin's inputs on its own stream just before its kernel
 * (largest bin first, one copy after the other), so later bins' copies overlap earlier
 * bins' solves; that solve is enqueued eagerly, not from the captured graph.  Two LPs may share a new array only
 * where the uploaded batch shared one (the upload stores each distinct array once). */
int blpg_batch_repack(blpg_ctx *ctx, blpg_batch *batch, const blpg_lp *lps, int64_t count);

/* Bulk results of the last blpg_batch_solve without per-LP copies: waits for the solve,
 * copies the batch's results with one DMA into host memory the batch owns (pinned) and
 * points `res` at it.  The arrays stay valid until the next solve, fetch or free of the
 * batch.  LP k's x is x[x_offset[k] .. + n_k), its basis basis[basis_offset[k] .. + m_k). */
typedef struct {
  int64_t count;
  const int32_t *status, *pivots, *phase1_pivots, *fallback_used;
  const double *objective, *kkt_primal, *kkt_dual;
  const double *x;
  const int32_t *basis;
  const int64_t *x_offset, *basis_offset;
} blpg_batch_results;
int blpg_batch_fetch(blpg_ctx *ctx, blpg_batch *batch, blpg_batch_results *res);

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
  double warmstart_ms;  /* learned warm start inference (features, network, basis pick) */
  int32_t param[16];    /* per bin: tile variant (TILE) or threads per workgroup */
} blpg_batch_info;
int blpg_batch_info_get(blpg_ctx *ctx, const blpg_batch *batch, blpg_batch_info *info);

/* ---- measured autotuning (no reference counterpart: the reference fixes one block size
 * per launch, simplex.cu:558) ----
 * blpg_autotune times, for every LP shape in `lps`, each kernel instance that can hold it
 * (register-tile variants, LDS-resident and HBM-streamed tableaux at 64..512 threads) on
 * those LPs in the context's precision (best of `reps` solves after one warm-up), and
 * routes later batches of that shape and precision to the fastest; the default route's
 * tile variant is kept unless another instance is more than 3% faster (closer timings are
 * run-to-run noise, and a route that flips moves shapes between bins).  The choice changes
 * speed only: all instances compute bit-identical results.  With the FP64 check on, it also
 * times the shape's check with and without its first LDS tier (31 KB per LP) where a later
 * tier exists, and skips the first tier when that is more than 3% faster (workloads whose
 * final bases never fit it); decisions are the same either way.  blpg_tune_set pins a route
 * by hand (param: tile variant, or threads per workgroup); export / import carry the table as
 * text lines "precision m n kernel param ms" (kernel 0: the check plan, param 1 = skip). */
int blpg_autotune(blpg_ctx *ctx, const blpg_lp *lps, int64_t count, int32_t reps);
int blpg_tune_set(blpg_ctx *ctx, int32_t precision, int32_t m, int32_t n, int32_t kernel,
                  int32_t param);
int blpg_tune_export(blpg_ctx *ctx, char *buf, int64_t cap, int64_t *len);
int blpg_tune_import(blpg_ctx *ctx, const char *text);

/* ---- learned warm start (no reference counterpart) ----
 * A per-family network maps the normalised (b, c) of an LP of a parametric family (fixed
 * A and shape, varying b and c) to one logit per variable (structurals 0..n-1, then the
 * slacks n..n+m-1): x -> hidden (ReLU) -> hidden (ReLU) -> n+m, bf16 weights and
 * activations on MFMA with FP32 accumulation and FP32 master weights (Adam).  It is
 * trained on the GPU from optimal bases the solver itself produced.  At solve time the m
 * variables with the largest logits form the crash basis (most confident first); the crash
 * and its repair (dual simplex when the crashed basis is dual feasible, else the auxiliary
 * phase I) run inside the solve kernels, so a wrong prediction costs pivots, never the
 * answer. */
typedef struct blpg_ws_model blpg_ws_model;
typedef struct {
  int32_t hidden;      /* hidden width, a multiple of 64; 0 = 256 */
  int32_t steps;       /* full-batch training steps; 0 = 300 */
  float lr;            /* Adam learning rate; 0 = 1e-3 */
  float weight_decay;  /* decoupled weight decay; default 0 */
  uint64_t seed;       /* weight initialisation */
} blpg_ws_config;

int blpg_ws_create(blpg_ctx *ctx, int32_t families, int32_t m, int32_t n,
                   const blpg_ws_config *cfg, blpg_ws_model **model);
void blpg_ws_destroy(blpg_ctx *ctx, blpg_ws_model *model);
/* Trains on `count` LPs (lps[k].family in [0, families), shape m x n) whose optimal bases
 * (m variable ids each) are `bases`; fits the per-family input normalisation first.
 * loss (optional, cfg.steps entries): mean over families of the cross-entropy per step. */
int blpg_ws_train(blpg_ctx *ctx, blpg_ws_model *model, const blpg_lp *lps,
                  const int32_t *bases, int64_t count, float *loss);
/* Logits (count x (n+m), optional) and warm-start lists (count x m, optional). */
int blpg_ws_predict(blpg_ctx *ctx, blpg_ws_model *model, const blpg_lp *lps, int64_t count,
                    float *logits, int32_t *warm);
/* Attach (or detach with NULL) a model to a device batch: every later blpg_batch_solve
 * evaluates it on the batch's LPs with family >= 0 and the model's shape and crashes
 * from its prediction. */
int blpg_batch_set_warmstart(blpg_ctx *ctx, blpg_batch *batch, blpg_ws_model *model);
/* FP32 master parameters of one family, unpadded: W1 (hidden x (m+n)), b1, W2 (hidden x
 * hidden), b2, W3 ((n+m) x hidden), b3.  `get` / `set` may be NULL; *count receives the
 * number of floats. */
int blpg_ws_params(blpg_ctx *ctx, blpg_ws_model *model, int32_t family, float *get,
                   const float *set, int64_t *count);
/* Per-family input normalisation (mu, 1/sd over the m+n inputs b then c). */
int blpg_ws_norm(blpg_ctx *ctx, const blpg_ws_model *model, int32_t family, float *mu,
                 float *inv_sd);
/* One forward + backward pass without an update (gradient check): fits the normalisation
 * to `lps` like blpg_ws_train, then writes grad (families x count floats, the
 * blpg_ws_params layout) and loss (families entries). */
int blpg_ws_gradient(blpg_ctx *ctx, blpg_ws_model *model, const blpg_lp *lps,
                     const int32_t *bases, int64_t count, float *grad, float *loss);

/* ---- serving (no reference counterpart: the reference solves one batch per call) ----
 * A server admits single LPs, keeps them resident in device arenas (one per LP shape,
 * fixed slots, a device-side LP count), forms batches from its queue -- a batch launches
 * when it is full or when its oldest request has waited max_wait_us -- and replays each
 * arena's captured hipGraph for any batch size.  Requests are admitted against an HBM
 * budget: a submit that would exceed it returns BLPG_E_BUDGET (retry after polling).
 * A worker thread owned by the server does the batching; submit / poll are thread-safe. */
typedef struct blpg_server blpg_server;
typedef struct {
  int64_t hbm_budget;   /* bytes of device memory for arenas and resident requests; 0 = 64 GiB */
  int32_t max_batch;    /* LPs per launch (arena capacity per shape); 0 = 2048 */
  int32_t max_wait_us;  /* batching delay bound; 0 = 200 */
  int32_t precision;    /* BLPG_FP64 or BLPG_FP32 (FP64 check + fallback) */
  int32_t arenas;       /* arenas per shape (batches of one shape in flight); 0 = 2 */
} blpg_server_config;
typedef struct {
  int64_t ticket;       /* as returned by blpg_server_submit */
  int32_t status, pivots;
  double objective;
  double latency_us;    /* submit (or scheduled arrival, blpg_server_replay) to completion */
} blpg_server_result;
typedef struct {
  int64_t submitted, completed, rejected, batches;
  double mean_batch;            /* LPs per launch */
  double p50_us, p99_us, max_us; /* completion latency over all completed requests */
  int64_t resident_bytes;       /* admitted, not yet polled */
} blpg_server_stats;

int blpg_server_create(blpg_ctx *ctx, const blpg_server_config *cfg, blpg_server **server);
/* Waits for the arenas' work (at most 60 s in all; an arena still running then is not
 * freed) and releases the server. */
void blpg_server_destroy(blpg_server *server);
/* Copies the LP (shape m x n, '<=' rows only or with senses).  x / basis of the result are
 * written to the caller's buffers (n doubles, m int32; may be NULL) before the ticket is
 * returned by poll; the buffers must stay valid until then. */
int blpg_server_submit(blpg_server *server, const blpg_lp *lp, double *x, int32_t *basis,
                       int64_t *ticket);
/* Up to `max` completed results; waits up to wait_us for the first one. */
int blpg_server_poll(blpg_server *server, blpg_server_result *out, int32_t max, int32_t wait_us,
                     int32_t *n);
int blpg_server_stats_get(blpg_server *server, blpg_server_stats *stats);
/* Learned warm start for family-tagged requests: with `model` attached, a request of the
 * model's shape whose lp->family is one of the model's families is batched with other
 * tagged requests of that shape (any mix of families per launch); the model runs inside
 * those batches' graphs and each LP is crashed from its family's prediction (repaired on
 * the device like blpg_batch_set_warmstart).  Other requests are unaffected.  Call before
 * the first tagged request (afterwards: BLPG_E_INVALID); NULL detaches.  The model must
 * stay alive, and untrained, until the server is destroyed. */
int blpg_server_set_warmstart(blpg_server *server, blpg_ws_model *model);
/* Load generator: submits lps[k] at arrival_us[k] (microseconds after the call, ascending)
 * from up to 8 client threads (arrival k on thread k mod P), retrying admission, and waits
 * until all completed; latencies count from the scheduled arrival.  Results per LP go to
 * out[k] (optional).  With no result for 60 s it stops and returns BLPG_E_HIP, the server's
 * state in blpg_last_error(NULL) (per calling thread).  A request whose launch fails (or
 * whose stream reports an error) completes as BLPG_UNDECIDED. */
int blpg_server_replay(blpg_server *server, const blpg_lp *lps, const double *arrival_us,
                       int64_t count, blpg_server_result *out);

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
