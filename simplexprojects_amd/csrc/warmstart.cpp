// warmstart.cpp -- the learned warm start: per-family network, training on the GPU, and
// inference into a device batch's crash lists.  No reference counterpart (the reference
// always starts from the slack basis, simplex.cu:506-534).
//
// Parameter layout per family (FP32 master copy, bf16 mirror, gradients and Adam moments
// share it), with d_in = m+n inputs padded to d_in_pad and d_out = n+m outputs padded to
// d_out_pad (multiples of 64; padding entries stay 0):
//   W1 [H][d_in_pad] | b1 [H] | W2 [H][H] | b2 [H] | W3 [d_out_pad][H] | b3 [d_out_pad]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "kernels.hpp"
#include "mlp.hpp"

using namespace blpg;

struct blpg_ws_model {
  int nfam = 0, m = 0, n = 0, H = 256;
  int din = 0, din_pad = 0, dout = 0, dout_pad = 0;
  int64_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0, P = 0;
  void* mem = nullptr;
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  uint16_t* pb = nullptr;
  float *mu = nullptr, *inv_sd = nullptr;
  int64_t step = 0;
  blpg_ws_config cfg{};
};

namespace {

constexpr int64_t kAl = 64;
int64_t up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }
size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

uint64_t mix(uint64_t z) {  // splitmix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uni(uint64_t seed, uint64_t a, uint64_t b) {  // U[-1, 1)
  const uint64_t r = mix(seed ^ mix(a * 0x100000001B3ull + b));
  return (double)(r >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Extra buffers of a training pass.
struct TrainBufs {
  uint8_t* Y = nullptr;
  float *dZ = nullptr, *dA2 = nullptr, *dA1 = nullptr;
  uint16_t *dZb = nullptr, *dA2b = nullptr, *dA1b = nullptr;
  float* loss = nullptr;  // steps x nfam
};

int alloc_rows(blpg_ctx* ctx, const blpg_ws_model* M, WsRows& R, TrainBufs* tb, int64_t raw_rows,
               int steps);

// Groups `count` rows by family (fam[k] < 0: skipped) into 64-row-padded segments and
// allocates the device buffers of a forward pass (+ training buffers when tb != nullptr).
int plan_rows(blpg_ctx* ctx, const blpg_ws_model* M, const std::vector<int32_t>& fam, WsRows& R,
              TrainBufs* tb, int64_t raw_rows, int steps, std::vector<int32_t>& row_lp,
              std::vector<int32_t>& row_fam) {
  R.nfam = M->nfam;
  R.fam_count.assign(M->nfam, 0);
  for (int32_t f : fam)
    if (f >= 0) R.fam_count[f]++;
  R.seg_off.assign(M->nfam, 0);
  R.seg_rows.assign(M->nfam, 0);
  int64_t rows = 0;
  R.max_seg = 0;
  for (int f = 0; f < M->nfam; ++f) {
    R.seg_off[f] = (int32_t)rows;
    R.seg_rows[f] = (int32_t)up(R.fam_count[f], kAl);
    rows += R.seg_rows[f];
    R.max_seg = std::max(R.max_seg, R.seg_rows[f]);
  }
  if (rows > INT32_MAX / 4) return fail(ctx, BLPG_E_INVALID, "too many warm-start rows");
  R.rows = rows;
  row_lp.assign(rows, -1);
  row_fam.assign(rows, 0);
  std::vector<int32_t> fill(M->nfam, 0);
  for (int64_t k = 0; k < (int64_t)fam.size(); ++k) {
    const int f = fam[k];
    if (f < 0) continue;
    const int64_t r = R.seg_off[f] + fill[f]++;
    row_lp[r] = (int32_t)k;
  }
  for (int f = 0; f < M->nfam; ++f)
    for (int r = 0; r < R.seg_rows[f]; ++r) row_fam[R.seg_off[f] + r] = f;
  int rc = alloc_rows(ctx, M, R, tb, raw_rows, steps);
  if (rc) return rc;
  std::vector<int32_t> so(3 * M->nfam);
  for (int f = 0; f < M->nfam; ++f) {
    so[f] = R.seg_off[f];
    so[M->nfam + f] = R.seg_rows[f];
    so[2 * M->nfam + f] = R.fam_count[f];
  }
  HIPCHK(ctx, hipMemcpy(R.row_lp, row_lp.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(R.row_fam, row_fam.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(R.d_seg_off, so.data(), sizeof(int32_t) * so.size(), hipMemcpyHostToDevice));
  return BLPG_OK;
}

// Device buffers of a forward pass over R.rows rows (+ training buffers when tb != nullptr).
// The row tables come first, in one range of R.table_bytes: row_lp [rows], row_fam [rows],
// then seg_off, seg_rows, fam_count [nfam each], every array 256-byte aligned.
int alloc_rows(blpg_ctx* ctx, const blpg_ws_model* M, WsRows& R, TrainBufs* tb, int64_t raw_rows,
               int steps) {
  const int64_t rows = R.rows;
  size_t off = 0;
  const size_t o_rl = off; off += al256(sizeof(int32_t) * rows);
  const size_t o_rf = off; off += al256(sizeof(int32_t) * rows);
  const size_t o_so = off; off += al256(sizeof(int32_t) * M->nfam * 3);
  R.table_bytes = off;
  const size_t o_X = off; off += al256(2 * (size_t)rows * M->din_pad);
  const size_t o_H1 = off; off += al256(2 * (size_t)rows * M->H);
  const size_t o_H2 = off; off += al256(2 * (size_t)rows * M->H);
  const size_t o_Z = off; off += al256(4 * (size_t)rows * M->dout_pad);
  size_t o_Y = 0, o_dZ = 0, o_dZb = 0, o_dA2 = 0, o_dA2b = 0, o_dA1 = 0, o_dA1b = 0, o_raw = 0,
         o_loss = 0;
  if (tb) {
    o_Y = off; off += al256((size_t)rows * M->dout_pad);
    o_dZ = off; off += al256(4 * (size_t)rows * M->dout_pad);
    o_dZb = off; off += al256(2 * (size_t)rows * M->dout_pad);
    o_dA2 = off; off += al256(4 * (size_t)rows * M->H);
    o_dA2b = off; off += al256(2 * (size_t)rows * M->H);
    o_dA1 = off; off += al256(4 * (size_t)rows * M->H);
    o_dA1b = off; off += al256(2 * (size_t)rows * M->H);
    o_raw = off; off += al256(4 * (size_t)std::max<int64_t>(raw_rows, 1) * M->din);
    o_loss = off; off += al256(4 * (size_t)std::max(steps, 1) * M->nfam);
  } else if (raw_rows > 0) {
    o_raw = off; off += al256(4 * (size_t)raw_rows * M->din);
  }
  if (hipMalloc(&R.mem, off) != hipSuccess)
    return fail(ctx, BLPG_E_NOMEM, "hipMalloc(%zu) for warm-start rows failed", off);
  R.bytes = off;
  char* b = (char*)R.mem;
  R.row_lp = (int32_t*)(b + o_rl);
  R.row_fam = (int32_t*)(b + o_rf);
  R.d_seg_off = (int32_t*)(b + o_so);
  R.d_seg_rows = R.d_seg_off + M->nfam;
  R.d_fam_count = R.d_seg_off + 2 * M->nfam;
  R.X = (uint16_t*)(b + o_X);
  R.H1 = (uint16_t*)(b + o_H1);
  R.H2 = (uint16_t*)(b + o_H2);
  R.Z = (float*)(b + o_Z);
  if (raw_rows > 0) R.raw = (float*)(b + o_raw);
  if (tb) {
    tb->Y = (uint8_t*)(b + o_Y);
    tb->dZ = (float*)(b + o_dZ);
    tb->dZb = (uint16_t*)(b + o_dZb);
    tb->dA2 = (float*)(b + o_dA2);
    tb->dA2b = (uint16_t*)(b + o_dA2b);
    tb->dA1 = (float*)(b + o_dA1);
    tb->dA1b = (uint16_t*)(b + o_dA1b);
    tb->loss = (float*)(b + o_loss);
  }
  return BLPG_OK;
}

void free_rows(WsRows& R) {
  if (R.mem) (void)hipFree(R.mem);
  R = WsRows{};
}

// Z = W3 relu(W2 relu(W1 x + b1) + b2) + b3 for every row (X already built).
int forward(blpg_ctx* ctx, const blpg_ws_model* M, const WsRows& R, hipStream_t st) {
  if (R.rows == 0) return BLPG_OK;
  const int Mrows = R.max_seg;
  GemmArgs g;
  g.seg_off = R.d_seg_off; g.seg_rows = R.d_seg_rows; g.seg_dim = 1;
  // layer 1
  g.M = Mrows; g.N = M->H; g.K = M->din_pad;
  g.A = R.X; g.lda_i = M->din_pad; g.lda_k = 1; g.sa = 0;
  g.B = M->pb + M->oW1; g.ldb_k = 1; g.ldb_j = M->din_pad; g.sb = M->P;
  g.C = nullptr; g.Cb = R.H1; g.ldc = M->H; g.sc = 0;
  g.bias = M->p + M->ob1; g.sbias = M->P; g.relu = 1;
  HIPCHK(ctx, launch_gemm(g, M->nfam, st));
  // layer 2
  g.N = M->H; g.K = M->H;
  g.A = R.H1; g.lda_i = M->H;
  g.B = M->pb + M->oW2; g.ldb_j = M->H;
  g.Cb = R.H2; g.ldc = M->H;
  g.bias = M->p + M->ob2;
  HIPCHK(ctx, launch_gemm(g, M->nfam, st));
  // output layer
  g.N = M->dout_pad; g.K = M->H;
  g.A = R.H2; g.lda_i = M->H;
  g.B = M->pb + M->oW3; g.ldb_j = M->H;
  g.C = R.Z; g.Cb = nullptr; g.ldc = M->dout_pad;
  g.bias = M->p + M->ob3; g.relu = 0;
  HIPCHK(ctx, launch_gemm(g, M->nfam, st));
  return BLPG_OK;
}

// Gradients of the per-family mean cross-entropy into M->g (every entry written).
int backward(blpg_ctx* ctx, blpg_ws_model* M, const WsRows& R, const TrainBufs& T, float* loss,
             hipStream_t st) {
  BceArgs b;
  b.Z = R.Z; b.Y = T.Y; b.row_fam = R.row_fam; b.row_lp = R.row_lp; b.fam_count = R.d_fam_count;
  b.rows = R.rows; b.d_out = M->dout; b.d_out_pad = M->dout_pad;
  b.dZ = T.dZ; b.dZb = T.dZb; b.loss = loss;
  HIPCHK(ctx, launch_bce_grad(b, st));
  const int H = M->H;
  GemmArgs g;
  g.seg_off = R.d_seg_off; g.seg_rows = R.d_seg_rows;
  // dW3 = dZ^T H2  [d_out_pad x H], K over the family's rows
  g.seg_dim = 2; g.M = M->dout_pad; g.N = H; g.K = R.max_seg;
  g.A = T.dZb; g.lda_i = 1; g.lda_k = M->dout_pad; g.sa = 0;
  g.B = R.H2; g.ldb_k = H; g.ldb_j = 1; g.sb = 0;
  g.C = M->g + M->oW3; g.ldc = H; g.sc = M->P;
  HIPCHK(ctx, launch_gemm(g, M->nfam, st));
  HIPCHK(ctx, launch_colsum(T.dZ, M->dout_pad, R.d_seg_off, R.d_seg_rows, M->nfam, M->dout_pad,
                            M->g + M->ob3, M->P, st));
  // dA2 = (dZ W3) * [H2 > 0]
  GemmArgs h;
  h.seg_off = R.d_seg_off; h.seg_rows = R.d_seg_rows; h.seg_dim = 1;
  h.M = R.max_seg; h.N = H; h.K = M->dout_pad;
  h.A = T.dZb; h.lda_i = M->dout_pad; h.lda_k = 1;
  h.B = M->pb + M->oW3; h.ldb_k = H; h.ldb_j = 1; h.sb = M->P;
  h.C = T.dA2; h.Cb = T.dA2b; h.ldc = H;
  h.gate = R.H2; h.ldg = H;
  HIPCHK(ctx, launch_gemm(h, M->nfam, st));
  // dW2 = dA2^T H1, db2
  g.M = H; g.N = H;
  g.A = T.dA2b; g.lda_i = 1; g.lda_k = H;
  g.B = R.H1; g.ldb_k = H; g.ldb_j = 1;
  g.C = M->g + M->oW2; g.ldc = H;
  HIPCHK(ctx, launch_gemm(g, M->nfam, st));
  HIPCHK(ctx, launch_colsum(T.dA2, H, R.d_seg_off, R.d_seg_rows, M->nfam, H, M->g + M->ob2, M->P, st));
  // dA1 = (dA2 W2) * [H1 > 0]
  h.K = H;
  h.A = T.dA2b; h.lda_i = H;
  h.B = M->pb + M->oW2;
  h.C = T.dA1; h.Cb = T.dA1b;
  h.gate = R.H1;
  HIPCHK(ctx, launch_gemm(h, M->nfam, st));
  // dW1 = dA1^T X, db1
  g.M = H; g.N = M->din_pad;
  g.A = T.dA1b; g.lda_i = 1; g.lda_k = H;
  g.B = R.X; g.ldb_k = M->din_pad; g.ldb_j = 1;
  g.C = M->g + M->oW1; g.ldc = M->din_pad;
  HIPCHK(ctx, launch_gemm(g, M->nfam, st));
  HIPCHK(ctx, launch_colsum(T.dA1, H, R.d_seg_off, R.d_seg_rows, M->nfam, H, M->g + M->ob1, M->P, st));
  return BLPG_OK;
}

int check_lps(blpg_ctx* ctx, const blpg_ws_model* M, const blpg_lp* lps, int64_t count,
              bool need_family) {
  for (int64_t k = 0; k < count; ++k) {
    const blpg_lp& L = lps[k];
    if (L.m != M->m || L.n != M->n || !L.b || !L.c)
      return fail(ctx, BLPG_E_INVALID, "LP %lld: shape %dx%d, model is %dx%d", (long long)k, L.m,
                  L.n, M->m, M->n);
    if (need_family && (L.family < 0 || L.family >= M->nfam))
      return fail(ctx, BLPG_E_INVALID, "LP %lld: family %d outside [0, %d)", (long long)k,
                  L.family, M->nfam);
    for (int i = 0; i < L.m; ++i)
      if (!std::isfinite(L.b[i])) return fail(ctx, BLPG_E_INVALID, "LP %lld: b has NaN/Inf", (long long)k);
    for (int j = 0; j < L.n; ++j)
      if (!std::isfinite(L.c[j])) return fail(ctx, BLPG_E_INVALID, "LP %lld: c has NaN/Inf", (long long)k);
  }
  return BLPG_OK;
}

std::vector<float> raw_features(const blpg_ws_model* M, const blpg_lp* lps, int64_t count) {
  std::vector<float> raw((size_t)count * M->din);
  for (int64_t k = 0; k < count; ++k) {
    float* r = raw.data() + (size_t)k * M->din;
    for (int i = 0; i < M->m; ++i) r[i] = (float)lps[k].b[i];
    for (int j = 0; j < M->n; ++j) r[M->m + j] = (float)lps[k].c[j];
  }
  return raw;
}

// per-family mean and 1/sd of the raw features (population statistics, FP64 on the host)
int fit_norm(blpg_ctx* ctx, blpg_ws_model* M, const std::vector<float>& raw,
             const blpg_lp* lps, int64_t count) {
  const int d = M->din;
  std::vector<double> s1((size_t)M->nfam * d, 0), s2((size_t)M->nfam * d, 0);
  std::vector<int64_t> cnt(M->nfam, 0);
  for (int64_t k = 0; k < count; ++k) {
    const int f = lps[k].family;
    cnt[f]++;
    for (int j = 0; j < d; ++j) {
      const double v = raw[(size_t)k * d + j];
      s1[(size_t)f * d + j] += v;
      s2[(size_t)f * d + j] += v * v;
    }
  }
  std::vector<float> mu((size_t)M->nfam * d, 0.f), isd((size_t)M->nfam * d, 0.f);
  for (int f = 0; f < M->nfam; ++f) {
    if (!cnt[f]) continue;
    for (int j = 0; j < d; ++j) {
      const double mean = s1[(size_t)f * d + j] / (double)cnt[f];
      const double var = std::max(0.0, s2[(size_t)f * d + j] / (double)cnt[f] - mean * mean);
      const double sd = std::sqrt(var);
      mu[(size_t)f * d + j] = (float)mean;
      isd[(size_t)f * d + j] = sd > 1e-12 * (1.0 + std::fabs(mean)) ? (float)(1.0 / sd) : 0.f;
    }
  }
  HIPCHK(ctx, hipMemcpy(M->mu, mu.data(), sizeof(float) * mu.size(), hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(M->inv_sd, isd.data(), sizeof(float) * isd.size(), hipMemcpyHostToDevice));
  return BLPG_OK;
}

// Shared set-up of a training / gradient pass: rows, raw features, normalisation, labels.
int setup_pass(blpg_ctx* ctx, blpg_ws_model* M, const blpg_lp* lps, const int32_t* bases,
               int64_t count, int steps, WsRows& R, TrainBufs& T) {
  int rc = check_lps(ctx, M, lps, count, true);
  if (rc) return rc;
  for (int64_t k = 0; k < count; ++k)
    for (int i = 0; i < M->m; ++i) {
      const int v = bases[k * M->m + i];
      if (v < 0 || v >= M->dout)
        return fail(ctx, BLPG_E_INVALID, "basis of LP %lld holds %d", (long long)k, v);
    }
  std::vector<int32_t> fam(count);
  for (int64_t k = 0; k < count; ++k) fam[k] = lps[k].family;
  std::vector<int32_t> row_lp, row_fam;
  rc = plan_rows(ctx, M, fam, R, &T, count, steps, row_lp, row_fam);
  if (rc) return rc;
  const std::vector<float> raw = raw_features(M, lps, count);
  rc = fit_norm(ctx, M, raw, lps, count);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpy(R.raw, raw.data(), sizeof(float) * raw.size(), hipMemcpyHostToDevice));
  std::vector<uint8_t> Y((size_t)R.rows * M->dout_pad, 0);
  for (int64_t r = 0; r < R.rows; ++r) {
    const int k = row_lp[r];
    if (k < 0) continue;
    for (int i = 0; i < M->m; ++i) Y[(size_t)r * M->dout_pad + bases[(int64_t)k * M->m + i]] = 1;
  }
  HIPCHK(ctx, hipMemcpy(T.Y, Y.data(), Y.size(), hipMemcpyHostToDevice));
  FeatureArgs fa;
  fa.raw = R.raw; fa.ld_raw = M->din; fa.row_lp = R.row_lp; fa.row_fam = R.row_fam;
  fa.rows = R.rows; fa.d_in = M->din; fa.d_in_pad = M->din_pad; fa.mu = M->mu;
  fa.inv_sd = M->inv_sd; fa.X = R.X;
  HIPCHK(ctx, launch_features(fa, ctx->stream));
  return BLPG_OK;
}

}  // namespace

namespace blpg {

// Evaluates the batch's attached model and writes the crash lists of its family LPs
// (on the context stream, before the solve kernels).
int ws_run_batch(blpg_ctx* ctx, blpg_batch* B, hipStream_t st) {
  blpg_ws_model* M = B->model;
  if (!M || B->ws.rows == 0) return BLPG_OK;
  FeatureArgs fa;
  fa.desc = B->d_desc; fa.pool = B->d_pool; fa.row_lp = B->ws.row_lp; fa.row_fam = B->ws.row_fam;
  fa.rows = B->ws.rows; fa.d_in = M->din; fa.d_in_pad = M->din_pad; fa.mu = M->mu;
  fa.inv_sd = M->inv_sd; fa.X = B->ws.X;
  HIPCHK(ctx, launch_features(fa, st));
  const int rc = forward(ctx, M, B->ws, st);
  if (rc) return rc;
  HIPCHK(ctx, launch_pick_basis(B->ws.Z, M->dout_pad, B->ws.row_lp, B->ws.rows, B->d_desc, 0, 0,
                                B->d_warm, M->dout, st));
  return BLPG_OK;
}

void ws_free_rows(WsRows& R) { free_rows(R); }

// Serving arenas.  A launch holds up to `cap` LPs of any mix of families, so the rows are
// laid out for the worst case (every family's segment padded to 64 rows: at most
// cap + 63 * min(families, cap) rows) and the forward GEMMs are sized for the largest
// possible segment (cap rows); tiles beyond a family's rows in this launch exit at once
// (k_gemm reads the segment sizes from the device).  The row tables are rewritten before
// every launch, so one captured graph serves every batch.
int ws_arena_attach(blpg_ctx* ctx, blpg_batch* B, blpg_ws_model* M, int cap) {
  if (!M || cap < 1) return fail(ctx, BLPG_E_INVALID, "invalid arguments to ws_arena_attach");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  free_rows(B->ws);
  B->model = nullptr;
  B->g_gen[0] = B->g_gen[1] = -1;
  WsRows& R = B->ws;
  R.nfam = M->nfam;
  R.rows = up(cap + (kAl - 1) * std::min<int64_t>(M->nfam, cap), kAl);
  R.max_seg = (int)up(cap, kAl);
  R.arena = true;
  int rc = alloc_rows(ctx, M, R, nullptr, 0, 0);
  if (rc) { free_rows(R); return rc; }
  // no LP until the first launch's tables arrive
  HIPCHK(ctx, launch_fill_i32(R.row_lp, -1, R.rows, ctx->stream));
  HIPCHK(ctx, launch_fill_i32(R.d_seg_off, 0, 3 * (int64_t)M->nfam, ctx->stream));
  if (!B->d_warm) {
    if (hipMalloc(&B->warm_mem, sizeof(int32_t) * std::max<int64_t>(B->basis_elems, 1)) != hipSuccess) {
      free_rows(R);
      return fail(ctx, BLPG_E_NOMEM, "hipMalloc for warm lists failed");
    }
    B->d_warm = (int32_t*)B->warm_mem;
    HIPCHK(ctx, launch_fill_i32(B->d_warm, -1, B->basis_elems, ctx->stream));
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  B->model = M;
  warm_route(B);
  return BLPG_OK;
}

void ws_arena_tables(const blpg_batch* B, const int32_t* fam, int k, void* image) {
  const WsRows& R = B->ws;
  const int nf = R.nfam;
  char* img = (char*)image;
  int32_t* row_lp = (int32_t*)img;
  int32_t* row_fam = (int32_t*)(img + al256(sizeof(int32_t) * R.rows));
  int32_t* so = (int32_t*)(img + 2 * al256(sizeof(int32_t) * R.rows));
  std::fill(so, so + 3 * nf, 0);
  int32_t* seg_off = so;
  int32_t* seg_rows = so + nf;
  int32_t* count = so + 2 * nf;
  for (int s = 0; s < k; ++s) count[fam[s]]++;
  int64_t rows = 0;
  for (int f = 0; f < nf; ++f) {
    seg_off[f] = (int32_t)rows;
    seg_rows[f] = (int32_t)up(count[f], kAl);
    rows += seg_rows[f];
  }
  std::fill(row_lp, row_lp + R.rows, -1);
  std::fill(row_fam, row_fam + R.rows, 0);
  for (int f = 0; f < nf; ++f)
    std::fill(row_fam + seg_off[f], row_fam + seg_off[f] + seg_rows[f], f);
  std::vector<int32_t> fill(nf, 0);
  for (int s = 0; s < k; ++s) row_lp[seg_off[fam[s]] + fill[fam[s]]++] = s;
}

void ws_model_shape(const blpg_ws_model* M, int* families, int* m, int* n) {
  *families = M->nfam;
  *m = M->m;
  *n = M->n;
}

}  // namespace blpg

extern "C" {

int blpg_ws_create(blpg_ctx* ctx, int32_t families, int32_t m, int32_t n, const blpg_ws_config* cfg,
                   blpg_ws_model** out) {
  if (!ctx || !out || families < 1 || m < 1 || n < 1 || m > 4096 || n > 4096)
    return fail(ctx, BLPG_E_INVALID, "invalid arguments to blpg_ws_create");
  *out = nullptr;
  blpg_ws_model* M = new blpg_ws_model();
  if (cfg) M->cfg = *cfg;
  M->H = M->cfg.hidden > 0 ? M->cfg.hidden : 256;
  if (M->H % 64) { delete M; return fail(ctx, BLPG_E_INVALID, "hidden width must be a multiple of 64"); }
  if (M->cfg.steps <= 0) M->cfg.steps = 300;
  if (!(M->cfg.lr > 0)) M->cfg.lr = 1e-3f;
  M->nfam = families; M->m = m; M->n = n;
  M->din = m + n; M->dout = n + m;
  M->din_pad = (int)up(M->din, 64); M->dout_pad = (int)up(M->dout, 64);
  const int64_t H = M->H;
  M->oW1 = 0; M->ob1 = H * M->din_pad; M->oW2 = M->ob1 + H; M->ob2 = M->oW2 + H * H;
  M->oW3 = M->ob2 + H; M->ob3 = M->oW3 + (int64_t)M->dout_pad * H; M->P = up(M->ob3 + M->dout_pad, 64);
  const size_t np = (size_t)M->P * families;
  const size_t bytes = al256(4 * np) * 4 + al256(2 * np) + al256(4 * (size_t)families * M->din) * 2;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (hipMalloc(&M->mem, bytes) != hipSuccess) {
    delete M;
    return fail(ctx, BLPG_E_NOMEM, "hipMalloc(%zu) for the warm-start model failed", bytes);
  }
  char* b = (char*)M->mem;
  M->p = (float*)b; b += al256(4 * np);
  M->g = (float*)b; b += al256(4 * np);
  M->m1 = (float*)b; b += al256(4 * np);
  M->m2 = (float*)b; b += al256(4 * np);
  M->pb = (uint16_t*)b; b += al256(2 * np);
  M->mu = (float*)b; b += al256(4 * (size_t)families * M->din);
  M->inv_sd = (float*)b;
  // He-uniform hidden layers, Glorot-uniform output layer, zero biases and padding
  std::vector<float> p(np, 0.f);
  const uint64_t seed = M->cfg.seed;
  for (int f = 0; f < families; ++f) {
    float* q = p.data() + (size_t)f * M->P;
    const double a1 = std::sqrt(6.0 / M->din), a2 = std::sqrt(6.0 / H),
                 a3 = std::sqrt(6.0 / (H + M->dout));
    for (int64_t h = 0; h < H; ++h)
      for (int j = 0; j < M->din; ++j) q[M->oW1 + h * M->din_pad + j] = (float)(a1 * uni(seed, f * 8 + 1, h * 65536 + j));
    for (int64_t h = 0; h < H; ++h)
      for (int64_t j = 0; j < H; ++j) q[M->oW2 + h * H + j] = (float)(a2 * uni(seed, f * 8 + 2, h * 65536 + j));
    for (int64_t o = 0; o < M->dout; ++o)
      for (int64_t j = 0; j < H; ++j) q[M->oW3 + o * H + j] = (float)(a3 * uni(seed, f * 8 + 3, o * 65536 + j));
  }
  if (hipMemcpy(M->p, p.data(), 4 * np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(M->m1, 0, 4 * np) != hipSuccess || hipMemset(M->m2, 0, 4 * np) != hipSuccess ||
      hipMemset(M->g, 0, 4 * np) != hipSuccess ||
      hipMemset(M->mu, 0, 4 * (size_t)families * M->din) != hipSuccess ||
      hipMemset(M->inv_sd, 0, 4 * (size_t)families * M->din) != hipSuccess ||
      launch_to_bf16(M->p, M->pb, (int64_t)np, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    blpg_ws_destroy(ctx, M);
    return fail(ctx, BLPG_E_HIP, "warm-start model initialisation failed");
  }
  *out = M;
  return BLPG_OK;
}

void blpg_ws_destroy(blpg_ctx* ctx, blpg_ws_model* M) {
  if (!M) return;
  if (ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  if (M->mem) (void)hipFree(M->mem);
  delete M;
}

int blpg_ws_train(blpg_ctx* ctx, blpg_ws_model* M, const blpg_lp* lps, const int32_t* bases,
                  int64_t count, float* loss) {
  if (!ctx || !M || (count > 0 && (!lps || !bases)) || count < 1)
    return fail(ctx, BLPG_E_INVALID, "invalid arguments to blpg_ws_train");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int steps = M->cfg.steps;
  WsRows R;
  TrainBufs T;
  int rc = setup_pass(ctx, M, lps, bases, count, steps, R, T);
  if (rc) { free_rows(R); return rc; }
  const int64_t np = M->P * M->nfam;
  if (launch_fill_i32((int32_t*)T.loss, 0, (int64_t)steps * M->nfam, st) != hipSuccess) {
    free_rows(R);
    return fail(ctx, BLPG_E_HIP, "loss buffer clear failed");
  }
  for (int s = 0; s < steps && !rc; ++s) {
    rc = forward(ctx, M, R, st);
    if (!rc) rc = backward(ctx, M, R, T, T.loss + (int64_t)s * M->nfam, st);
    if (!rc) {
      ++M->step;
      AdamArgs a;
      a.p = M->p; a.g = M->g; a.m1 = M->m1; a.m2 = M->m2; a.pb = M->pb; a.n = np;
      a.lr = M->cfg.lr; a.wd = M->cfg.weight_decay;
      a.c1 = 1.f - std::pow(a.b1, (float)M->step);
      a.c2 = 1.f - std::pow(a.b2, (float)M->step);
      if (launch_adam(a, st) != hipSuccess) rc = fail(ctx, BLPG_E_HIP, "adam launch failed");
    }
  }
  if (!rc) {
    std::vector<float> lh((size_t)steps * M->nfam);
    if (hipMemcpyAsync(lh.data(), T.loss, sizeof(float) * lh.size(), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = fail(ctx, BLPG_E_HIP, "training failed");
    else if (loss) {
      int nf = 0;
      for (int f = 0; f < M->nfam; ++f) nf += R.fam_count[f] > 0;
      for (int s = 0; s < steps; ++s) {
        double acc = 0;
        for (int f = 0; f < M->nfam; ++f) acc += lh[(size_t)s * M->nfam + f];
        loss[s] = (float)(acc / std::max(nf, 1));
      }
    }
  }
  free_rows(R);
  return rc;
}

int blpg_ws_gradient(blpg_ctx* ctx, blpg_ws_model* M, const blpg_lp* lps, const int32_t* bases,
                     int64_t count, float* grad, float* loss) {
  if (!ctx || !M || !lps || !bases || count < 1)
    return fail(ctx, BLPG_E_INVALID, "invalid arguments to blpg_ws_gradient");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  WsRows R;
  TrainBufs T;
  int rc = setup_pass(ctx, M, lps, bases, count, 1, R, T);
  if (!rc && launch_fill_i32((int32_t*)T.loss, 0, M->nfam, st) != hipSuccess)
    rc = fail(ctx, BLPG_E_HIP, "loss buffer clear failed");
  if (!rc) rc = forward(ctx, M, R, st);
  if (!rc) rc = backward(ctx, M, R, T, T.loss, st);
  if (!rc) {
    std::vector<float> g((size_t)M->P * M->nfam), l(M->nfam);
    if (hipMemcpyAsync(g.data(), M->g, 4 * g.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(l.data(), T.loss, 4 * l.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = fail(ctx, BLPG_E_HIP, "gradient download failed");
    } else {
      const int64_t H = M->H;
      const int64_t cnt = H * M->din + H + H * H + H + (int64_t)M->dout * H + M->dout;
      for (int f = 0; f < M->nfam; ++f) {
        const float* q = g.data() + (size_t)f * M->P;
        float* o = grad ? grad + (size_t)f * cnt : nullptr;
        if (o) {
          for (int64_t h = 0; h < H; ++h)
            for (int j = 0; j < M->din; ++j) *o++ = q[M->oW1 + h * M->din_pad + j];
          for (int64_t h = 0; h < H; ++h) *o++ = q[M->ob1 + h];
          for (int64_t e = 0; e < H * H; ++e) *o++ = q[M->oW2 + e];
          for (int64_t h = 0; h < H; ++h) *o++ = q[M->ob2 + h];
          for (int64_t r = 0; r < M->dout; ++r)
            for (int64_t j = 0; j < H; ++j) *o++ = q[M->oW3 + r * H + j];
          for (int64_t r = 0; r < M->dout; ++r) *o++ = q[M->ob3 + r];
        }
        if (loss) loss[f] = l[f];
      }
    }
  }
  free_rows(R);
  return rc;
}

int blpg_ws_params(blpg_ctx* ctx, blpg_ws_model* M, int32_t family, float* get, const float* set,
                   int64_t* count) {
  if (!ctx || !M || family < 0 || family >= M->nfam)
    return fail(ctx, BLPG_E_INVALID, "invalid arguments to blpg_ws_params");
  const int64_t H = M->H;
  const int64_t cnt = H * M->din + H + H * H + H + (int64_t)M->dout * H + M->dout;
  if (count) *count = cnt;
  if (!get && !set) return BLPG_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::vector<float> q(M->P);
  float* dev = M->p + (size_t)family * M->P;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipMemcpy(q.data(), dev, 4 * q.size(), hipMemcpyDeviceToHost));
  // walk the unpadded layout; `io` copies one entry in the requested direction
  int64_t pos = 0;
  auto io = [&](int64_t at) {
    if (get) get[pos] = q[at];
    if (set) q[at] = set[pos];
    ++pos;
  };
  for (int64_t h = 0; h < H; ++h)
    for (int j = 0; j < M->din; ++j) io(M->oW1 + h * M->din_pad + j);
  for (int64_t h = 0; h < H; ++h) io(M->ob1 + h);
  for (int64_t e = 0; e < H * H; ++e) io(M->oW2 + e);
  for (int64_t h = 0; h < H; ++h) io(M->ob2 + h);
  for (int64_t r = 0; r < M->dout; ++r)
    for (int64_t j = 0; j < H; ++j) io(M->oW3 + r * H + j);
  for (int64_t r = 0; r < M->dout; ++r) io(M->ob3 + r);
  if (set) {
    // new parameters restart the optimiser of this family
    HIPCHK(ctx, hipMemcpy(dev, q.data(), 4 * q.size(), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemset(M->m1 + (size_t)family * M->P, 0, 4 * (size_t)M->P));
    HIPCHK(ctx, hipMemset(M->m2 + (size_t)family * M->P, 0, 4 * (size_t)M->P));
    HIPCHK(ctx, launch_to_bf16(dev, M->pb + (size_t)family * M->P, M->P, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  return BLPG_OK;
}

int blpg_ws_norm(blpg_ctx* ctx, const blpg_ws_model* M, int32_t family, float* mu, float* inv_sd) {
  if (!ctx || !M || family < 0 || family >= M->nfam)
    return fail(ctx, BLPG_E_INVALID, "invalid arguments to blpg_ws_norm");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (mu) HIPCHK(ctx, hipMemcpy(mu, M->mu + (size_t)family * M->din, 4 * M->din, hipMemcpyDeviceToHost));
  if (inv_sd)
    HIPCHK(ctx, hipMemcpy(inv_sd, M->inv_sd + (size_t)family * M->din, 4 * M->din, hipMemcpyDeviceToHost));
  return BLPG_OK;
}

int blpg_ws_predict(blpg_ctx* ctx, blpg_ws_model* M, const blpg_lp* lps, int64_t count,
                    float* logits, int32_t* warm) {
  if (!ctx || !M || (count > 0 && !lps) || count < 0)
    return fail(ctx, BLPG_E_INVALID, "invalid arguments to blpg_ws_predict");
  if (count == 0) return BLPG_OK;
  int rc = check_lps(ctx, M, lps, count, true);
  if (rc) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  std::vector<int32_t> fam(count), row_lp, row_fam;
  for (int64_t k = 0; k < count; ++k) fam[k] = lps[k].family;
  WsRows R;
  rc = plan_rows(ctx, M, fam, R, nullptr, count, 0, row_lp, row_fam);
  if (rc) { free_rows(R); return rc; }
  int32_t* d_warm = nullptr;
  const std::vector<float> raw = raw_features(M, lps, count);
  float* d_raw = R.raw;
  if (hipMemcpy(d_raw, raw.data(), 4 * raw.size(), hipMemcpyHostToDevice) != hipSuccess) {
    free_rows(R);
    return fail(ctx, BLPG_E_HIP, "feature upload failed");
  }
  FeatureArgs fa;
  fa.raw = d_raw; fa.ld_raw = M->din; fa.row_lp = R.row_lp; fa.row_fam = R.row_fam;
  fa.rows = R.rows; fa.d_in = M->din; fa.d_in_pad = M->din_pad; fa.mu = M->mu;
  fa.inv_sd = M->inv_sd; fa.X = R.X;
  if (launch_features(fa, st) != hipSuccess) rc = fail(ctx, BLPG_E_HIP, "features launch failed");
  if (!rc) rc = forward(ctx, M, R, st);
  if (!rc && warm) {
    if (hipMalloc(&d_warm, sizeof(int32_t) * (size_t)count * M->m) != hipSuccess)
      rc = fail(ctx, BLPG_E_NOMEM, "hipMalloc for warm lists failed");
    else if (launch_pick_basis(R.Z, M->dout_pad, R.row_lp, R.rows, nullptr, M->m, M->n, d_warm,
                               M->dout, st) != hipSuccess)
      rc = fail(ctx, BLPG_E_HIP, "pick launch failed");
  }
  if (!rc) {
    std::vector<float> Z((size_t)R.rows * M->dout_pad);
    if (hipMemcpyAsync(Z.data(), R.Z, 4 * Z.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
        (warm && hipMemcpyAsync(warm, d_warm, sizeof(int32_t) * (size_t)count * M->m,
                                hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = fail(ctx, BLPG_E_HIP, "prediction download failed");
    } else if (logits) {
      for (int64_t r = 0; r < R.rows; ++r) {
        const int k = row_lp[r];
        if (k < 0) continue;
        std::memcpy(logits + (size_t)k * M->dout, Z.data() + (size_t)r * M->dout_pad,
                    4 * (size_t)M->dout);
      }
    }
  }
  if (d_warm) (void)hipFree(d_warm);
  free_rows(R);
  return rc;
}

int blpg_batch_set_warmstart(blpg_ctx* ctx, blpg_batch* B, blpg_ws_model* M) {
  if (!ctx || !B) return fail(ctx, BLPG_E_INVALID, "invalid arguments to blpg_batch_set_warmstart");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  free_rows(B->ws);
  B->model = nullptr;
  B->g_gen[0] = B->g_gen[1] = -1;  // captured graphs hold the old rows: capture again
  if (!M) return BLPG_OK;
  std::vector<int32_t> fam(B->count, -1), row_lp, row_fam;
  for (int64_t k = 0; k < B->count; ++k) {
    const int f = B->family[k];
    if (f >= 0 && f < M->nfam && B->desc[k].m == M->m && B->desc[k].n == M->n) fam[k] = f;
  }
  int rc = plan_rows(ctx, M, fam, B->ws, nullptr, 0, 0, row_lp, row_fam);
  if (rc) { free_rows(B->ws); return rc; }
  if (!B->d_warm) {
    if (hipMalloc(&B->warm_mem, sizeof(int32_t) * std::max<int64_t>(B->basis_elems, 1)) != hipSuccess) {
      free_rows(B->ws);
      return fail(ctx, BLPG_E_NOMEM, "hipMalloc for warm lists failed");
    }
    B->d_warm = (int32_t*)B->warm_mem;
    HIPCHK(ctx, launch_fill_i32(B->d_warm, -1, B->basis_elems, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  B->model = M;
  warm_route(B);
  return BLPG_OK;
}

}  // extern "C"
