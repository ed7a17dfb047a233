// abi.cpp -- the C ABI of libblpg.so: contexts, packed device batches, size-bin routing,
// solve / verify / fallback pipeline and result download.
//
// Replaces the host half of the reference's Simplex class (simplex.cu:387-978): instead of
// building full tableaux on the host with OpenMP and copying them to the device per call
// (simplex.cu:506-534, 944-953), the batch is packed once as compact (A, b, c, sense)
// records, the kernels build their tableaux on the device, and only results come back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/blpg/blpg.h"
#include "blpg_internal.hpp"
#include "kernels.hpp"
#include "ctx.hpp"
#include "tuner.hpp"

using namespace blpg;

// ------------------------------------------------------------------------------------
namespace {

void resolve(const blpg_options& o, double& eps_d, double& eps_p, double& tol, double& kkt) {
  const bool f32 = o.precision == BLPG_FP32;
  eps_d = o.eps_d > 0 ? o.eps_d : (f32 ? 1e-6 : 1e-9);
  eps_p = o.eps_p > 0 ? o.eps_p : (f32 ? 1e-6 : 1e-9);
  tol = o.tol_feas > 0 ? o.tol_feas : (f32 ? 1e-4 : 1e-7);
  kkt = o.kkt_tol > 0 ? o.kkt_tol : 1e-9;
}

int bland_after(const blpg_options& o) { return o.bland_after > 0 ? o.bland_after : 50; }

bool finite_all(const double* p, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

int grow(blpg_ctx* ctx, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes) return BLPG_OK;
  ++ctx->gen;  // captured graphs that hold the old buffer must be captured again
  if (b.p) {
    (void)hipDeviceSynchronize();
    (void)hipFree(b.p);
  }
  b.p = nullptr;
  b.bytes = 0;
  if (hipMalloc(&b.p, bytes) != hipSuccess)
    return fail(ctx, BLPG_E_NOMEM, "hipMalloc(%zu) failed", bytes);
  b.bytes = bytes;
  return BLPG_OK;
}

int persistent_grid(const blpg_ctx* ctx, int per_cu, int64_t count) {
  const int64_t g = (int64_t)ctx->props.multiProcessorCount * std::max(per_cu, 1);
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, count));
}

// Copies the LPs' A, b, c and senses into the batch's pinned staging at the offsets the
// upload chose (every distinct array once).
void stage_inputs(blpg_batch* B, const blpg_lp* lps) {
  double* pool = (double*)(B->h_stage + B->o_pool);
  int8_t* spool = (int8_t*)(B->h_stage + B->o_spool);
  for (int64_t k = 0; k < B->count; ++k) {
    const blpg_lp& L = lps[k];
    const LpDesc& d = B->desc[k];
    std::memcpy(pool + d.a, L.A, sizeof(double) * (size_t)L.m * L.n);
    std::memcpy(pool + d.b, L.b, sizeof(double) * L.m);
    std::memcpy(pool + d.c, L.c, sizeof(double) * L.n);
    if (L.sense) std::memcpy(spool + d.s, L.sense, L.m);
  }
}

// Warm-start lists of the LPs that carry one (-1 elsewhere and for invalid ids).
hipError_t upload_warm(blpg_batch* B, const blpg_lp* lps) {
  std::vector<int32_t> w(B->basis_elems, -1);
  for (int64_t k = 0; k < B->count; ++k)
    if (lps[k].warm_basis)
      for (int i = 0; i < lps[k].m; ++i) {
        const int32_t v = lps[k].warm_basis[i];
        w[B->desc[k].basis + i] = (v >= 0 && v < lps[k].n + lps[k].m) ? v : -1;
      }
  return hipMemcpy(B->d_warm, w.data(), sizeof(int32_t) * w.size(), hipMemcpyHostToDevice);
}

// Packs the staging pool's ranges for the copy of the next solve (see blpg_batch::ZRange).
// Without room for the packed buffers every range is copied dense.
void pack_ranges(blpg_batch* B) {
  auto a8 = [](size_t x) { return (x + 7) & ~(size_t)7; };
  auto a256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
  std::vector<std::pair<int64_t, int64_t>> rng;
  rng.push_back({0, B->common_elems});
  for (const Bin& b : B->bins) rng.push_back({b.pool_lo, b.pool_hi});
  size_t cap = 0;
  for (const auto& r : rng) {
    const int64_t n = r.second - r.first, nw = (n + 63) / 64;
    cap += a256(8 * (size_t)nw + a8(4 * (size_t)nw) + 8 * (size_t)n);
  }
  B->zr.clear();
  if (cap > B->z_cap) {
    if (B->h_z) (void)hipHostFree(B->h_z);
    if (B->d_z) (void)hipFree(B->d_z);
    B->h_z = nullptr;
    B->d_z = nullptr;
    B->z_cap = 0;
    if (hipHostMalloc((void**)&B->h_z, cap, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&B->d_z, cap) != hipSuccess) {
      if (B->h_z) (void)hipHostFree(B->h_z);
      B->h_z = nullptr;
      B->d_z = nullptr;
      return;
    }
    B->z_cap = cap;
  }
  const double* pool = (const double*)(B->h_stage + B->o_pool);
  size_t off = 0;
  for (const auto& r : rng) {
    blpg_batch::ZRange z;
    z.lo = r.first;
    z.hi = r.second;
    z.off = off;
    const int64_t n = z.hi - z.lo, nw = (n + 63) / 64;
    uint64_t* mask = (uint64_t*)(B->h_z + off);
    int32_t* pref = (int32_t*)(B->h_z + off + 8 * nw);
    double* vals = (double*)(B->h_z + off + 8 * nw + a8(4 * (size_t)nw));
    int64_t nnz = 0;
    for (int64_t w = 0; w < nw && n < INT32_MAX; ++w) {
      uint64_t mw = 0;
      pref[w] = (int32_t)nnz;
      const int64_t e = std::min<int64_t>(64, n - 64 * w);
      for (int64_t u = 0; u < e; ++u) {
        const double v = pool[z.lo + 64 * w + u];
        uint64_t bits;
        std::memcpy(&bits, &v, 8);
        if (bits != 0) {  // -0.0 is kept as a value: only +0.0 is implied
          mw |= 1ull << u;
          vals[nnz++] = v;
        }
      }
      mask[w] = mw;
    }
    z.bytes = 8 * (size_t)nw + a8(4 * (size_t)nw) + 8 * (size_t)nnz;
    z.packed = n > 0 && n < INT32_MAX && z.bytes * 10 <= (size_t)n * 8 * 9;
    off += a256(8 * (size_t)nw + a8(4 * (size_t)nw) + 8 * (size_t)n);
    B->zr.push_back(z);
  }
}

}  // namespace

// message of the last failed call that takes no context, per host thread
static thread_local char g_free_err[512] = "no context";

int blpg::set_free_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_free_err, sizeof g_free_err, fmt, ap);
  va_end(ap);
  return code;
}

void blpg::warm_route(blpg_batch* B) {
  for (size_t i = 0; i < B->bins.size(); ++i) {
    Bin& b = B->bins[i];
    if (b.kernel != BLPG_KERNEL_WAVE) continue;
    const int tv = Tuner::best_tile(b.rcap - 2, b.ccap - 2, B->precision);
    if (tv < 0) continue;
    b.kernel = BLPG_KERNEL_TILE;
    b.variant = tv;
    B->info.kernel[i] = b.kernel;
    B->info.param[i] = b.variant;
  }
}

static void drop_graph(blpg_batch* B, int g) {
  if (B->gexec[g]) (void)hipGraphExecDestroy(B->gexec[g]);
  if (B->graph[g]) (void)hipGraphDestroy(B->graph[g]);
  B->gexec[g] = nullptr;
  B->graph[g] = nullptr;
}

// ------------------------------------------------------------------------------------
extern "C" {

void blpg_default_options(blpg_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->precision = BLPG_FP64;
  o->verify = 1;
  o->kernel = BLPG_KERNEL_AUTO;
  o->use_graph = 1;
}

const char* blpg_version(void) { return "blpg-mi355x 0.1 (gfx950)"; }

int blpg_create(blpg_ctx** out, const blpg_options* opt) {
  if (!out) return BLPG_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BLPG_E_NODEVICE;
  blpg_ctx* ctx = new blpg_ctx();
  if (opt) ctx->opt = *opt;
  else blpg_default_options(&ctx->opt);
  ctx->device = ctx->opt.device;
  if (ctx->device < 0 || ctx->device >= ndev) { delete ctx; return BLPG_E_NODEVICE; }
  if (hipSetDevice(ctx->device) != hipSuccess ||
      hipGetDeviceProperties(&ctx->props, ctx->device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return BLPG_E_HIP;
  }
  ctx->tuner.init(ctx->props);
  // the aux streams are created when a batch first needs them (a one-bin batch, such as a
  // serving arena, never does: its context then holds a single stream, and the process's
  // few hardware queues are not taken by idle streams)
  if (hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming) != hipSuccess) {
    blpg_destroy(ctx);
    return BLPG_E_HIP;
  }
  *out = ctx;
  return BLPG_OK;
}

void blpg_destroy(blpg_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) { (void)hipStreamSynchronize(ctx->stream); (void)hipStreamDestroy(ctx->stream); }
  for (int i = 0; i < blpg_ctx::kAux; ++i)
    if (ctx->aux[i]) { (void)hipStreamSynchronize(ctx->aux[i]); (void)hipStreamDestroy(ctx->aux[i]); }
  if (ctx->fork) (void)hipEventDestroy(ctx->fork);
  if (ctx->ws.p) (void)hipFree(ctx->ws.p);
  if (ctx->ws2.p) (void)hipFree(ctx->ws2.p);
  if (ctx->vws.p) (void)hipFree(ctx->vws.p);
  delete ctx;
}

const char* blpg_last_error(const blpg_ctx* ctx) { return ctx ? ctx->err.c_str() : g_free_err; }

int blpg_set_options(blpg_ctx* ctx, const blpg_options* opt) {
  if (!ctx || !opt) return BLPG_E_INVALID;
  if (opt->device != ctx->device) return fail(ctx, BLPG_E_INVALID, "device cannot change");
  ctx->opt = *opt;
  return BLPG_OK;
}

int blpg_batch_upload(blpg_ctx* ctx, const blpg_lp* lps, int64_t count, blpg_batch** out) {
  if (!ctx || !out || (count > 0 && !lps) || count < 0)
    return fail(ctx, BLPG_E_INVALID, "invalid arguments to blpg_batch_upload");
  if (count > INT32_MAX / 2) return fail(ctx, BLPG_E_INVALID, "batch too large");
  *out = nullptr;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  blpg_batch* B = new blpg_batch();
  B->count = count;
  B->precision = ctx->opt.precision;
  B->desc.resize(count);
  B->family.resize(count);
  for (int64_t k = 0; k < count; ++k) B->family[k] = lps ? lps[k].family : -1;
  // shapes, senses and output offsets; the input pools are laid out once the bins are known
  for (int64_t k = 0; k < count; ++k) {
    const blpg_lp& L = lps[k];
    if (L.m < 1 || L.n < 1 || !L.A || !L.b || !L.c) {
      delete B;
      return fail(ctx, BLPG_E_INVALID, "LP %lld: m, n must be >= 1 and A, b, c non-NULL",
                  (long long)k);
    }
    if (L.m > 4096 || L.n > 4096) {
      delete B;
      return fail(ctx, BLPG_E_INVALID, "LP %lld: m, n above 4096 not supported", (long long)k);
    }
    if (L.sense)
      for (int i = 0; i < L.m; ++i)
        if (L.sense[i] < 0 || L.sense[i] > 2) {
          delete B;
          return fail(ctx, BLPG_E_INVALID, "LP %lld: sense[%d] not in {0,1,2}", (long long)k, i);
        }
    LpDesc& d = B->desc[k];
    d.m = L.m;
    d.n = L.n;
    d.x = B->x_elems;
    B->x_elems += L.n;
    d.basis = B->basis_elems;
    B->basis_elems += L.m;
    d.t = B->tab_elems;
    B->tab_elems += (int64_t)(L.m + 2) * (L.n + 2) + (L.n + 3) / 2;
    B->mmax = std::max(B->mmax, L.m);
    B->nmax = std::max(B->nmax, L.n);
  }

  // ---- size bins, routed by the tuner: one bin per kernel instance (a tile variant or a
  //      streamed kernel handles every LP up to its capacity); LDS bins are split by size
  //      class and folded into the HBM-streamed kernel if the batch would need > 16 bins ----
  std::vector<Route> routes(count);
  for (int64_t k = 0; k < count; ++k)
    routes[k] = ctx->tuner.route(B->desc[k].m, B->desc[k].n, ctx->opt.precision, ctx->opt.kernel);
  auto bin_key = [](const Route& r) -> long {
    if (r.kernel == BLPG_KERNEL_TILE) return 1000000L + r.variant;
    if (r.kernel == BLPG_KERNEL_WAVE) return 4000000L + r.variant;
    if (r.kernel == BLPG_KERNEL_STREAM) return 2000000L + r.nt;
    return 3000000L + (long)r.nt * 100000L + r.bin;  // LDS: per size class
  };
  for (int pass = 0; pass < 2; ++pass) {
    std::map<long, int> binof;
    B->bins.clear();
    for (int64_t k = 0; k < count; ++k) {
      const Route& r = routes[k];
      const long key = bin_key(r);
      auto it = binof.find(key);
      int bi;
      if (it == binof.end()) {
        bi = (int)B->bins.size();
        binof[key] = bi;
        Bin b;
        b.kernel = r.kernel;
        b.variant = r.variant;
        b.nt = r.nt;
        B->bins.push_back(b);
      } else {
        bi = it->second;
      }
      Bin& b = B->bins[bi];
      b.ids.push_back((int32_t)k);
      b.rcap = std::max(b.rcap, B->desc[k].m + 2);
      b.ccap = std::max(b.ccap, B->desc[k].n + 2);
    }
    if (B->bins.size() <= 16) break;
    for (auto& r : routes)
      if (r.kernel == BLPG_KERNEL_LDS) r.kernel = BLPG_KERNEL_STREAM;
  }
  if (B->bins.size() > 16) {
    delete B;
    return fail(ctx, BLPG_E_INVALID, "batch needs more than 16 kernel instances");
  }

  // ---- input pools.  Every distinct input array, keyed by (pointer, length), is stored
  //      once (shared-A batches and parametric families share A; two prefixes of one buffer
  //      with different lengths are different arrays).  An array that only one bin reads
  //      lives in that bin's contiguous range, so a solve can copy each bin's inputs to the
  //      device on the bin's own stream just before its kernel; arrays several bins read
  //      form a common range in front, copied first. ----
  {
    std::map<SrcKey, int> key_bin;  // the one bin reading the array, or -1 (several)
    for (int bi = 0; bi < (int)B->bins.size(); ++bi)
      for (int32_t k : B->bins[bi].ids) {
        const blpg_lp& L = lps[k];
        for (const SrcKey& key : {SrcKey(L.A, (int64_t)L.m * L.n), SrcKey(L.b, L.m), SrcKey(L.c, L.n)}) {
          auto it = key_bin.find(key);
          if (it == key_bin.end()) key_bin[key] = bi;
          else if (it->second != bi) it->second = -1;
        }
      }
    int64_t bad = -1;
    auto place = [&](const double* p, int64_t len, int64_t k) {
      const SrcKey key(p, len);
      if (B->pool_off.count(key)) return;
      if (bad < 0 && !finite_all(p, len)) bad = k;
      B->pool_off[key] = B->pool_elems;
      B->pool_elems += (len + 7) & ~7ll;
    };
    auto place_lp = [&](int64_t k, int want_bin) {
      const blpg_lp& L = lps[k];
      if (key_bin[SrcKey(L.A, (int64_t)L.m * L.n)] == want_bin) place(L.A, (int64_t)L.m * L.n, k);
      if (key_bin[SrcKey(L.b, L.m)] == want_bin) place(L.b, L.m, k);
      if (key_bin[SrcKey(L.c, L.n)] == want_bin) place(L.c, L.n, k);
    };
    for (int64_t k = 0; k < count; ++k) place_lp(k, -1);
    B->common_elems = B->pool_elems;
    for (int bi = 0; bi < (int)B->bins.size(); ++bi) {
      Bin& b = B->bins[bi];
      b.pool_lo = B->pool_elems;
      for (int32_t k : b.ids) place_lp(k, bi);
      b.pool_hi = B->pool_elems;
    }
    if (bad >= 0) {
      delete B;
      return fail(ctx, BLPG_E_INVALID, "LP %lld: A, b or c has NaN/Inf", (long long)bad);
    }
    for (int64_t k = 0; k < count; ++k) {
      const blpg_lp& L = lps[k];
      LpDesc& d = B->desc[k];
      d.a = B->pool_off[SrcKey(L.A, (int64_t)L.m * L.n)];
      d.b = B->pool_off[SrcKey(L.b, L.m)];
      d.c = B->pool_off[SrcKey(L.c, L.n)];
      if (L.sense) {
        const SrcKey key(L.sense, L.m);
        auto it = B->spool_off.find(key);
        if (it != B->spool_off.end()) d.s = it->second;
        else {
          d.s = B->spool_elems;
          B->spool_elems += (L.m + 15) & ~15ll;
          B->spool_off[key] = d.s;
        }
      } else {
        d.s = -1;
      }
    }
  }

  // ---- one device allocation, carved ----
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const int64_t n = count;
  const int nb = (int)B->bins.size();
  for (int i = 0; i < std::min(nb - 1, blpg_ctx::kAux); ++i)
    if (!ctx->aux[i] && hipStreamCreateWithFlags(&ctx->aux[i], hipStreamNonBlocking) != hipSuccess) {
      delete B;
      return fail(ctx, BLPG_E_HIP, "stream creation failed");
    }
  size_t off = 0;
  const size_t o_desc = off; off += al(sizeof(LpDesc) * n);
  const size_t o_pool = off; off += al(sizeof(double) * B->pool_elems);
  const size_t o_spool = off; off += al(B->spool_elems + 16);
  const size_t o_i32 = off; off += al(sizeof(int32_t) * n * 12);    // status pivots p1 unb ok fb all fail idx defer defer2 lastpiv
  const size_t o_f64 = off; off += al(sizeof(double) * n * 6);      // obj wfin chkobj kktp kktd amax
  const size_t o_x = off; off += al(sizeof(double) * B->x_elems);
  const size_t o_basis = off; off += al(sizeof(int32_t) * B->basis_elems);
  const size_t o_tab = off; off += al(8 * (size_t)B->tab_elems);
  const size_t o_cnt = off; off += al(sizeof(int32_t) * 256);
  bool any_warm = false;
  for (int64_t k = 0; k < count; ++k) any_warm |= lps[k].warm_basis != nullptr;
  const size_t o_warm = off;
  if (any_warm) off += al(sizeof(int32_t) * B->basis_elems);
  B->mem_bytes = off;
  if (ctx->opt.hbm_budget > 0 && (int64_t)off > ctx->opt.hbm_budget) {
    delete B;
    return fail(ctx, BLPG_E_BUDGET, "batch needs %zu bytes, budget %lld", off,
                (long long)ctx->opt.hbm_budget);
  }
  if (hipMalloc(&B->mem, off) != hipSuccess) {
    delete B;
    return fail(ctx, BLPG_E_NOMEM, "hipMalloc(%zu) for the batch failed", off);
  }
  char* base = (char*)B->mem;
  B->d_desc = (LpDesc*)(base + o_desc);
  B->d_pool = (double*)(base + o_pool);
  B->d_spool = (int8_t*)(base + o_spool);
  int32_t* i32 = (int32_t*)(base + o_i32);
  B->out.status = i32;
  B->out.pivots = i32 + n;
  B->out.p1 = i32 + 2 * n;
  B->out.unb = i32 + 3 * n;
  B->chk.ok = i32 + 4 * n;
  B->d_fb = i32 + 5 * n;
  B->chk.fb = B->d_fb;
  B->d_all = i32 + 6 * n;
  B->d_fail = i32 + 7 * n;
  B->d_defer = i32 + 9 * n;
  B->d_defer2 = i32 + 10 * n;
  B->out.lastpiv = i32 + 11 * n;
  int32_t* d_binidx = i32 + 8 * n;
  double* f64 = (double*)(base + o_f64);
  B->out.obj = f64;
  B->out.wfin = f64 + n;
  B->chk.obj = f64 + 2 * n;
  B->chk.kkt_p = f64 + 3 * n;
  B->chk.kkt_d = f64 + 4 * n;
  B->out.amax = f64 + 5 * n;
  B->out.x = (double*)(base + o_x);
  B->res_off = o_i32;
  B->res_bytes = o_basis + sizeof(int32_t) * B->basis_elems - o_i32;
  B->x_off.resize(n);
  B->basis_off.resize(n);
  for (int64_t k = 0; k < n; ++k) { B->x_off[k] = B->desc[k].x; B->basis_off[k] = B->desc[k].basis; }
  B->out.basis = (int32_t*)(base + o_basis);
  B->d_tab = base + o_tab;
  B->out.tab = nullptr;
  B->d_counters = (int32_t*)(base + o_cnt);
  if (any_warm) B->d_warm = (int32_t*)(base + o_warm);

  // ---- host staging (pinned, kept with the batch for repacks) and one upload ----
  B->o_pool = o_pool;
  B->o_spool = o_spool;
  B->stage_bytes = o_i32 + sizeof(int32_t) * n * 9;
  if (hipHostMalloc((void**)&B->h_stage, B->stage_bytes, hipHostMallocDefault) != hipSuccess) {
    B->h_stage = nullptr;
    (void)hipFree(B->mem);
    delete B;
    return fail(ctx, BLPG_E_NOMEM, "hipHostMalloc(%zu) failed", B->stage_bytes);
  }
  char* h = B->h_stage;
  std::memcpy(h + o_desc, B->desc.data(), sizeof(LpDesc) * n);
  stage_inputs(B, lps);
  int32_t* hi = (int32_t*)(h + o_i32);
  std::memset(hi, 0, sizeof(int32_t) * n * 8);
  for (int64_t k = 0; k < n; ++k) hi[6 * n + k] = (int32_t)k;
  int64_t pos = 0;
  for (auto& b : B->bins) {
    b.d_idx = d_binidx + pos;
    std::memcpy(hi + 8 * n + pos, b.ids.data(), sizeof(int32_t) * b.ids.size());
    pos += (int64_t)b.ids.size();
  }
  hipError_t e = hipMemcpyAsync(B->mem, h, B->stage_bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && any_warm) e = upload_warm(B, lps);
  if (e != hipSuccess) {
    (void)hipHostFree(B->h_stage);
    (void)hipFree(B->mem);
    delete B;
    return fail(ctx, BLPG_E_HIP, "batch upload failed: %s", hipGetErrorString(e));
  }
  for (int i = 0; i < 96; ++i)
    if (hipEventCreate(&B->ev[i]) != hipSuccess) {
      blpg_batch_free(ctx, B);
      return fail(ctx, BLPG_E_HIP, "hipEventCreate failed");
    }
  for (int i = 0; i < 16; ++i)
    if (hipEventCreateWithFlags(&B->join[i], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&B->copied[i], hipEventDisableTiming) != hipSuccess) {
      blpg_batch_free(ctx, B);
      return fail(ctx, BLPG_E_HIP, "hipEventCreate failed");
    }
  B->info.nbins = nb;
  if (any_warm) warm_route(B);
  for (int i = 0; i < nb; ++i) {
    B->info.kernel[i] = B->bins[i].kernel;
    B->info.rcap[i] = B->bins[i].rcap;
    B->info.ccap[i] = B->bins[i].ccap;
    B->info.count[i] = (int64_t)B->bins[i].ids.size();
    B->info.param[i] = (B->bins[i].kernel == BLPG_KERNEL_TILE || B->bins[i].kernel == BLPG_KERNEL_WAVE)
                           ? B->bins[i].variant : B->bins[i].nt;
  }
  *out = B;
  return BLPG_OK;
}

void blpg_batch_free(blpg_ctx* ctx, blpg_batch* B) {
  if (!B) return;
  if (ctx) (void)hipSetDevice(ctx->device);
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  drop_graph(B, 0);
  drop_graph(B, 1);
  for (int i = 0; i < 96; ++i)
    if (B->ev[i]) (void)hipEventDestroy(B->ev[i]);
  for (int i = 0; i < 16; ++i) {
    if (B->join[i]) (void)hipEventDestroy(B->join[i]);
    if (B->copied[i]) (void)hipEventDestroy(B->copied[i]);
  }
  if (B->h_stage) (void)hipHostFree(B->h_stage);
  if (B->h_z) (void)hipHostFree(B->h_z);
  if (B->d_z) (void)hipFree(B->d_z);
  ws_free_rows(B->ws);
  if (B->h_res) (void)hipHostFree(B->h_res);
  if (B->warm_mem) (void)hipFree(B->warm_mem);
  if (B->mem) (void)hipFree(B->mem);
  delete B;
}

// Launch plan of one size bin: kernel, persistent grid and HBM slab size.
struct BinPlan {
  int grid = 1;
  size_t ws_bytes = 0;
  int64_t ws_stride = 0;
};

static int plan_bin(blpg_ctx* ctx, const Bin& b, int precision, int64_t count, BinPlan& P) {
  if (b.kernel == BLPG_KERNEL_WAVE) {
    const int per_cu = wave_max_blocks_per_cu(precision, b.variant);
    if (per_cu <= 0) return fail(ctx, BLPG_E_INVALID, "wave variant %d cannot be resident", b.variant);
    P.grid = persistent_grid(ctx, per_cu, count);
    return BLPG_OK;
  }
  if (b.kernel == BLPG_KERNEL_TILE) {
    const int per_cu = tile_max_blocks_per_cu(precision, b.variant);
    if (per_cu <= 0) return fail(ctx, BLPG_E_INVALID, "tile variant %d cannot be resident", b.variant);
    P.grid = persistent_grid(ctx, per_cu, count);
    return BLPG_OK;
  }
  const bool lds = b.kernel == BLPG_KERNEL_LDS;
  const int per_cu = stream_max_blocks_per_cu(precision, lds, b.nt, b.rcap, b.ccap);
  if (per_cu <= 0)
    return fail(ctx, BLPG_E_INVALID, "kernel cannot be resident (rcap %d ccap %d lds %d)",
                b.rcap, b.ccap, (int)lds);
  P.grid = persistent_grid(ctx, per_cu, count);
  if (!lds) {
    const size_t esz = precision == BLPG_FP64 ? 8 : 4;
    P.ws_stride = ((int64_t)b.rcap * b.ccap + 63) & ~63ll;
    P.ws_bytes = (esz * (size_t)P.ws_stride * P.grid + 255) & ~(size_t)255;
  }
  return BLPG_OK;
}

static int launch_bin(blpg_ctx* ctx, blpg_batch* B, const Bin& b, const BinPlan& P,
                      int precision, const int32_t* idx, int32_t count,
                      const int32_t* count_dev, int32_t* queue, void* ws,
                      const int32_t* warm, hipStream_t st) {
  double eps_d, eps_p, tol, kkt;
  blpg_options o = ctx->opt;
  o.precision = precision;
  resolve(o, eps_d, eps_p, tol, kkt);
  HIPCHK(ctx, launch_fill_i32(queue, 0, 1, st));
  if (b.kernel == BLPG_KERNEL_TILE || b.kernel == BLPG_KERNEL_WAVE) {
    TileLaunch L{};
    L.precision = precision;
    L.variant = b.variant;
    L.desc = B->d_desc; L.pool = B->d_pool; L.spool = B->d_spool;
    L.idx = idx; L.count = count; L.count_dev = count_dev; L.queue = queue; L.out = B->out;
    L.eps_d = eps_d; L.eps_p = eps_p; L.tol = tol;
    L.max_iter = o.max_iter; L.bland_after = bland_after(o);
    L.grid = P.grid;
    L.warm = warm;
    if (b.kernel == BLPG_KERNEL_WAVE) HIPCHK(ctx, launch_wave(L, st));
    else HIPCHK(ctx, launch_tile(L, st));
    return BLPG_OK;
  }
  StreamLaunch L{};
  L.precision = precision;
  L.lds_tab = b.kernel == BLPG_KERNEL_LDS;
  L.nt = b.nt;
  L.desc = B->d_desc; L.pool = B->d_pool; L.spool = B->d_spool;
  L.idx = idx; L.count = count; L.count_dev = count_dev; L.queue = queue; L.out = B->out;
  L.eps_d = eps_d; L.eps_p = eps_p; L.tol = tol;
  L.max_iter = o.max_iter; L.bland_after = bland_after(o);
  L.rcap = b.rcap; L.ccap = b.ccap;
  L.grid = P.grid;
  L.ws = ws;
  L.ws_stride = P.ws_stride;
  L.warm = warm;
  HIPCHK(ctx, launch_stream(L, st));
  return BLPG_OK;
}

// Launch plans of every bin (grids, HBM workspace offsets, check-pass sizes), with the
// context workspaces grown to fit: the HBM slabs of streamed bins, of the FP64 fallback,
// and the LU slabs of the check (bins run concurrently, so each gets its own region).
struct SolvePlan {
  std::vector<BinPlan> plans, fplans;
  std::vector<Bin> fbins;
  std::vector<size_t> ws_off, fws_off, vws_off;
  std::vector<std::vector<int>> vtiers;  // per bin: (tier, grid) pairs of the check, in order
  std::vector<int64_t> vstride;
};

static int plan_solve(blpg_ctx* ctx, blpg_batch* B, SolvePlan& P) {
  const int nb = (int)B->bins.size();
  const int precision = ctx->opt.precision;
  const bool do_verify = ctx->opt.verify || precision == BLPG_FP32;
  const bool fp32 = precision == BLPG_FP32;
  P.plans.assign(nb, BinPlan{}); P.fplans.assign(nb, BinPlan{}); P.fbins.assign(nb, Bin{});
  P.ws_off.assign(nb, 0); P.fws_off.assign(nb, 0); P.vws_off.assign(nb, 0);
  P.vtiers.assign(nb, {});
  P.vstride.assign(nb, 0);
  std::vector<BinPlan>& plans = P.plans;
  std::vector<BinPlan>& fplans = P.fplans;
  std::vector<Bin>& fbins = P.fbins;
  std::vector<size_t>& ws_off = P.ws_off;
  std::vector<size_t>& fws_off = P.fws_off;
  std::vector<size_t>& vws_off = P.vws_off;
  std::vector<int64_t>& vstride = P.vstride;
  size_t ws_total = 0, fws_total = 0, vws_total = 0;
  for (int i = 0; i < nb; ++i) {
    const Bin& b = B->bins[i];
    const int64_t cnt = (int64_t)b.ids.size();
    int rc = plan_bin(ctx, b, precision, cnt, plans[i]);
    if (rc) return rc;
    ws_off[i] = ws_total;
    ws_total += plans[i].ws_bytes;
    if (do_verify) {
      const int mc = b.rcap - 2, nc = b.ccap - 2;
      // tiers of growing LDS per LP; the last one used takes whatever the others defer
      const int tsz = fp32 ? 4 : 8;
      std::vector<int>& tv = P.vtiers[i];
      const bool skip0 = (ctx->tuner.check_flags(precision, mc, nc) & 1) != 0;
      for (int tier = 0; tier < 3; ++tier) {
        if (!verify_tier_used(tier, mc, nc, tsz) || (tier == 0 && skip0)) continue;
        tv.push_back(tier);
        tv.push_back(persistent_grid(ctx, verify_max_blocks_per_cu(tier, mc, nc), cnt));
        if (tier < 2 && verify_tier_covers(tier, mc, nc, tsz)) break;
      }
      vstride[i] = (int64_t)verify_ws_stride(mc, nc);
      vws_off[i] = vws_total;
      if (tv[tv.size() - 2] == 2) vws_total += (size_t)vstride[i] * tv.back();
    }
    if (fp32) {
      // FP64 fallback kernel of this bin
      Bin& f = fbins[i];
      f.rcap = b.rcap; f.ccap = b.ccap; f.nt = Tuner::stream_threads(b.rcap - 2, b.ccap - 2);
      const int wv = Tuner::best_wave(b.rcap - 2, b.ccap - 2, BLPG_FP64);
      const int tv = Tuner::best_tile(b.rcap - 2, b.ccap - 2, BLPG_FP64);
      if (wv >= 0) { f.kernel = BLPG_KERNEL_WAVE; f.variant = wv; }
      else if (tv >= 0) { f.kernel = BLPG_KERNEL_TILE; f.variant = tv; }
      else if (stream_lds_bytes(BLPG_FP64, true, f.nt, f.rcap, f.ccap) <= ctx->tuner.lds_limit() &&
               stream_max_blocks_per_cu(BLPG_FP64, true, f.nt, f.rcap, f.ccap) > 0)
        f.kernel = BLPG_KERNEL_LDS;
      else f.kernel = BLPG_KERNEL_STREAM;
      rc = plan_bin(ctx, f, BLPG_FP64, cnt, fplans[i]);
      if (rc) return rc;
      fws_off[i] = fws_total;
      fws_total += fplans[i].ws_bytes;
    }
  }
  if (ws_total) { const int rc = grow(ctx, ctx->ws, ws_total); if (rc) return rc; }
  if (fws_total) { const int rc = grow(ctx, ctx->ws2, fws_total); if (rc) return rc; }
  if (vws_total) { const int rc = grow(ctx, ctx->vws, vws_total); if (rc) return rc; }
  return BLPG_OK;
}

static int plan_workspaces(blpg_ctx* ctx, blpg_batch* B) {
  SolvePlan P;
  return plan_solve(ctx, B, P);
}

// Event slots of a batch: ev[0] start; per bin i: ev[1+4i] solve start, ev[2+4i] solve end,
// ev[3+4i] verify end, ev[4+4i] fallback end; ev[65] end of everything; ev[66], ev[67]
// bracket the warm-start inference.
static constexpr int kEvEnd = 65;
static constexpr int kEvWsStart = 66, kEvWsEnd = 67;  // learned warm start inference

// Enqueues one solve of the batch on the context stream (and the aux streams it forks
// to).  With `timing` every stage is bracketed by timing events; without, only the
// fork / join events are recorded, so the sequence can be captured into a hipGraph.
// Workspaces are grown first: nothing in the enqueued sequence allocates.
static int enqueue_solve(blpg_ctx* ctx, blpg_batch* B, bool timing, bool copy) {
  hipStream_t st = ctx->stream;
  const int precision = ctx->opt.precision;
  B->precision = precision;
  int32_t* q = B->d_counters;
  const int nb = (int)B->bins.size();
  double eps_d, eps_p, tol, kkt;
  resolve(ctx->opt, eps_d, eps_p, tol, kkt);
  const bool do_verify = ctx->opt.verify || precision == BLPG_FP32;
  const bool fp32 = precision == BLPG_FP32;

  SolvePlan P;
  {
    const int rc = plan_solve(ctx, B, P);
    if (rc) return rc;
  }
  std::vector<BinPlan>& plans = P.plans;
  std::vector<BinPlan>& fplans = P.fplans;
  std::vector<Bin>& fbins = P.fbins;
  std::vector<size_t>& ws_off = P.ws_off;
  std::vector<size_t>& fws_off = P.fws_off;
  std::vector<size_t>& vws_off = P.vws_off;

  if (timing) HIPCHK(ctx, hipEventRecord(B->ev[0], st));
  HIPCHK(ctx, launch_fill_i32(B->d_fb, 0, B->count, st));
  HIPCHK(ctx, launch_fill_i32(q, 0, 256, st));
  // the solve kernels dump their final tableaux for the FP64 check
  B->out.tab = do_verify ? B->d_tab : nullptr;
  // inputs from the pinned staging (after a repack): the arrays several bins read and the
  // senses first; each bin's own range on its stream just before its kernel (below), in
  // launch order, one copy after the other -- unless the warm-start network, which reads
  // every family LP's b and c, runs first
  double* hpool = copy ? (double*)(B->h_stage + B->o_pool) : nullptr;
  // range r (0: the shared arrays, 1 + i: bin i) to the device: packed (the copy, then its
  // expansion into the pool on the same stream) or dense
  // (`done`, if given, is recorded once the host-to-device copy itself is enqueued, so the
  // next copy need not wait for the expansion)
  auto copy_range = [&](int r, int64_t lo, int64_t hi, hipStream_t cs,
                        hipEvent_t done = nullptr) -> hipError_t {
    hipError_t e = hipSuccess;
    if (hi > lo && r < (int)B->zr.size() && B->zr[r].packed) {
      const blpg_batch::ZRange& z = B->zr[r];
      const int64_t nw = (z.hi - z.lo + 63) / 64;
      const size_t o_pref = 8 * (size_t)nw, o_vals = o_pref + ((4 * (size_t)nw + 7) & ~(size_t)7);
      e = hipMemcpyAsync(B->d_z + z.off, B->h_z + z.off, z.bytes, hipMemcpyHostToDevice, cs);
      if (e == hipSuccess && done) e = hipEventRecord(done, cs);
      if (e == hipSuccess)
        e = launch_expand_f64((const uint64_t*)(B->d_z + z.off),
                              (const int32_t*)(B->d_z + z.off + o_pref),
                              (const double*)(B->d_z + z.off + o_vals), z.hi - z.lo,
                              B->d_pool + z.lo, cs);
      return e;
    }
    if (hi > lo)
      e = hipMemcpyAsync(B->d_pool + lo, hpool + lo, sizeof(double) * (hi - lo),
                         hipMemcpyHostToDevice, cs);
    if (e == hipSuccess && done) e = hipEventRecord(done, cs);
    return e;
  };
  if (copy) {
    HIPCHK(ctx, copy_range(0, 0, B->common_elems, st));
    if (B->spool_elems > 0)
      HIPCHK(ctx, hipMemcpyAsync(B->d_spool, B->h_stage + B->o_spool, (size_t)B->spool_elems,
                                 hipMemcpyHostToDevice, st));
    if (B->model)
      for (size_t i = 0; i < B->bins.size(); ++i)
        HIPCHK(ctx, copy_range(1 + (int)i, B->bins[i].pool_lo, B->bins[i].pool_hi, st));
  }
  const bool bin_copies = copy && !B->model;
  int prev_copy = -1;
  // learned warm start: crash lists of the family LPs, before any solve kernel
  if (timing) HIPCHK(ctx, hipEventRecord(B->ev[kEvWsStart], st));
  if (B->model) {
    const int rc = ws_run_batch(ctx, B, st);
    if (rc) return rc;
  }
  if (timing) HIPCHK(ctx, hipEventRecord(B->ev[kEvWsEnd], st));
  // ---- fork: each size bin runs solve -> check -> fallback on one of kAux+1 streams;
  //      bins are placed longest first on the least loaded stream (cost ~ LPs * R * C) ----
  HIPCHK(ctx, hipEventRecord(ctx->fork, st));
  constexpr int kStreams = blpg_ctx::kAux + 1;
  std::vector<int> order(nb), lane_of(nb, 0);
  for (int i = 0; i < nb; ++i) order[i] = i;
  auto cost = [&](int i) {
    return (double)B->bins[i].ids.size() * B->bins[i].rcap * B->bins[i].ccap;
  };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
  double load[kStreams] = {};
  for (int i : order) {
    int best = 0;
    for (int l = 1; l < kStreams; ++l)
      if (load[l] < load[best]) best = l;
    lane_of[i] = best;
    load[best] += cost(i);
  }
  std::vector<int64_t> first(nb, 0);
  for (int i = 1; i < nb; ++i) first[i] = first[i - 1] + (int64_t)B->bins[i - 1].ids.size();
  bool used[kStreams] = {};
  // enqueued largest first: persistent grids take the CUs in launch order, so the bin that
  // sets the critical path starts at once and the small bins fill the CUs it frees
  for (int i : order) {
    const Bin& b = B->bins[i];
    const int32_t cnt = (int32_t)b.ids.size();
    const int64_t pos = first[i];
    int32_t* qb = q + 16 * i;  // [0] solve queue [2] fail count [3] fallback queue
                               // [4] re-check queue [5] [6] deferred-check counts
                               // [8 + u] queue of check tier u
    hipStream_t s = lane_of[i] == 0 ? st : ctx->aux[lane_of[i] - 1];
    if (s != st && !used[lane_of[i]]) HIPCHK(ctx, hipStreamWaitEvent(s, ctx->fork, 0));
    used[lane_of[i]] = true;
    if (bin_copies) {
      if (prev_copy >= 0) HIPCHK(ctx, hipStreamWaitEvent(s, B->copied[prev_copy], 0));
      HIPCHK(ctx, copy_range(1 + i, b.pool_lo, b.pool_hi, s, B->copied[i]));
      prev_copy = i;
    }
    if (timing) HIPCHK(ctx, hipEventRecord(B->ev[1 + 4 * i], s));
    // arena batches (serving) take their LP count from the device, so one captured graph
    // serves every batch size up to the capacity
    int rc = launch_bin(ctx, B, b, plans[i], precision, b.d_idx, cnt, B->d_count, qb,
                        (char*)ctx->ws.p + ws_off[i], B->d_warm, s);
    if (rc) return rc;
    if (timing) HIPCHK(ctx, hipEventRecord(B->ev[2 + 4 * i], s));
    if (do_verify) {
      VerifyLaunch V{};
      V.desc = B->d_desc; V.pool = B->d_pool; V.spool = B->d_spool;
      V.idx = b.d_idx; V.count = cnt; V.count_dev = B->d_count;
      V.out = B->out; V.chk = B->chk; V.tol = kkt;
      V.fp32_source = fp32;
      V.mcap = b.rcap - 2;
      V.ncap = b.ccap - 2;
      V.ws = (char*)ctx->vws.p + vws_off[i];
      V.fail_list = B->d_fail + pos;
      V.fail_count = qb + 2;
      // tiers in sequence: the first takes the bin's LPs, each later one the LPs the
      // previous deferred (lists alternate, counts read on the device)
      const std::vector<int>& tv = P.vtiers[i];
      const int ntier = (int)tv.size() / 2;
      for (int u = 0; u < ntier; ++u) {
        VerifyLaunch T = V;
        T.tier = tv[2 * u];
        T.grid = tv[2 * u + 1];
        T.queue = qb + 8 + u;
        if (u > 0) {
          T.idx = ((u - 1) % 2 == 0 ? B->d_defer : B->d_defer2) + pos;
          T.count = -1;
          T.count_dev = qb + ((u - 1) % 2 == 0 ? 5 : 6);
        }
        if (u + 1 < ntier) {
          T.defer_list = (u % 2 == 0 ? B->d_defer : B->d_defer2) + pos;
          T.defer_count = qb + (u % 2 == 0 ? 5 : 6);
        }
        HIPCHK(ctx, launch_verify(T, s));
      }
      if (timing) HIPCHK(ctx, hipEventRecord(B->ev[3 + 4 * i], s));
      if (fp32) {
        // FP64 re-solve of the LPs of this bin whose FP32 result failed; the count is
        // read on the device, so no host synchronisation is needed.
        rc = launch_bin(ctx, B, fbins[i], fplans[i], BLPG_FP64, B->d_fail + pos, -1, qb + 2,
                        qb + 3, (char*)ctx->ws2.p + fws_off[i], B->out.basis, s);
        if (rc) return rc;
        VerifyLaunch V2 = V;
        V2.tier = 2;
        // (the same grid as a tier-2 pass over the bin, which sized the HBM vectors)
        V2.grid = persistent_grid(ctx, verify_max_blocks_per_cu(2, V.mcap, V.ncap), cnt);
        V2.idx = B->d_fail + pos;
        V2.count = -1;
        V2.count_dev = qb + 2;
        V2.fp32_source = 0;
        V2.queue = qb + 4;
        V2.fail_list = nullptr;
        V2.fail_count = nullptr;
        HIPCHK(ctx, launch_verify(V2, s));
      }
      if (timing) HIPCHK(ctx, hipEventRecord(B->ev[4 + 4 * i], s));
    }
    if (s != st) {
      HIPCHK(ctx, hipEventRecord(B->join[i], s));
      HIPCHK(ctx, hipStreamWaitEvent(st, B->join[i], 0));
    }
  }
  if (timing) {
    HIPCHK(ctx, hipEventRecord(B->ev[kEvEnd], st));
    B->nev = kEvEnd + 1;
  }
  B->timed = timing;
  return BLPG_OK;
}

int blpg_batch_solve(blpg_ctx* ctx, blpg_batch* B) {
  if (!ctx || !B) return BLPG_E_INVALID;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (B->count == 0) return BLPG_OK;
  // After a repack the solve starts with the inputs' host-to-device copies, enqueued
  // eagerly: copies recorded into a hipGraph run as graph memcpy nodes, which took 7.7 ms
  // where the same copies enqueued on the bins' streams took 4.3 ms (203 MB, overlapped
  // with the earlier bins' kernels).
  const int g = 0;
  if (!ctx->opt.use_graph || B->dirty) {
    const int rc = enqueue_solve(ctx, B, true, B->dirty);
    if (!rc) B->dirty = false;
    return rc;
  }
  // hipGraph path: the whole sequence (clears, warm-start network, every bin's solve /
  // check / fallback on its stream, joins) is captured once per batch and replayed; it is
  // captured again when options, the attached model or the context's workspaces change.
  blpg_options key = ctx->opt;
  key.use_graph = 1;
  const bool stale = !B->gexec[g] || std::memcmp(&key, &B->g_opt[g], sizeof key) != 0 ||
                     B->g_model[g] != B->model || B->g_gen[g] != ctx->gen;
  if (stale) {
    drop_graph(B, g);
    int rc = plan_workspaces(ctx, B);  // allocations happen outside the capture
    if (rc) return rc;
    HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_solve(ctx, B, false, false);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &gr);
    if (rc) { if (gr) (void)hipGraphDestroy(gr); return rc; }
    if (e != hipSuccess || !gr) return fail(ctx, BLPG_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
    B->graph[g] = gr;
    if (hipGraphInstantiate(&B->gexec[g], gr, nullptr, nullptr, 0) != hipSuccess) {
      drop_graph(B, g);
      return fail(ctx, BLPG_E_HIP, "hipGraphInstantiate failed");
    }
    B->g_opt[g] = key;
    B->g_model[g] = B->model;
    B->g_gen[g] = ctx->gen;
  }
  HIPCHK(ctx, hipEventRecord(B->ev[0], ctx->stream));
  HIPCHK(ctx, hipGraphLaunch(B->gexec[g], ctx->stream));
  HIPCHK(ctx, hipEventRecord(B->ev[kEvEnd], ctx->stream));
  B->nev = kEvEnd + 1;
  B->timed = false;
  return BLPG_OK;
}

// New LPs of the same shapes into an uploaded batch's pinned staging (host work only, the
// reference's setConstratint, simplex.cu:463-550); the next blpg_batch_solve copies them to
// the device before its kernels.  Arrays the upload stored once must stay shared the same
// way (two LPs may share a new array only where they shared one before).
int blpg_batch_repack(blpg_ctx* ctx, blpg_batch* B, const blpg_lp* lps, int64_t count) {
  if (!ctx || !B || count != B->count || (count > 0 && !lps))
    return fail(ctx, BLPG_E_INVALID, "blpg_batch_repack: the batch holds %lld LPs",
                B ? (long long)B->count : 0LL);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::map<int64_t, SrcKey> owner;  // pool offset -> the array staged there
  std::map<int64_t, SrcKey> sowner;
  for (int64_t k = 0; k < count; ++k) {
    const blpg_lp& L = lps[k];
    const LpDesc& d = B->desc[k];
    if (L.m != d.m || L.n != d.n || !L.A || !L.b || !L.c || (L.sense != nullptr) != (d.s >= 0) ||
        (L.warm_basis != nullptr && !B->d_warm))
      return fail(ctx, BLPG_E_INVALID, "blpg_batch_repack: LP %lld differs in shape, senses or "
                  "warm list from the uploaded batch", (long long)k);
    const SrcKey ka(L.A, (int64_t)L.m * L.n), kb(L.b, L.m), kc(L.c, L.n);
    for (const auto& ok : {std::make_pair(d.a, ka), std::make_pair(d.b, kb), std::make_pair(d.c, kc)}) {
      auto it = owner.find(ok.first);
      if (it == owner.end()) {
        if (!finite_all((const double*)ok.second.first, ok.second.second))
          return fail(ctx, BLPG_E_INVALID, "LP %lld: A, b or c has NaN/Inf", (long long)k);
        owner[ok.first] = ok.second;
      } else if (it->second != ok.second) {
        return fail(ctx, BLPG_E_INVALID, "blpg_batch_repack: LP %lld shares an array the "
                    "uploaded batch stored once with a different array", (long long)k);
      }
    }
    if (L.sense) {
      for (int i = 0; i < L.m; ++i)
        if (L.sense[i] < 0 || L.sense[i] > 2)
          return fail(ctx, BLPG_E_INVALID, "LP %lld: sense[%d] not in {0,1,2}", (long long)k, i);
      const SrcKey ks(L.sense, L.m);
      auto it = sowner.find(d.s);
      if (it == sowner.end()) sowner[d.s] = ks;
      else if (it->second != ks)
        return fail(ctx, BLPG_E_INVALID, "blpg_batch_repack: LP %lld shares a sense array the "
                    "uploaded batch stored once with a different one", (long long)k);
    }
  }
  // the previous solve's copies read the staging: wait for them
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  stage_inputs(B, lps);
  pack_ranges(B);
  if (B->d_warm) {
    const hipError_t e = upload_warm(B, lps);
    if (e != hipSuccess) return fail(ctx, BLPG_E_HIP, "warm lists: %s", hipGetErrorString(e));
  }
  B->dirty = true;
  return BLPG_OK;
}

int blpg_batch_sync(blpg_ctx* ctx) {
  if (!ctx) return BLPG_E_INVALID;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return BLPG_OK;
}

int blpg_batch_info_get(blpg_ctx* ctx, const blpg_batch* B, blpg_batch_info* info) {
  if (!ctx || !B || !info) return BLPG_E_INVALID;
  *info = B->info;
  const int nb = (int)B->bins.size();
  const bool checked = ctx->opt.verify || B->precision == BLPG_FP32;
  float ms = 0;
  info->verify_ms = info->fallback_ms = 0;
  info->warmstart_ms = 0;
  info->total_ms = 0;
  if (B->nev > 0 && hipEventElapsedTime(&ms, B->ev[0], B->ev[kEvEnd]) == hipSuccess)
    info->total_ms = ms;
  if (!B->timed) {  // graph replay: only the whole solve is timed
    for (int i = 0; i < nb; ++i) info->kernel_ms[i] = -1.0;
    info->verify_ms = info->fallback_ms = info->warmstart_ms = -1.0;
    return BLPG_OK;
  }
  for (int i = 0; i < nb; ++i) {
    if (hipEventElapsedTime(&ms, B->ev[1 + 4 * i], B->ev[2 + 4 * i]) == hipSuccess)
      info->kernel_ms[i] = ms;
    if (checked) {
      if (hipEventElapsedTime(&ms, B->ev[2 + 4 * i], B->ev[3 + 4 * i]) == hipSuccess)
        info->verify_ms = std::max(info->verify_ms, (double)ms);
      if (hipEventElapsedTime(&ms, B->ev[3 + 4 * i], B->ev[4 + 4 * i]) == hipSuccess)
        info->fallback_ms = std::max(info->fallback_ms, (double)ms);
    }
  }
  if (B->nev > 0 && B->model &&
      hipEventElapsedTime(&ms, B->ev[kEvWsStart], B->ev[kEvWsEnd]) == hipSuccess)
    info->warmstart_ms = ms;
  return BLPG_OK;
}

int blpg_batch_fetch(blpg_ctx* ctx, blpg_batch* B, blpg_batch_results* R) {
  if (!ctx || !B || !R) return BLPG_E_INVALID;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int64_t n = B->count;
  std::memset(R, 0, sizeof *R);
  R->count = n;
  if (n == 0) return BLPG_OK;
  if (!B->h_res && hipHostMalloc((void**)&B->h_res, B->res_bytes, hipHostMallocDefault) != hipSuccess) {
    B->h_res = nullptr;
    return fail(ctx, BLPG_E_NOMEM, "hipHostMalloc(%zu) for results failed", B->res_bytes);
  }
  HIPCHK(ctx, hipMemcpyAsync(B->h_res, (char*)B->mem + B->res_off, B->res_bytes,
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  // the host span mirrors the device carve: int32 block, FP64 block, x pool, basis pool
  int32_t* i32 = (int32_t*)B->h_res;
  double* f64 = (double*)(B->h_res + ((char*)B->out.obj - (char*)B->out.status));
  const double* x = (const double*)(B->h_res + ((char*)B->out.x - (char*)B->out.status));
  const int32_t* basis = (const int32_t*)(B->h_res + ((char*)B->out.basis - (char*)B->out.status));
  const bool verified = ctx->opt.verify || B->precision == BLPG_FP32;
  int64_t fbc = 0, piv = 0;
  for (int64_t k = 0; k < n; ++k) {
    int32_t& status = i32[k];
    const int ok = i32[4 * n + k];
    fbc += i32[5 * n + k];
    piv += i32[n + k];
    // objective: the FP64 re-solve's when verified, else the tableau value
    if (verified && ok == 1) f64[k] = f64[2 * n + k];
    if (!verified) { f64[3 * n + k] = -1.0; f64[4 * n + k] = -1.0; }
    if (B->precision == BLPG_FP32 && verified && ok == 0 && status == BLPG_OPTIMAL)
      status = BLPG_NUMERIC;
  }
  B->info.fallback_count = fbc;
  B->info.total_pivots = piv;
  R->status = i32;
  R->pivots = i32 + n;
  R->phase1_pivots = i32 + 2 * n;
  R->fallback_used = i32 + 5 * n;
  R->objective = f64;
  R->kkt_primal = f64 + 3 * n;
  R->kkt_dual = f64 + 4 * n;
  R->x = x;
  R->basis = basis;
  R->x_offset = B->x_off.data();
  R->basis_offset = B->basis_off.data();
  return BLPG_OK;
}

}  // extern "C"

// Results of slots 0..k-1 only (arena batches): the same host layout as blpg_batch_fetch,
// filled with one small copy per array instead of the whole span.
int blpg::batch_fetch_prefix(blpg_ctx* ctx, blpg_batch* B, int64_t k, blpg_batch_results* R) {
  if (!ctx || !B || !R || k < 0 || k > B->count) return BLPG_E_INVALID;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int64_t n = B->count;
  std::memset(R, 0, sizeof *R);
  R->count = k;
  if (!B->h_res && hipHostMalloc((void**)&B->h_res, B->res_bytes, hipHostMallocDefault) != hipSuccess) {
    B->h_res = nullptr;
    return fail(ctx, BLPG_E_NOMEM, "hipHostMalloc(%zu) for results failed", B->res_bytes);
  }
  hipStream_t st = ctx->stream;
  char* h = B->h_res;
  char* d = (char*)B->mem + B->res_off;
  auto cp = [&](const void* dptr, size_t bytes) -> hipError_t {
    const size_t off = (const char*)dptr - d;
    return bytes ? hipMemcpyAsync(h + off, dptr, bytes, hipMemcpyDeviceToHost, st) : hipSuccess;
  };
  if (k > 0) {
    const LpDesc& last = B->desc[k - 1];
    for (int a = 0; a < 6; ++a) HIPCHK(ctx, cp(B->out.status + a * n, sizeof(int32_t) * k));
    for (int a = 0; a < 5; ++a) HIPCHK(ctx, cp(B->out.obj + a * n, sizeof(double) * k));
    HIPCHK(ctx, cp(B->out.x, sizeof(double) * (last.x + last.n)));
    HIPCHK(ctx, cp(B->out.basis, sizeof(int32_t) * (last.basis + last.m)));
  }
  HIPCHK(ctx, hipStreamSynchronize(st));
  int32_t* i32 = (int32_t*)h;
  double* f64 = (double*)(h + ((char*)B->out.obj - (char*)B->out.status));
  const bool verified = ctx->opt.verify || B->precision == BLPG_FP32;
  for (int64_t q = 0; q < k; ++q) {
    const int ok = i32[4 * n + q];
    if (verified && ok == 1) f64[q] = f64[2 * n + q];
    if (!verified) { f64[3 * n + q] = -1.0; f64[4 * n + q] = -1.0; }
    if (B->precision == BLPG_FP32 && verified && ok == 0 && i32[q] == BLPG_OPTIMAL) i32[q] = BLPG_NUMERIC;
  }
  R->status = i32;
  R->pivots = i32 + n;
  R->phase1_pivots = i32 + 2 * n;
  R->fallback_used = i32 + 5 * n;
  R->objective = f64;
  R->kkt_primal = f64 + 3 * n;
  R->kkt_dual = f64 + 4 * n;
  R->x = (const double*)(h + ((char*)B->out.x - (char*)B->out.status));
  R->basis = (const int32_t*)(h + ((char*)B->out.basis - (char*)B->out.status));
  R->x_offset = B->x_off.data();
  R->basis_offset = B->basis_off.data();
  return BLPG_OK;
}

extern "C" {

int blpg_batch_download(blpg_ctx* ctx, blpg_batch* B, blpg_result* out) {
  if (!ctx || !B || (B->count > 0 && !out)) return BLPG_E_INVALID;
  blpg_batch_results R;
  const int rc = blpg_batch_fetch(ctx, B, &R);
  if (rc) return rc;
  for (int64_t k = 0; k < B->count; ++k) {
    blpg_result& r = out[k];
    const LpDesc& d = B->desc[k];
    r.status = R.status[k];
    r.pivots = R.pivots[k];
    r.phase1_pivots = R.phase1_pivots[k];
    r.fallback_used = R.fallback_used[k];
    r.objective = R.objective[k];
    r.kkt_primal = R.kkt_primal[k];
    r.kkt_dual = R.kkt_dual[k];
    if (r.x) std::memcpy(r.x, R.x + d.x, sizeof(double) * d.n);
    if (r.basis) std::memcpy(r.basis, R.basis + d.basis, sizeof(int32_t) * d.m);
  }
  return BLPG_OK;
}

}  // extern "C"

// Solves lps[0..count) as one device batch; a batch that exceeds the HBM budget or the
// device's free memory is split in halves, recursively, and the parts solved one after
// the other (out-of-core batching: no device reset, results in place).  A single LP that
// does not fit still fails with BLPG_E_BUDGET / BLPG_E_NOMEM.
static int solve_range(blpg_ctx* ctx, const blpg_lp* lps, int64_t count, blpg_result* out) {
  blpg_batch* B = nullptr;
  int rc = blpg_batch_upload(ctx, lps, count, &B);
  if ((rc == BLPG_E_BUDGET || rc == BLPG_E_NOMEM) && count > 1) {
    const int64_t h = count / 2;
    rc = solve_range(ctx, lps, h, out);
    if (!rc) rc = solve_range(ctx, lps + h, count - h, out + h);
    return rc;
  }
  if (rc) return rc;
  rc = blpg_batch_solve(ctx, B);
  if (!rc) rc = blpg_batch_sync(ctx);
  if (!rc) rc = blpg_batch_download(ctx, B, out);
  blpg_batch_free(ctx, B);
  return rc;
}

extern "C" {

int blpg_solve_batch(blpg_ctx* ctx, const blpg_lp* lps, int64_t count, blpg_result* out) {
  if (!ctx) return BLPG_E_INVALID;
  if (count == 0) return BLPG_OK;
  return solve_range(ctx, lps, count, out);
}

int blpg_tune_set(blpg_ctx* ctx, int32_t precision, int32_t m, int32_t n, int32_t kernel,
                  int32_t param) {
  if (!ctx || precision < 0 || precision > 1 || m < 1 || n < 1 || kernel < BLPG_KERNEL_STREAM ||
      kernel > BLPG_KERNEL_WAVE)
    return fail(ctx, BLPG_E_INVALID, "invalid arguments to blpg_tune_set");
  if (kernel == BLPG_KERNEL_TILE &&
      (param < 0 || param >= tile_variant_count(precision) || !Tuner::fits_tile(param, m, n, precision)))
    return fail(ctx, BLPG_E_INVALID, "tile variant %d cannot hold %dx%d", param, m, n);
  if (kernel == BLPG_KERNEL_WAVE &&
      (param < 0 || param >= wave_variant_count() || !Tuner::fits_wave(param, m, n, precision)))
    return fail(ctx, BLPG_E_INVALID, "wave variant %d cannot hold %dx%d", param, m, n);
  if (kernel != BLPG_KERNEL_TILE && kernel != BLPG_KERNEL_WAVE && param != 64 && param != 128 &&
      param != 256 && param != 512)
    return fail(ctx, BLPG_E_INVALID, "threads per workgroup must be 64, 128, 256 or 512");
  ctx->tuner.set_preference(precision, m, n, kernel, param);
  return BLPG_OK;
}

int blpg_tune_export(blpg_ctx* ctx, char* buf, int64_t cap, int64_t* len) {
  if (!ctx) return BLPG_E_INVALID;
  const std::string t = ctx->tuner.export_text();
  if (len) *len = (int64_t)t.size();
  if (buf && cap > 0) {
    const size_t k = std::min<size_t>(t.size(), (size_t)cap - 1);
    std::memcpy(buf, t.data(), k);
    buf[k] = 0;
  }
  return BLPG_OK;
}

int blpg_tune_import(blpg_ctx* ctx, const char* text) {
  if (!ctx || !text) return BLPG_E_INVALID;
  if (!ctx->tuner.import_text(text)) return fail(ctx, BLPG_E_INVALID, "malformed tuning table");
  return BLPG_OK;
}

// The check plan of one shape: when its check starts with the first tier (31 KB per LP) and
// has a later one, time the check (eager, HIP events) with and without the first tier on
// the shape's LPs and keep the faster; the default (first tier used) stays unless skipping
// it is more than 3% faster.  Workloads whose final bases never fit the first tier pay a
// pass that only hands every LP on.
static int tune_check(blpg_ctx* ctx, std::vector<blpg_lp>& lps, int prec, int m, int n,
                      int reps) {
  const int tsz = prec == BLPG_FP32 ? 4 : 8;
  ctx->tuner.erase_check(prec, m, n);
  if (!verify_tier_used(0, m, n, tsz) || verify_tier_covers(0, m, n, tsz)) return BLPG_OK;
  double t[2] = {1e30, 1e30};
  for (int skip = 0; skip < 2; ++skip) {
    ctx->tuner.set_check(prec, m, n, skip);
    blpg_batch* B = nullptr;
    int rc = blpg_batch_upload(ctx, lps.data(), (int64_t)lps.size(), &B);
    for (int r = 0; r <= reps && !rc; ++r) {
      rc = blpg_batch_solve(ctx, B);
      if (!rc) rc = blpg_batch_sync(ctx);
      blpg_batch_info info;
      if (!rc) rc = blpg_batch_info_get(ctx, B, &info);
      if (!rc && r > 0 && info.verify_ms > 0) t[skip] = std::min(t[skip], info.verify_ms);
    }
    if (B) blpg_batch_free(ctx, B);
    if (rc) {
      ctx->tuner.erase_check(prec, m, n);
      return rc;
    }
  }
  const int skip = t[1] < 0.97 * t[0] ? 1 : 0;
  if (t[skip] < 1e30) ctx->tuner.set_check(prec, m, n, skip, t[skip]);
  else ctx->tuner.erase_check(prec, m, n);
  return BLPG_OK;
}

int blpg_autotune(blpg_ctx* ctx, const blpg_lp* lps, int64_t count, int32_t reps) {
  if (!ctx || (count > 0 && !lps) || count < 0)
    return fail(ctx, BLPG_E_INVALID, "invalid arguments to blpg_autotune");
  if (reps < 1) reps = 3;
  const int prec = ctx->opt.precision;
  const blpg_options saved = ctx->opt;
  ctx->opt.kernel = BLPG_KERNEL_AUTO;
  ctx->opt.use_graph = 0;  // per-bin kernel times come from the timed (eager) path
  std::map<std::pair<int, int>, std::vector<blpg_lp>> groups;
  for (int64_t k = 0; k < count; ++k) {
    blpg_lp L = lps[k];
    L.warm_basis = nullptr;  // time cold solves
    groups[{L.m, L.n}].push_back(L);
  }
  int rc = BLPG_OK;
  for (auto& g : groups) {
    const int m = g.first.first, n = g.first.second;
    // candidates: every instance that can hold the shape and be resident
    std::vector<std::pair<int, int>> cand;
    for (int v = 0; v < wave_variant_count(); ++v)
      if (Tuner::fits_wave(v, m, n, prec) && wave_max_blocks_per_cu(prec, v) > 0)
        cand.push_back({BLPG_KERNEL_WAVE, v});
    for (int v = 0; v < tile_variant_count(prec); ++v)
      if (Tuner::fits_tile(v, m, n, prec) && tile_max_blocks_per_cu(prec, v) > 0)
        cand.push_back({BLPG_KERNEL_TILE, v});
    for (int nt : {64, 128, 256, 512}) {
      if (stream_lds_bytes(prec, true, nt, m + 2, n + 2) <= ctx->tuner.lds_limit() &&
          stream_max_blocks_per_cu(prec, true, nt, m + 2, n + 2) > 0)
        cand.push_back({BLPG_KERNEL_LDS, nt});
      if (stream_max_blocks_per_cu(prec, false, nt, m + 2, n + 2) > 0)
        cand.push_back({BLPG_KERNEL_STREAM, nt});
    }
    // the static route is timed first and kept unless another instance is more than 3%
    // faster: closer timings are run-to-run noise, and a choice that flips between runs
    // moves shapes between bins (and the batch's critical path with them)
    const Route r0 = ctx->tuner.route(m, n, prec, BLPG_KERNEL_AUTO);
    const std::pair<int, int> c0{r0.kernel, (r0.kernel == BLPG_KERNEL_TILE || r0.kernel == BLPG_KERNEL_WAVE)
                                                ? r0.variant : r0.nt};
    std::stable_partition(cand.begin(), cand.end(),
                          [&](const std::pair<int, int>& c) { return c == c0; });
    double best_ms = 1e30;
    std::pair<int, int> best{-1, -1};
    // the shape's route before the measurement, restored if a candidate fails
    int pk = 0, pp = 0;
    double pms = -1.0;
    const bool had = ctx->tuner.get_preference(prec, m, n, pk, pp, pms);
    for (const auto& c : cand) {
      ctx->tuner.set_preference(prec, m, n, c.first, c.second);
      blpg_batch* B = nullptr;
      rc = blpg_batch_upload(ctx, g.second.data(), (int64_t)g.second.size(), &B);
      if (rc) break;
      double t = 1e30;
      for (int r = 0; r <= reps && !rc; ++r) {
        rc = blpg_batch_solve(ctx, B);
        if (!rc) rc = blpg_batch_sync(ctx);
        blpg_batch_info info;
        if (!rc) rc = blpg_batch_info_get(ctx, B, &info);
        if (!rc && r > 0 && info.nbins == 1) t = std::min(t, info.kernel_ms[0]);
      }
      blpg_batch_free(ctx, B);
      if (rc) break;
      const bool incumbent_static = best == c0;
      if (t < best_ms * (incumbent_static ? 0.97 : 1.0)) { best_ms = t; best = c; }
    }
    if (rc) {
      if (had) ctx->tuner.set_preference(prec, m, n, pk, pp, pms);
      else ctx->tuner.erase_preference(prec, m, n);
      break;
    }
    if (best.first >= 0) ctx->tuner.set_preference(prec, m, n, best.first, best.second, best_ms);
    if (ctx->opt.verify || prec == BLPG_FP32) {
      rc = tune_check(ctx, g.second, prec, m, n, reps);
      if (rc) break;
    }
  }
  ctx->opt = saved;
  return rc;
}

int blpg_solve_shared_A(blpg_ctx* ctx, int32_t m, int32_t n, const double* A, const double* b,
                        const int8_t* sense, const double* C, int64_t count, blpg_result* out) {
  if (!ctx || !A || !b || !C || count < 0 || m < 1 || n < 1)
    return fail(ctx, BLPG_E_INVALID, "invalid arguments to blpg_solve_shared_A");
  std::vector<blpg_lp> lps(count);
  for (int64_t k = 0; k < count; ++k) {
    lps[k].m = m;
    lps[k].n = n;
    lps[k].A = A;
    lps[k].b = b;
    lps[k].c = C + k * (int64_t)n;
    lps[k].sense = sense;
    lps[k].family = -1;
    lps[k].warm_basis = nullptr;
  }
  if (!ctx->opt.sibling_warm || count < 2) return blpg_solve_batch(ctx, lps.data(), count, out);
  // sibling basis: LP 0 first; a basis it ends phase II on is primal feasible for all
  std::vector<int32_t> pilot((size_t)m);
  blpg_result r0 = out[0];
  r0.basis = pilot.data();
  int rc = blpg_solve_batch(ctx, lps.data(), 1, &r0);
  if (rc) return rc;
  if (out[0].basis) std::memcpy(out[0].basis, pilot.data(), sizeof(int32_t) * (size_t)m);
  r0.basis = out[0].basis;
  out[0] = r0;
  if (r0.status == BLPG_OPTIMAL || r0.status == BLPG_UNBOUNDED)
    for (int64_t k = 1; k < count; ++k) lps[k].warm_basis = pilot.data();
  return blpg_solve_batch(ctx, lps.data() + 1, count - 1, out + 1);
}

}  // extern "C"
