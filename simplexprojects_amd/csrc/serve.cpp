// serve.cpp -- serving loop: single-LP requests, device-resident arenas, batches formed
// from a queue and launched through captured hipGraphs.  No reference counterpart (the
// reference solves one batch per ComputeLP call, GPU_Tester.cu:244-259).
//
// Per LP shape (m, n, with / without senses) the server keeps `arenas` device batches of
// `max_batch` slots each (fixed pool offsets per slot, a device-side LP count), each with
// its own context and stream, and a pinned mirror of its input pool.  A request is copied
// once, by the submitting thread, straight into the next free slot of its shape's filling
// arena (the slot is reserved under the lock, the copy runs outside it).  When no arena
// can take it (all busy or full) it waits in the shape's overflow queue, which the worker
// stages into the next free arena first, so requests launch in arrival order.  The worker
//   - completes finished arenas (hipStreamQuery), copying only the used slots' results;
//   - launches the filling arena when it is full, when its oldest request has waited
//     max_wait_us, or when no arena of that shape is busy, once every reserved slot has
//     been copied: one H2D per pool, the count to the device, the arena's graph replayed.
// Admission: arena bytes + bytes of admitted, unpolled requests <= hbm_budget.
// Learned warm start: with a model attached (blpg_server_set_warmstart), requests of the
// model's shape tagged with one of its families form shapes of their own, whose arenas run
// the model inside their graph; before each launch the worker copies the arena's row tables
// (which slot is which family) next to the inputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "ctx.hpp"

using namespace blpg;
using Clock = std::chrono::steady_clock;

namespace {

// what completion needs of a request
struct Meta {
  int64_t ticket = 0;
  int m = 0, n = 0;
  int family = -1;  // >= 0: tagged for the attached warm-start model
  double* x = nullptr;
  int32_t* basis = nullptr;
  Clock::time_point t0;
  int64_t bytes = 0;
};

// a request waiting in the overflow queue (its data copied to the heap)
struct Req {
  Meta meta;
  std::vector<double> data;  // A (m*n) | b (m) | c (n)
  std::vector<int8_t> sense;
};

struct Arena {
  blpg_ctx* ctx = nullptr;
  blpg_batch* B = nullptr;
  double* h_pool = nullptr;  // pinned mirror of the device pool
  int8_t* h_sense = nullptr; // pinned mirror of the sense pool
  int32_t* h_count = nullptr;
  int32_t* d_count = nullptr;
  void* h_ws = nullptr;      // pinned image of the warm-start row tables (tagged arenas)
  int64_t pool_elems = 0, spool_elems = 0;
  std::vector<Meta> filled;     // reserved slots 0..filled.size()-1 (not launched yet)
  std::atomic<int> copied{0};   // reserved slots whose data is in h_pool
  std::vector<Meta> inflight;
  bool busy = false;
  bool failed = false;  // the launch failed: its requests complete as undecided
  size_t bytes = 0;  // device bytes
};

struct Shape {
  int m = 0, n = 0;
  bool sense = false;
  bool tagged = false;    // family-tagged requests for the warm-start model
  int fill = -1;          // arena taking new requests (-1: none chosen)
  std::deque<Req> queue;  // overflow, in arrival order
  std::vector<std::unique_ptr<Arena>> arenas;
};

}  // namespace

struct blpg_server {
  blpg_ctx* parent = nullptr;
  blpg_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<std::tuple<int, int, int, int>, std::unique_ptr<Shape>> shapes;  // m, n, sense, tagged
  blpg_ws_model* model = nullptr;  // warm-start model of tagged requests
  int ws_families = 0, ws_m = 0, ws_n = 0;
  std::deque<blpg_server_result> done;
  std::vector<double> lat;  // all completion latencies
  int64_t next_ticket = 1, submitted = 0, completed = 0, rejected = 0, batches = 0,
          batched_lps = 0;
  int64_t arena_bytes = 0, pending_bytes = 0;
  std::string err;
  std::mutex create_mu;  // serialises the creation of a new shape's arenas
  std::atomic<bool> stop{false};
  std::thread worker;
};

namespace {

constexpr int64_t kLargeTableau = 150 * 150;  // (m+2)(n+2) from which an arena is masked
constexpr int kStallSeconds = 60;             // bound of the replay / destroy waits

// The server's state in one line (for a stalled replay's error).
std::string describe(blpg_server* S) {
  std::lock_guard<std::mutex> lk(S->mu);
  std::string out = "err='" + S->err + "'";
  char buf[160];
  for (auto& kv : S->shapes) {
    const Shape& sh = *kv.second;
    std::snprintf(buf, sizeof buf, "; %dx%d%s%s fill=%d queued=%zu", sh.m, sh.n,
                  sh.sense ? " senses" : "", sh.tagged ? " tagged" : "", sh.fill,
                  sh.queue.size());
    out += buf;
    for (auto& a : sh.arenas) {
      std::snprintf(buf, sizeof buf, " [busy=%d failed=%d filled=%zu copied=%d inflight=%zu "
                    "stream=%s]", a->busy ? 1 : 0, a->failed ? 1 : 0, a->filled.size(),
                    a->copied.load(), a->inflight.size(),
                    a->busy ? hipGetErrorName(hipStreamQuery(a->ctx->stream)) : "-");
      out += buf;
    }
  }
  return out;
}

int64_t req_bytes(int m, int n, bool sense) {
  return 8 * ((int64_t)m * n + m + n) + (sense ? m : 0) + 12 * (int64_t)(m + n) + 64;
}

void destroy_arena(Arena& a) {
  if (a.B) blpg_batch_free(a.ctx, a.B);
  if (a.d_count) (void)hipFree(a.d_count);
  if (a.h_pool) (void)hipHostFree(a.h_pool);
  if (a.h_sense) (void)hipHostFree(a.h_sense);
  if (a.h_count) (void)hipHostFree(a.h_count);
  if (a.h_ws) (void)hipHostFree(a.h_ws);
  if (a.ctx) blpg_destroy(a.ctx);
  a.ctx = nullptr; a.B = nullptr; a.h_pool = nullptr; a.h_sense = nullptr;
  a.h_count = nullptr; a.d_count = nullptr; a.h_ws = nullptr;
}

// Records an error of arena creation (which runs without the server lock).
int note(blpg_server* S, const char* msg, int rc) {
  std::lock_guard<std::mutex> lk(S->mu);
  S->err = msg;
  return rc;
}

// An arena: a device batch of `cap` distinct slots of shape (m, n), uploaded once with
// zero data (the pool layout is what matters), plus its pinned staging mirrors.
int make_arena(blpg_server* S, Shape& sh, Arena& a) {
  const int cap = S->cfg.max_batch;
  blpg_options o;
  blpg_default_options(&o);
  o.precision = S->cfg.precision;
  o.device = S->parent->opt.device;
  o.use_graph = 1;
  int rc = blpg_create(&a.ctx, &o);
  if (rc) return rc;
  // Arenas of large shapes (LPs that hold a CU for milliseconds) run on three quarters of
  // the CUs, so batches of small shapes always find free CUs instead of queueing behind
  // them: a stream with a CU mask that leaves out every fourth group of 8 CUs (spread over
  // the XCDs however CU indices map to them).
  if ((int64_t)(sh.m + 2) * (sh.n + 2) >= kLargeTableau) {
    const int ncu = a.ctx->props.multiProcessorCount;
    std::vector<uint32_t> mask((ncu + 31) / 32, 0u);
    for (int cu = 0; cu < ncu; ++cu)
      if ((cu / 8) % 4 != 3) mask[cu / 32] |= 1u << (cu % 32);
    hipStream_t masked = nullptr;
    if (hipExtStreamCreateWithCUMask(&masked, (uint32_t)ncu, mask.data()) == hipSuccess) {
      (void)hipStreamDestroy(a.ctx->stream);
      a.ctx->stream = masked;
    }
  }
  const std::string tune = S->parent->tuner.export_text();
  if (!tune.empty()) (void)blpg_tune_import(a.ctx, tune.c_str());
  const int m = sh.m, n = sh.n;
  std::vector<double> zeros((size_t)cap * ((size_t)m * n + m + n), 0.0);
  std::vector<int8_t> szero(sh.sense ? (size_t)cap * m : 1, 0);
  std::vector<blpg_lp> lps(cap);
  for (int s = 0; s < cap; ++s) {
    double* base = zeros.data() + (size_t)s * ((size_t)m * n + m + n);
    lps[s].m = m;
    lps[s].n = n;
    lps[s].A = base;
    lps[s].b = base + (size_t)m * n;
    lps[s].c = base + (size_t)m * n + m;
    lps[s].sense = sh.sense ? szero.data() + (size_t)s * m : nullptr;
    lps[s].family = -1;
    lps[s].warm_basis = nullptr;
  }
  rc = blpg_batch_upload(a.ctx, lps.data(), cap, &a.B);
  if (rc) return note(S, blpg_last_error(a.ctx), rc);
  a.pool_elems = a.B->pool_elems;
  a.spool_elems = a.B->spool_elems;
  if (hipMalloc(&a.d_count, sizeof(int32_t)) != hipSuccess ||
      hipHostMalloc((void**)&a.h_count, sizeof(int32_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&a.h_pool, sizeof(double) * a.pool_elems, hipHostMallocDefault) != hipSuccess ||
      (sh.sense && hipHostMalloc((void**)&a.h_sense, std::max<int64_t>(a.spool_elems, 1),
                                 hipHostMallocDefault) != hipSuccess))
    return BLPG_E_NOMEM;
  std::memset(a.h_pool, 0, sizeof(double) * a.pool_elems);
  if (a.h_sense) std::memset(a.h_sense, 0, a.spool_elems);
  a.B->d_count = a.d_count;
  a.bytes = a.B->mem_bytes + a.B->res_bytes;
  if (sh.tagged) {
    rc = ws_arena_attach(a.ctx, a.B, S->model, cap);
    if (rc) return note(S, blpg_last_error(a.ctx), rc);
    if (hipHostMalloc(&a.h_ws, a.B->ws.table_bytes, hipHostMallocDefault) != hipSuccess)
      return BLPG_E_NOMEM;
    a.bytes += a.B->ws.bytes + sizeof(int32_t) * (size_t)a.B->basis_elems;
  }
  return BLPG_OK;
}

// Copies one LP into slot `slot` of the arena's pinned pool mirror.
void copy_slot(Arena& a, int slot, const double* A, const double* b, const double* c,
               const int8_t* sense) {
  const LpDesc& d = a.B->desc[slot];
  const size_t mn = (size_t)d.m * d.n;
  std::memcpy(a.h_pool + d.a, A, sizeof(double) * mn);
  std::memcpy(a.h_pool + d.b, b, sizeof(double) * d.m);
  std::memcpy(a.h_pool + d.c, c, sizeof(double) * d.n);
  if (a.h_sense && d.s >= 0 && sense) std::memcpy(a.h_sense + d.s, sense, d.m);
}

// Uploads slots 0..k-1 of the arena (requests `metas`) and replays its graph (asynchronous).
int fire(Arena& a, const std::vector<Meta>& metas) {
  blpg_batch* B = a.B;
  const int k = (int)metas.size();
  const LpDesc& last = B->desc[k - 1];
  const int64_t end = last.c + last.n;
  const int64_t send = (a.h_sense && last.s >= 0) ? last.s + last.m : 0;
  *a.h_count = k;
  hipStream_t st = a.ctx->stream;
  if (hipMemcpyAsync(B->d_pool, a.h_pool, sizeof(double) * end, hipMemcpyHostToDevice, st) != hipSuccess ||
      (send > 0 && hipMemcpyAsync(B->d_spool, a.h_sense, send, hipMemcpyHostToDevice, st) != hipSuccess) ||
      hipMemcpyAsync(a.d_count, a.h_count, sizeof(int32_t), hipMemcpyHostToDevice, st) != hipSuccess)
    return BLPG_E_HIP;
  if (a.h_ws) {
    std::vector<int32_t> fam(k);
    for (int s = 0; s < k; ++s) fam[s] = metas[s].family;
    ws_arena_tables(B, fam.data(), k, a.h_ws);
    if (hipMemcpyAsync(B->ws.mem, a.h_ws, B->ws.table_bytes, hipMemcpyHostToDevice, st) != hipSuccess)
      return BLPG_E_HIP;
  }
  return blpg_batch_solve(a.ctx, B);
}

// Results of a finished arena (only the used slots are copied back).
int finish(blpg_server* S, Arena& a) {
  blpg_batch_results R;
  const int64_t k = (int64_t)a.inflight.size();
  const int rc = a.failed ? BLPG_E_HIP : batch_fetch_prefix(a.ctx, a.B, k, &R);
  a.failed = false;
  const auto now = Clock::now();
  std::lock_guard<std::mutex> lk(S->mu);
  for (int64_t s = 0; s < k; ++s) {
    Meta& r = a.inflight[s];
    blpg_server_result o{};
    o.ticket = r.ticket;
    o.status = rc ? BLPG_UNDECIDED : R.status[s];
    o.pivots = rc ? 0 : R.pivots[s];
    o.objective = rc ? 0.0 : R.objective[s];
    o.latency_us = std::chrono::duration<double, std::micro>(now - r.t0).count();
    if (!rc) {
      if (r.x) std::memcpy(r.x, R.x + R.x_offset[s], sizeof(double) * r.n);
      if (r.basis) std::memcpy(r.basis, R.basis + R.basis_offset[s], sizeof(int32_t) * r.m);
    }
    S->lat.push_back(o.latency_us);
    S->done.push_back(o);
    S->pending_bytes -= r.bytes;
    ++S->completed;
  }
  a.inflight.clear();
  a.busy = false;
  S->cv_done.notify_all();
  return rc;
}

void worker_loop(blpg_server* S) {
  (void)hipSetDevice(S->parent->device);
  const auto max_wait = std::chrono::microseconds(S->cfg.max_wait_us);
  while (!S->stop.load()) {
    bool progress = false;
    // completions: a stream that reports an error (not just "not ready") completes its
    // arena too, as failed, so its requests return (undecided) instead of waiting forever
    std::vector<Arena*> ready;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      for (auto& kv : S->shapes)
        for (auto& a : kv.second->arenas) {
          if (!a->busy) continue;
          const hipError_t q = hipStreamQuery(a->ctx->stream);
          if (q == hipErrorNotReady) continue;
          if (q != hipSuccess) {
            a->failed = true;
            S->err = std::string("arena stream error: ") + hipGetErrorString(q);
          }
          ready.push_back(a.get());
        }
    }
    for (Arena* a : ready) { finish(S, *a); progress = true; }
    // launches: the filling arena of each shape, and overflow requests (staged here)
    std::vector<std::pair<Arena*, std::vector<Req>>> staged;
    std::vector<Arena*> ready_fill;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      const auto now = Clock::now();
      for (auto& kv : S->shapes) {
        Shape& sh = *kv.second;
        bool any_busy = false;
        for (auto& a : sh.arenas) any_busy = any_busy || a->busy;
        // arenas holding requests (older than any in the overflow queue): the filling one,
        // and any the filling moved on from when it became full before this loop launched
        // it (a submit picks a new empty arena then), which is launched at once
        for (size_t ai = 0; ai < sh.arenas.size(); ++ai) {
          Arena* fa = sh.arenas[ai].get();
          if (fa->busy || fa->filled.empty()) continue;
          const bool filling = (int)ai == sh.fill;
          const bool full = (int)fa->filled.size() >= S->cfg.max_batch;
          if ((full || !filling || !sh.queue.empty() || !any_busy ||
               now - fa->filled.front().t0 >= max_wait) &&
              fa->copied.load(std::memory_order_acquire) == (int)fa->filled.size()) {
            fa->busy = true;
            any_busy = true;
            fa->inflight = std::move(fa->filled);
            fa->filled.clear();
            fa->copied.store(0, std::memory_order_relaxed);
            if (filling) sh.fill = -1;
            ready_fill.push_back(fa);
          }
        }
        // overflow requests into a free, empty arena
        if (!sh.queue.empty()) {
          Arena* freea = nullptr;
          for (auto& a : sh.arenas)
            if (!a->busy && a->filled.empty() && !freea) freea = a.get();
          const bool full = (int)sh.queue.size() >= S->cfg.max_batch;
          const bool aged = now - sh.queue.front().meta.t0 >= max_wait;
          if (freea && (full || aged || !any_busy)) {
            const int k = std::min<int>((int)sh.queue.size(), S->cfg.max_batch);
            std::vector<Req> reqs;
            reqs.reserve(k);
            for (int i = 0; i < k; ++i) {
              reqs.push_back(std::move(sh.queue.front()));
              sh.queue.pop_front();
            }
            staged.push_back({freea, std::move(reqs)});
            freea->busy = true;  // reserved
          }
        }
      }
    }
    auto record = [&](Arena* a, int k, int rc) {
      std::lock_guard<std::mutex> lk(S->mu);
      if (rc) {
        S->err = "launch failed";
        a->failed = true;
      } else {
        ++S->batches;
        S->batched_lps += k;
      }
      progress = true;
    };
    for (auto& t : staged) {
      Arena* a = t.first;
      const int k = (int)t.second.size();
      std::vector<Meta> metas;
      metas.reserve(k);
      for (int s = 0; s < k; ++s) {
        const Req& r = t.second[s];
        const size_t mn = (size_t)r.meta.m * r.meta.n;
        copy_slot(*a, s, r.data.data(), r.data.data() + mn, r.data.data() + mn + r.meta.m,
                  r.sense.empty() ? nullptr : r.sense.data());
        metas.push_back(r.meta);
      }
      const int rc = fire(*a, metas);
      {
        std::lock_guard<std::mutex> lk(S->mu);
        a->inflight = std::move(metas);  // busy stays set: finished by the poll above
      }
      record(a, k, rc);
    }
    for (Arena* a : ready_fill) {
      const int k = (int)a->inflight.size();
      record(a, k, fire(*a, a->inflight));
    }
    if (!progress) {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait_for(lk, std::chrono::microseconds(20));
    }
  }
}

}  // namespace

extern "C" {

int blpg_server_create(blpg_ctx* ctx, const blpg_server_config* cfg, blpg_server** out) {
  if (!ctx || !out) return BLPG_E_INVALID;
  *out = nullptr;
  blpg_server* S = new blpg_server();
  S->parent = ctx;
  if (cfg) S->cfg = *cfg;
  if (S->cfg.hbm_budget <= 0) S->cfg.hbm_budget = 64ll << 30;
  if (S->cfg.max_batch <= 0) S->cfg.max_batch = 2048;
  if (S->cfg.max_wait_us <= 0) S->cfg.max_wait_us = 200;
  if (S->cfg.arenas <= 0) S->cfg.arenas = 2;
  if (S->cfg.precision != BLPG_FP32) S->cfg.precision = BLPG_FP64;
  S->worker = std::thread(worker_loop, S);
  *out = S;
  return BLPG_OK;
}

void blpg_server_destroy(blpg_server* S) {
  if (!S) return;
  S->stop.store(true);
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  // the arenas' work is waited for, all within one bound: an arena still running after
  // kStallSeconds is left allocated rather than freed under a running kernel
  const auto deadline = Clock::now() + std::chrono::seconds(kStallSeconds);
  for (auto& kv : S->shapes)
    for (auto& a : kv.second->arenas) {
      bool idle = !a->ctx;
      while (!idle) {
        idle = hipStreamQuery(a->ctx->stream) != hipErrorNotReady;
        if (idle || Clock::now() > deadline) break;
        std::this_thread::sleep_for(std::chrono::microseconds(100));
      }
      if (idle) destroy_arena(*a);
    }
  delete S;
}

int blpg_server_set_warmstart(blpg_server* S, blpg_ws_model* M) {
  if (!S) return BLPG_E_INVALID;
  std::lock_guard<std::mutex> ck(S->create_mu);  // no shape is being created meanwhile
  std::lock_guard<std::mutex> lk(S->mu);
  for (auto& kv : S->shapes)
    if (std::get<3>(kv.first)) {
      S->err = "blpg_server_set_warmstart: tagged requests were already admitted";
      return BLPG_E_INVALID;
    }
  S->model = M;
  S->ws_families = S->ws_m = S->ws_n = 0;
  if (M) ws_model_shape(M, &S->ws_families, &S->ws_m, &S->ws_n);
  return BLPG_OK;
}

}  // extern "C"

namespace {

int submit_at(blpg_server* S, const blpg_lp* lp, double* x, int32_t* basis, Clock::time_point t0,
              int64_t* ticket) {
  if (!S || !lp || !lp->A || !lp->b || !lp->c || lp->m < 1 || lp->n < 1 || lp->m > 4096 ||
      lp->n > 4096)
    return BLPG_E_INVALID;
  const int m = lp->m, n = lp->n;
  const size_t mn = (size_t)m * n;
  for (size_t e = 0; e < mn; ++e)
    if (!std::isfinite(lp->A[e])) return BLPG_E_INVALID;
  for (int i = 0; i < m; ++i)
    if (!std::isfinite(lp->b[i]) || (lp->sense && (lp->sense[i] < 0 || lp->sense[i] > 2)))
      return BLPG_E_INVALID;
  for (int j = 0; j < n; ++j)
    if (!std::isfinite(lp->c[j])) return BLPG_E_INVALID;
  const bool has_sense = lp->sense != nullptr;
  Meta r;
  {
    std::lock_guard<std::mutex> lk0(S->mu);
    if (S->model && lp->family >= 0 && lp->family < S->ws_families && m == S->ws_m && n == S->ws_n)
      r.family = lp->family;
  }
  const auto key = std::make_tuple(m, n, (int)has_sense, r.family >= 0 ? 1 : 0);
  r.m = m;
  r.n = n;
  r.x = x;
  r.basis = basis;
  r.t0 = t0;
  r.bytes = req_bytes(m, n, has_sense);
  bool need;
  {
    std::lock_guard<std::mutex> lk0(S->mu);
    need = S->shapes.find(key) == S->shapes.end();
  }
  std::unique_lock<std::mutex> ck(S->create_mu, std::defer_lock);
  if (need) ck.lock();  // (never while holding mu: the creator takes mu after create_mu)
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->shapes.find(key);
  if (it == S->shapes.end()) {
    // a new shape: its arenas are created now (admitted against the budget too)
    auto sh = std::make_unique<Shape>();
    sh->m = m;
    sh->n = n;
    sh->sense = has_sense;
    sh->tagged = r.family >= 0;
    const int64_t arena_need = (int64_t)S->cfg.arenas * S->cfg.max_batch * r.bytes;
    if (S->arena_bytes + S->pending_bytes + arena_need > S->cfg.hbm_budget) {
      ++S->rejected;
      return BLPG_E_BUDGET;
    }
    lk.unlock();  // arena creation allocates and uploads: do it without the lock
    for (int i = 0; i < S->cfg.arenas; ++i) {
      auto a = std::make_unique<Arena>();
      const int rc = make_arena(S, *sh, *a);
      if (rc) {
        destroy_arena(*a);
        for (auto& b : sh->arenas) destroy_arena(*b);
        return rc;
      }
      sh->arenas.push_back(std::move(a));
    }
    lk.lock();
    for (auto& a : sh->arenas) S->arena_bytes += (int64_t)a->bytes;
    it = S->shapes.emplace(key, std::move(sh)).first;
  }
  if (S->arena_bytes + S->pending_bytes + r.bytes > S->cfg.hbm_budget) {
    ++S->rejected;
    return BLPG_E_BUDGET;
  }
  r.ticket = S->next_ticket++;
  if (ticket) *ticket = r.ticket;
  S->pending_bytes += r.bytes;
  ++S->submitted;
  Shape& sh = *it->second;
  // the filling arena, or a free empty one (only while nothing waits in the overflow)
  Arena* a = nullptr;
  int slot = -1;
  if (sh.queue.empty()) {
    if (sh.fill >= 0) {
      Arena* fa = sh.arenas[sh.fill].get();
      if (!fa->busy && (int)fa->filled.size() < S->cfg.max_batch) a = fa;
    }
    if (!a)
      for (size_t i = 0; i < sh.arenas.size(); ++i)
        if (!sh.arenas[i]->busy && sh.arenas[i]->filled.empty()) {
          sh.fill = (int)i;
          a = sh.arenas[i].get();
          break;
        }
  }
  if (a) {
    slot = (int)a->filled.size();
    a->filled.push_back(r);
    lk.unlock();
    copy_slot(*a, slot, lp->A, lp->b, lp->c, lp->sense);  // outside the lock
    a->copied.fetch_add(1, std::memory_order_release);
  } else {
    Req q;
    q.meta = r;
    q.data.resize(mn + m + n);
    std::memcpy(q.data.data(), lp->A, sizeof(double) * mn);
    std::memcpy(q.data.data() + mn, lp->b, sizeof(double) * m);
    std::memcpy(q.data.data() + mn + m, lp->c, sizeof(double) * n);
    if (has_sense) q.sense.assign(lp->sense, lp->sense + m);
    sh.queue.push_back(std::move(q));
    lk.unlock();
  }
  S->cv_work.notify_one();
  return BLPG_OK;
}

}  // namespace

extern "C" {

int blpg_server_submit(blpg_server* S, const blpg_lp* lp, double* x, int32_t* basis,
                       int64_t* ticket) {
  return submit_at(S, lp, x, basis, Clock::now(), ticket);
}

int blpg_server_poll(blpg_server* S, blpg_server_result* out, int32_t max, int32_t wait_us,
                     int32_t* n) {
  if (!S || (max > 0 && !out) || !n) return BLPG_E_INVALID;
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->done.empty() && wait_us > 0)
    S->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [&] { return !S->done.empty(); });
  int32_t k = 0;
  while (k < max && !S->done.empty()) {
    out[k++] = S->done.front();
    S->done.pop_front();
  }
  *n = k;
  return BLPG_OK;
}

int blpg_server_stats_get(blpg_server* S, blpg_server_stats* st) {
  if (!S || !st) return BLPG_E_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  std::memset(st, 0, sizeof *st);
  st->submitted = S->submitted;
  st->completed = S->completed;
  st->rejected = S->rejected;
  st->batches = S->batches;
  st->mean_batch = S->batches ? (double)S->batched_lps / (double)S->batches : 0.0;
  st->resident_bytes = S->arena_bytes + S->pending_bytes;
  if (!S->lat.empty()) {
    std::vector<double> v = S->lat;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) { return v[std::min(v.size() - 1, (size_t)(p * (double)(v.size() - 1) + 0.5))]; };
    st->p50_us = pct(0.50);
    st->p99_us = pct(0.99);
    st->max_us = v.back();
  }
  return BLPG_OK;
}

int blpg_server_replay(blpg_server* S, const blpg_lp* lps, const double* arrival_us, int64_t count,
                       blpg_server_result* out) {
  if (!S || (count > 0 && (!lps || !arrival_us)) || count < 0) return BLPG_E_INVALID;
  std::vector<int64_t> tickets(count, -1);
  std::atomic<int> err{0};
  const auto start = Clock::now();
  // several client threads (arrival k goes to thread k mod P), each submitting at the
  // scheduled times: the copy of a request into its arena happens on the client's thread
  const unsigned hw = std::thread::hardware_concurrency();
  const int P = (int)std::max(1u, std::min(8u, hw > 2 ? hw / 2 : 1u));
  auto pace = [&](int t) {
    for (int64_t k = t; k < count && !err.load(); k += P) {
      const auto due = start + std::chrono::nanoseconds((int64_t)(arrival_us[k] * 1000.0));
      while (Clock::now() < due) std::this_thread::yield();
      while (!err.load()) {
        int64_t tk = -1;
        // latency counts from the scheduled arrival
        const int rc = submit_at(S, &lps[k], nullptr, nullptr, due, &tk);
        if (rc == BLPG_OK) {
          tickets[k] = tk;
          break;
        }
        if (rc != BLPG_E_BUDGET) { err.store(rc); break; }
        std::this_thread::yield();
      }
    }
  };
  std::vector<std::thread> pacers;
  for (int t = 0; t < P; ++t) pacers.emplace_back(pace, t);
  std::map<int64_t, blpg_server_result> got;
  std::vector<blpg_server_result> buf(1024);
  // every wait ends: with no result for kStallSeconds the replay stops and reports what
  // the server holds (the pacers stop submitting too)
  auto last = Clock::now();
  while ((int64_t)got.size() < count && !err.load()) {
    int32_t n = 0;
    blpg_server_poll(S, buf.data(), (int32_t)buf.size(), 1000, &n);
    for (int i = 0; i < n; ++i) got[buf[i].ticket] = buf[i];
    if (n > 0) {
      last = Clock::now();
    } else if (Clock::now() - last > std::chrono::seconds(kStallSeconds)) {
      err.store(BLPG_E_HIP);
      set_free_error(BLPG_E_HIP, "blpg_server_replay: no result for %d s (%lld of %lld "
                     "completed): %s", kStallSeconds, (long long)got.size(), (long long)count,
                     describe(S).c_str());
    }
  }
  for (auto& th : pacers) th.join();
  if (out)
    for (int64_t k = 0; k < count; ++k) {
      auto it = got.find(tickets[k]);
      out[k] = it == got.end() ? blpg_server_result{tickets[k], BLPG_UNDECIDED, 0, 0.0, -1.0} : it->second;
    }
  return err.load();
}

}  // extern "C"
