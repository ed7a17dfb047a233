// ctx.hpp -- host-side state behind the C ABI handles (context, device batch) and the
// error helpers shared by abi.cpp and warmstart.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/blpg/blpg.h"
#include "blpg_internal.hpp"
#include "tuner.hpp"

namespace blpg {

// Rows of a batched warm-start network evaluation: rows grouped by family, each family's
// segment padded to a multiple of 64 rows.
struct WsRows {
  int64_t rows = 0;            // padded total
  int nfam = 0;
  std::vector<int32_t> seg_off, seg_rows, fam_count;  // per family (host copies)
  void* mem = nullptr;         // device: row_lp, row_fam, seg_off, seg_rows, fam_count + buffers
  size_t bytes = 0;
  int32_t *row_lp = nullptr, *row_fam = nullptr, *d_seg_off = nullptr, *d_seg_rows = nullptr,
          *d_fam_count = nullptr;
  uint16_t *X = nullptr, *H1 = nullptr, *H2 = nullptr;  // bf16 activations
  float* Z = nullptr;                                   // logits
  float* raw = nullptr;        // raw (b, c) features (host-supplied rows), or nullptr
  int max_seg = 0;             // largest segment (rows)
  size_t table_bytes = 0;      // the row tables' range at the front of mem
  bool arena = false;          // tables rewritten before every launch (serving arenas)
};

struct Bin {
  int kernel = BLPG_KERNEL_STREAM;  // BLPG_KERNEL_*
  int variant = 0;                  // tile variant
  int nt = 256;                     // stream threads
  int rcap = 0, ccap = 0;
  std::vector<int32_t> ids;
  int32_t* d_idx = nullptr;
  int64_t pool_lo = 0, pool_hi = 0;  // the bin's own range of the input pool (elements)
};

// One input array of a batch: (host pointer, length).  Arrays are stored once per key.
typedef std::pair<const void*, int64_t> SrcKey;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace blpg

struct blpg_ctx {
  blpg_options opt;
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t props{};
  std::string err;
  blpg::DevBuf ws;   // HBM slabs of the streamed kernel (one region per concurrent bin)
  blpg::DevBuf ws2;  // HBM slabs of the fallback kernel
  blpg::DevBuf vws;  // LU workspace of the verify kernel
  blpg::Tuner tuner;
  // size bins run concurrently on the context stream and these (default priority); with
  // the context stream that is 4 streams, one per hardware queue of the process
  static constexpr int kAux = 3;
  hipStream_t aux[kAux] = {};
  hipEvent_t fork = nullptr;
  int64_t gen = 0;  // bumped whenever a workspace is reallocated (captured graphs go stale)
};

struct blpg_batch {
  int64_t count = 0;
  int precision = 0;
  std::vector<blpg::LpDesc> desc;
  int64_t pool_elems = 0, spool_elems = 0, x_elems = 0, basis_elems = 0, tab_elems = 0;
  int64_t common_elems = 0;  // pool[0, common_elems): arrays read by several bins
  std::map<blpg::SrcKey, int64_t> pool_off, spool_off;  // where each distinct array lives
  char* h_stage = nullptr;   // pinned host image of desc | pool | sense pool | int32 init
  size_t stage_bytes = 0, o_pool = 0, o_spool = 0;
  bool dirty = false;        // staging newer than the device copy: the next solve copies it
  // The copy after a repack sends each pool range (the shared arrays, then each bin's)
  // packed where that is at least 10% smaller: a bitmask of the elements that are not +0.0,
  // a running set-bit count per 64-element word and the packed values, expanded on the
  // device before the range's first reader.  zr[0] is the shared range, zr[1 + i] bin i's.
  struct ZRange {
    int64_t lo = 0, hi = 0;   // pool elements
    size_t off = 0;           // in h_z / d_z: mask [nw] u64, pref [nw] i32 (8-aligned), vals
    size_t bytes = 0;         // bytes to copy
    bool packed = false;      // else copied dense from the staging pool
  };
  std::vector<ZRange> zr;
  char* h_z = nullptr;        // pinned
  char* d_z = nullptr;
  size_t z_cap = 0;
  char* d_tab = nullptr;  // final-tableau dumps (read by the FP64 check)
  int mmax = 0, nmax = 0;
  void* mem = nullptr;
  size_t mem_bytes = 0;
  blpg::LpDesc* d_desc = nullptr;
  double* d_pool = nullptr;
  int8_t* d_spool = nullptr;
  blpg::Out out{};
  blpg::Check chk{};
  int32_t* d_fb = nullptr;        // fallback-used flag per LP
  int32_t* d_warm = nullptr;      // warm-start lists (basis pool layout, -1 = none), or nullptr
  std::vector<int32_t> family;    // per LP, -1 none
  blpg_ws_model* model = nullptr; // learned warm start attached to this batch
  blpg::WsRows ws;                // its inference rows and buffers
  void* warm_mem = nullptr;       // warm lists allocated when a model was attached
  int32_t* d_all = nullptr;       // 0..count-1
  int32_t* d_fail = nullptr;      // LP ids failing the check
  int32_t* d_defer = nullptr;     // LP ids a check tier hands on (two lists, alternating)
  int32_t* d_defer2 = nullptr;
  int32_t* d_count = nullptr;     // arena batches: number of LPs (slots 0..count-1) to solve
  int32_t* d_counters = nullptr;  // [0..nbins) queues, then verify/fallback queues and counts
  std::vector<blpg::Bin> bins;
  size_t res_off = 0, res_bytes = 0;  // device span of the results (one DMA)
  char* h_res = nullptr;              // pinned host copy of that span
  std::vector<int64_t> x_off, basis_off;
  blpg_batch_info info{};
  hipEvent_t ev[96] = {};
  hipEvent_t join[16] = {};  // per-bin join of an aux stream back into the context stream
  int nev = 0;
  bool timed = false;        // last solve recorded per-stage timing events
  hipEvent_t copied[16] = {};  // per bin: its inputs are on the device (solves with a copy)
  // captured solve (slot 0: inputs resident; solves that copy inputs first run eagerly)
  hipGraph_t graph[2] = {};
  hipGraphExec_t gexec[2] = {};
  blpg_options g_opt[2] = {};  // options / model / workspace generation each graph holds
  blpg_ws_model* g_model[2] = {};
  int64_t g_gen[2] = {-1, -1};
};


namespace blpg {

inline int fail(blpg_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

// error of a call that takes no context (the no-allocation entry point): kept per host
// thread, returned by blpg_last_error(NULL); returns `code`
int set_free_error(int code, const char* fmt, ...);

// results of slots 0..k-1 of a batch (abi.cpp; the serving arenas use it)
int batch_fetch_prefix(blpg_ctx* ctx, blpg_batch* B, int64_t k, blpg_batch_results* R);

// Batches with warm-start lists crash in blocks in the register-tile kernel: its bins
// routed to the column-per-lane kernel (which crashes one target at a time) move to the
// tile variant that holds their shape (abi.cpp).
void warm_route(blpg_batch* B);

// warm-start inference of a batch's attached model (warmstart.cpp)
int ws_run_batch(blpg_ctx* ctx, blpg_batch* B, hipStream_t st);
void ws_free_rows(WsRows& R);
// serving arenas: a model attached for up to `cap` LPs of any family mix per launch
int ws_arena_attach(blpg_ctx* ctx, blpg_batch* B, blpg_ws_model* M, int cap);
// host image of the row tables for slots 0..k-1 of families fam[0..k-1] (every fam[s] in
// [0, families)); B->ws.table_bytes bytes, copied to B->ws.mem before the launch
void ws_arena_tables(const blpg_batch* B, const int32_t* fam, int k, void* image);
void ws_model_shape(const blpg_ws_model* M, int* families, int* m, int* n);

}  // namespace blpg

#define HIPCHK(ctx, call)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return blpg::fail(ctx, BLPG_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
  } while (0)
