// kernels.hpp -- host-side launch interface of the HIP kernels (no device code).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "blpg_internal.hpp"

namespace blpg {

// Size-generic simplex kernel (kernel_stream.hip): tableau in LDS or in an HBM slab.
struct StreamLaunch {
  int precision;  // 0 fp64, 1 fp32
  bool lds_tab;
  int nt;         // threads per workgroup: 64, 128, 256, 512
  const LpDesc* desc;
  const double* pool;
  const int8_t* spool;
  const int32_t* idx;  // LP ids of this bin
  int32_t count;
  const int32_t* count_dev = nullptr;  // if set, the LP count is read on the device
  int32_t* queue;      // work-queue counter (must be 0 at launch)
  Out out;
  double eps_d, eps_p, tol;
  int32_t max_iter, bland_after;
  void* ws;            // HBM slabs (lds_tab == false): grid * ws_stride elements
  int64_t ws_stride;
  int rcap, ccap;      // largest m+2, n+2 of the bin
  int grid;
  const int32_t* warm = nullptr;  // warm-start bases (basis pool layout), or nullptr
};
size_t stream_lds_bytes(int precision, bool lds_tab, int nt, int rcap, int ccap);
int stream_max_blocks_per_cu(int precision, bool lds_tab, int nt, int rcap, int ccap);
hipError_t launch_stream(const StreamLaunch& L, hipStream_t st);

// Register-tiled simplex kernel (kernel_tile.hip): tableau resident in VGPRs.
struct TileLaunch {
  int precision;
  int variant;  // index into the tile variant table
  const LpDesc* desc;
  const double* pool;
  const int8_t* spool;
  const int32_t* idx;
  int32_t count;
  const int32_t* count_dev = nullptr;
  int32_t* queue;
  Out out;
  double eps_d, eps_p, tol;
  int32_t max_iter, bland_after;
  int grid;
  const int32_t* warm = nullptr;  // warm-start bases (basis pool layout), or nullptr
};
struct TileVariant {
  int nw, tr, rpt, cpt;  // waves per LP, row groups, rows / columns per thread
  int rcap() const { return tr * rpt; }
  int ccap() const { return (nw * 64 / tr) * cpt; }
};
// Column-per-lane register simplex kernel (kernel_wave.hip): one LP per NW wavefronts,
// each thread holds whole columns; launched with the TileLaunch arguments (variant = wave
// variant index).
struct WaveVariant {
  int nw, cq, rows;  // waves per LP, column slots per thread, row capacity
  bool enabled;      // instantiated in this precision
  int rcap() const { return rows; }
  int ccap() const { return 1 + 64 * nw * cq; }
};
int wave_variant_count();
WaveVariant wave_variant(int precision, int v);
int wave_max_blocks_per_cu(int precision, int v);
hipError_t launch_wave(const TileLaunch& L, hipStream_t st);

int tile_variant_count(int precision);
int tile_default_variants();  // variants 0..k-1 form the static routing order
TileVariant tile_variant(int precision, int v);
int tile_max_blocks_per_cu(int precision, int v);
hipError_t launch_tile(const TileLaunch& L, hipStream_t st);

// FP64 check of final bases (kernel_verify.hip), in up to three tiers of growing LDS per
// LP; a tier below 2 given a defer_list hands the LPs whose staged blocks it cannot hold
// to that list, which the next tier takes.
struct VerifyLaunch {
  const LpDesc* desc;
  const double* pool;
  const int8_t* spool;
  const int32_t* idx;  // LPs to check (nullptr: all)
  int32_t count;
  const int32_t* count_dev = nullptr;
  Out out;             // status / basis / unb / final tableaux read from here
  Check chk;
  double tol;          // relative KKT tolerance
  int fp32_source;     // 1: the tableaux are FP32 (else FP64)
  int tier = 2;
  void* ws;            // tier 2: grid * verify_ws_stride bytes of scratch (vectors beyond LDS)
  int mcap, ncap;      // largest m, n
  int grid;
  int32_t* queue;
  int32_t* fail_list;  // LP ids that failed the check (appended)
  int32_t* fail_count;
  int32_t* defer_list = nullptr;
  int32_t* defer_count = nullptr;
};
hipError_t launch_verify(const VerifyLaunch& L, hipStream_t st);
bool verify_tier_used(int tier, int mcap, int ncap, int tsz);    // tsz: tableau element bytes
bool verify_tier_covers(int tier, int mcap, int ncap, int tsz);  // every LP of the bin fits
int verify_max_blocks_per_cu(int tier, int mcap, int ncap);
size_t verify_ws_stride(int mcap, int ncap);  // 0: vectors in LDS

// small helpers (kernel_util.hip)
hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st);
// out[0, n) from a packed range (bitmask of the elements that are not +0.0, running set-bit
// count per 64-element word, the packed values)
hipError_t launch_expand_f64(const uint64_t* mask, const int32_t* pref, const double* vals,
                             int64_t n, double* out, hipStream_t st);

}  // namespace blpg
