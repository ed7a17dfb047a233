// mlp.hpp -- launch interface of the warm-start network kernels (kernel_mlp.hip).
// bf16 values are passed as uint16_t bit patterns.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "blpg_internal.hpp"

namespace blpg {

// Batched GEMM C = A*B (+bias)(ReLU | gate), batch index z = family.
// M, N multiples of 64, K a multiple of 32 (segment rows are multiples of 64).
struct GemmArgs {
  int M = 0, N = 0, K = 0;
  const uint16_t* A = nullptr;
  int64_t lda_i = 0, lda_k = 0, sa = 0;
  const uint16_t* B = nullptr;
  int64_t ldb_k = 0, ldb_j = 0, sb = 0;
  float* C = nullptr;        // FP32 output (optional)
  uint16_t* Cb = nullptr;    // bf16 output (optional), same layout
  int64_t ldc = 0, sc = 0;
  const float* bias = nullptr;
  int64_t sbias = 0;
  int relu = 0;
  const uint16_t* gate = nullptr;  // keep C[i][j] only where gate[i][j] > 0 (ReLU backward)
  int64_t ldg = 0, sg = 0;
  const int32_t* seg_off = nullptr;
  const int32_t* seg_rows = nullptr;
  int seg_dim = 0;  // 0: none; 1: rows of A and C are segment rows; 2: K runs over them
};
hipError_t launch_gemm(const GemmArgs& g, int batch, hipStream_t st);

struct FeatureArgs {
  const LpDesc* desc = nullptr;  // LP pool source (raw == nullptr)
  const double* pool = nullptr;
  const float* raw = nullptr;    // or raw features: row_lp[r] indexes rows of raw
  int64_t ld_raw = 0;
  const int32_t* row_lp = nullptr;
  const int32_t* row_fam = nullptr;
  int64_t rows = 0;
  int d_in = 0, d_in_pad = 0;
  const float* mu = nullptr;
  const float* inv_sd = nullptr;
  uint16_t* X = nullptr;
};
hipError_t launch_features(const FeatureArgs& a, hipStream_t st);

struct BceArgs {
  const float* Z = nullptr;
  const uint8_t* Y = nullptr;
  const int32_t* row_fam = nullptr;
  const int32_t* row_lp = nullptr;
  const int32_t* fam_count = nullptr;
  int64_t rows = 0;
  int d_out = 0, d_out_pad = 0;
  float* dZ = nullptr;
  uint16_t* dZb = nullptr;
  float* loss = nullptr;
};
hipError_t launch_bce_grad(const BceArgs& a, hipStream_t st);

hipError_t launch_colsum(const float* D, int64_t ld, const int32_t* seg_off, const int32_t* seg_rows,
                         int nfam, int ncol, float* db, int64_t sdb, hipStream_t st);

struct AdamArgs {
  float* p = nullptr;
  const float* g = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  uint16_t* pb = nullptr;
  int64_t n = 0;
  float lr = 1e-3f, b1 = 0.9f, b2 = 0.999f, eps = 1e-8f, c1 = 1.f, c2 = 1.f, wd = 0.f;
};
hipError_t launch_adam(const AdamArgs& a, hipStream_t st);
hipError_t launch_to_bf16(const float* p, uint16_t* pb, int64_t n, hipStream_t st);

hipError_t launch_pick_basis(const float* Z, int d_out_pad, const int32_t* row_lp, int64_t rows,
                             const LpDesc* desc, int m0, int n0, int32_t* warm_pool, int max_vars,
                             hipStream_t st);

}  // namespace blpg
