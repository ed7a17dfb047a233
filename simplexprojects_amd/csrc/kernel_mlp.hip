// kernel_mlp.hip -- the learned warm start's network on gfx950: a per-family MLP
//   x = normalised (b, c)  ->  256 (ReLU)  ->  256 (ReLU)  ->  n+m logits
// trained and evaluated by these kernels only (no library GEMM):
//   k_gemm        batched bf16 GEMM on MFMA (v_mfma_f32_32x32x16_bf16, FP32 accumulate),
//                 generic strides so one kernel serves X*W^T (forward), dZ*W (input
//                 gradients) and dZ^T*H (weight gradients); fused bias / ReLU / ReLU-gate
//                 epilogue writing FP32 and/or bf16
//   k_features    normalised bf16 inputs from the packed LP pool or from raw features
//   k_bce_grad    sigmoid cross-entropy loss and its gradient
//   k_colsum      bias gradients
//   k_adam        Adam on FP32 master weights, refreshing the bf16 copies
//   k_pick_basis  top-m logits -> warm-start basis list (variable ids, most confident first)
// Rows of one family form a segment (offset and row count a multiple of 64), so every
// family's GEMMs run in one launch (grid.z = family).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "blpg_internal.hpp"
#include "mlp.hpp"

namespace blpg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (inputs are finite)
  const uint32_t u = __float_as_uint(f);
  const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------------------------
// GEMM: C[i][j] = sum_k A(i,k) B(k,j) (+ bias[j]) (ReLU | gate), 64x64 tile per
// workgroup of 4 waves (each wave a 32x32 MFMA block), K in chunks of 32 through LDS.
// Operand element (i,k) of A is A[i*lda_i + k*lda_k]; (k,j) of B is B[k*ldb_k + j*ldb_j].
constexpr int kGT = 64;       // tile rows / cols
constexpr int kGK = 32;       // K chunk
constexpr int kGLd = kGK + 8; // LDS row pitch (bf16 elements): 80 B, staggers banks

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ __align__(16) uint16_t As[kGT * kGLd];
  __shared__ __align__(16) uint16_t Bs[kGT * kGLd];
  const int z = blockIdx.z;
  int M = g.M, K = g.K;
  const uint16_t* A = g.A + (int64_t)z * g.sa;
  const uint16_t* B = g.B + (int64_t)z * g.sb;
  int64_t coff = (int64_t)z * g.sc, goff = (int64_t)z * g.sg;
  if (g.seg_dim == 1) {  // rows of A / C are the family's segment
    const int o = g.seg_off[z];
    M = g.seg_rows[z];
    A += (int64_t)o * g.lda_i;
    coff += (int64_t)o * g.ldc;
    goff += (int64_t)o * g.ldg;
  } else if (g.seg_dim == 2) {  // K runs over the family's segment
    const int o = g.seg_off[z];
    K = g.seg_rows[z];
    A += (int64_t)o * g.lda_k;
    B += (int64_t)o * g.ldb_k;
  }
  const int i0 = blockIdx.y * kGT, j0 = blockIdx.x * kGT;
  if (i0 >= M) return;  // uniform per workgroup
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = (w >> 1) * 32, wj = (w & 1) * 32;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  const int lr = tid >> 2, lk = (tid & 3) * 8;  // loader: row (of 64) and k offset (of 32)
  for (int k0 = 0; k0 < K; k0 += kGK) {
    // A tile [64 x 32] -> As[i][k]
    {
      const uint16_t* src = A + (int64_t)(i0 + lr) * g.lda_i + (int64_t)(k0 + lk) * g.lda_k;
      uint16_t* dst = As + lr * kGLd + lk;
      if (g.lda_k == 1) {
        *(uint4*)dst = *(const uint4*)src;
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) dst[e] = src[(int64_t)e * g.lda_k];
      }
    }
    // B tile [32 x 64] -> Bs[j][k]
    {
      const uint16_t* src = B + (int64_t)(k0 + lk) * g.ldb_k + (int64_t)(j0 + lr) * g.ldb_j;
      uint16_t* dst = Bs + lr * kGLd + lk;
      if (g.ldb_k == 1) {
        *(uint4*)dst = *(const uint4*)src;
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) dst[e] = src[(int64_t)e * g.ldb_k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kGK; kk += 16) {
      // lane l: A row wi + l%32, k = kk + 8*(l/32) .. +7; B column wj + l%32, same k
      const bf16x8 a = *(const bf16x8*)(As + (wi + (lane & 31)) * kGLd + kk + 8 * (lane >> 5));
      const bf16x8 b = *(const bf16x8*)(Bs + (wj + (lane & 31)) * kGLd + kk + 8 * (lane >> 5));
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue: acc[r] is C[wi + 8*(r/4) + 4*(l/32) + r%4][wj + l%32]
  const int j = j0 + wj + (lane & 31);
  const float bj = g.bias ? g.bias[(int64_t)z * g.sbias + j] : 0.f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = i0 + wi + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    float val = acc[r] + bj;
    if (g.relu) val = val > 0.f ? val : 0.f;
    if (g.gate) val = bf2f(g.gate[goff + (int64_t)i * g.ldg + j]) > 0.f ? val : 0.f;
    const int64_t o = coff + (int64_t)i * g.ldc + j;
    if (g.C) g.C[o] = val;
    if (g.Cb) g.Cb[o] = f2bf(val);
  }
}

hipError_t launch_gemm(const GemmArgs& g, int batch, hipStream_t st) {
  if (g.M % kGT || g.N % kGT || g.K % kGK) return hipErrorInvalidValue;
  dim3 grid(g.N / kGT, g.M / kGT, batch);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, st, g);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Inputs: row r of the packed matrix belongs to LP row_lp[r] (-1: padding) of family
// row_fam[r]; x = ((b, c) - mu[f]) * inv_sd[f], zero beyond d_in and on padding rows.
__global__ void k_features(const LpDesc* __restrict__ desc, const double* __restrict__ pool,
                           const float* __restrict__ raw, int64_t ld_raw,
                           const int32_t* __restrict__ row_lp, const int32_t* __restrict__ row_fam,
                           int64_t rows, int d_in, int d_in_pad, const float* __restrict__ mu,
                           const float* __restrict__ inv_sd, uint16_t* __restrict__ X) {
  const int64_t r = blockIdx.x;
  if (r >= rows) return;
  const int lp = row_lp[r];
  const int f = row_fam[r];
  for (int j = threadIdx.x; j < d_in_pad; j += blockDim.x) {
    float x = 0.f;
    if (lp >= 0 && j < d_in) {
      float v;
      if (raw) {
        v = raw[(int64_t)lp * ld_raw + j];
      } else {
        const LpDesc d = desc[lp];
        v = j < d.m ? (float)pool[d.b + j] : (float)pool[d.c + (j - d.m)];
      }
      x = (v - mu[(int64_t)f * d_in + j]) * inv_sd[(int64_t)f * d_in + j];
    }
    X[r * d_in_pad + j] = f2bf(x);
  }
}

hipError_t launch_features(const FeatureArgs& a, hipStream_t st) {
  if (a.rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_features, dim3((unsigned)a.rows), dim3(256), 0, st, a.desc, a.pool, a.raw,
                     a.ld_raw, a.row_lp, a.row_fam, a.rows, a.d_in, a.d_in_pad, a.mu, a.inv_sd,
                     a.X);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Sigmoid cross-entropy over the d_out real outputs of the real rows of each family,
// averaged per family; dZ = (sigmoid(z) - y) / (rows_f * d_out).  loss[f] accumulates the
// family's mean loss (atomically, FP32).
__global__ void k_bce_grad(const float* __restrict__ Z, const uint8_t* __restrict__ Y,
                           const int32_t* __restrict__ row_fam, const int32_t* __restrict__ row_lp,
                           const int32_t* __restrict__ fam_count, int64_t rows, int d_out,
                           int d_out_pad, float* __restrict__ dZ, uint16_t* __restrict__ dZb,
                           float* __restrict__ loss) {
  const int64_t r = blockIdx.x;
  if (r >= rows) return;
  const int f = row_fam[r];
  const bool real = row_lp[r] >= 0;
  const float scale = real ? 1.f / ((float)fam_count[f] * (float)d_out) : 0.f;
  float part = 0.f;
  for (int j = threadIdx.x; j < d_out_pad; j += blockDim.x) {
    const int64_t o = r * d_out_pad + j;
    float g = 0.f;
    if (real && j < d_out) {
      const float z = Z[o];
      const float y = (float)Y[o];
      const float p = 1.f / (1.f + __expf(-z));
      g = (p - y) * scale;
      part += (fmaxf(z, 0.f) - z * y + log1pf(__expf(-fabsf(z)))) * scale;
    }
    dZ[o] = g;
    dZb[o] = f2bf(g);
  }
  // block reduction of the loss part
  __shared__ float red[256];
  red[threadIdx.x] = part;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && real) atomicAdd(loss + f, red[0]);
}

hipError_t launch_bce_grad(const BceArgs& a, hipStream_t st) {
  if (a.rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bce_grad, dim3((unsigned)a.rows), dim3(256), 0, st, a.Z, a.Y, a.row_fam,
                     a.row_lp, a.fam_count, a.rows, a.d_out, a.d_out_pad, a.dZ, a.dZb, a.loss);
  return hipGetLastError();
}

// db[f][j] = sum over the family's segment rows of D[row][j]
__global__ void k_colsum(const float* __restrict__ D, int64_t ld, const int32_t* __restrict__ seg_off,
                         const int32_t* __restrict__ seg_rows, int ncol, float* __restrict__ db,
                         int64_t sdb) {
  const int f = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  const int o = seg_off[f], n = seg_rows[f];
  float s = 0.f;
  for (int r = 0; r < n; ++r) s += D[(int64_t)(o + r) * ld + j];
  db[(int64_t)f * sdb + j] = s;
}

hipError_t launch_colsum(const float* D, int64_t ld, const int32_t* seg_off, const int32_t* seg_rows,
                         int nfam, int ncol, float* db, int64_t sdb, hipStream_t st) {
  dim3 grid((ncol + 255) / 256, nfam);
  hipLaunchKernelGGL(k_colsum, grid, dim3(256), 0, st, D, ld, seg_off, seg_rows, ncol, db, sdb);
  return hipGetLastError();
}

// Adam (bias-corrected) on the FP32 master parameters; refreshes the bf16 copy.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, uint16_t* __restrict__ pb, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  float q = p[i];
  q -= lr * ((a / c1) / (sqrtf(v / c2) + eps) + wd * q);
  p[i] = q;
  pb[i] = f2bf(q);
}

hipError_t launch_adam(const AdamArgs& a, hipStream_t st) {
  const int64_t blocks = (a.n + 255) / 256;
  hipLaunchKernelGGL(k_adam, dim3((unsigned)blocks), dim3(256), 0, st, a.p, a.g, a.m1, a.m2, a.pb,
                     a.n, a.lr, a.b1, a.b2, a.eps, a.c1, a.c2, a.wd);
  return hipGetLastError();
}

__global__ void k_to_bf16(const float* __restrict__ p, uint16_t* __restrict__ pb, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pb[i] = f2bf(p[i]);
}

hipError_t launch_to_bf16(const float* p, uint16_t* pb, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, pb, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Warm-start list: the m largest logits among the d_out = n+m outputs (ties: lower id)
// in decreasing order: warm[rank] = variable id.  One wavefront per row.
// With `desc` the list of row r goes to the LP's basis slot of the batch; without it, to
// warm_pool + row_lp[r] * m0 (LPs of shape m0 x n0).
// Each logit becomes one 64-bit key ordered as (value, then lower id first): the float's
// bits mapped to an increasing unsigned word (-0 taken as +0), then the inverted id.  The
// rank of v is the number of keys above its own, one 64-bit compare per pair, the row's
// keys read two per LDS load and held four per lane (rows of up to 256 variables; longer
// rows take each lane's variables 64 apart in turn).
__device__ __forceinline__ unsigned long long pick_key(float z, int v) {
  const uint32_t u = __float_as_uint(z + 0.0f);
  const uint32_t o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ((unsigned long long)o << 32) | (uint32_t)~(uint32_t)v;
}

__global__ __launch_bounds__(64) void k_pick_basis(const float* __restrict__ Z, int d_out_pad,
                                                   const int32_t* __restrict__ row_lp, int64_t rows,
                                                   const LpDesc* __restrict__ desc, int m0, int n0,
                                                   int32_t* __restrict__ warm_pool) {
  extern __shared__ __align__(16) unsigned long long zk[];
  const int64_t r = blockIdx.x;
  if (r >= rows) return;
  const int lp = row_lp[r];
  if (lp < 0) return;
  int m = m0, n = n0;
  int64_t off = (int64_t)lp * m0;
  if (desc) {
    const LpDesc d = desc[lp];
    m = d.m; n = d.n; off = d.basis;
  }
  const int nv = n + m;
  const int nv2 = (nv + 1) & ~1;  // key 0 pads to an even count: below every real key
  for (int j = threadIdx.x; j < nv2; j += 64) zk[j] = j < nv ? pick_key(Z[r * d_out_pad + j], j) : 0ull;
  __syncthreads();
  int32_t* w = warm_pool + off;
  const ulonglong2* z2 = reinterpret_cast<const ulonglong2*>(zk);
  for (int v0 = threadIdx.x; v0 < nv; v0 += 256) {
    unsigned long long kv[4];
    int rk[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int v = v0 + 64 * q;
      kv[q] = v < nv ? zk[v] : ~0ull;
      rk[q] = 0;
    }
    for (int u = 0; u < nv2; u += 2) {
      const ulonglong2 t = z2[u >> 1];
#pragma unroll
      for (int q = 0; q < 4; ++q) rk[q] += (t.x > kv[q] ? 1 : 0) + (t.y > kv[q] ? 1 : 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int v = v0 + 64 * q;
      if (v < nv && rk[q] < m) w[rk[q]] = v;
    }
  }
}

hipError_t launch_pick_basis(const float* Z, int d_out_pad, const int32_t* row_lp, int64_t rows,
                             const LpDesc* desc, int m0, int n0, int32_t* warm_pool, int max_vars,
                             hipStream_t st) {
  if (rows <= 0) return hipSuccess;
  const size_t lds = sizeof(unsigned long long) * (size_t)((max_vars + 1) & ~1);
  hipLaunchKernelGGL(k_pick_basis, dim3((unsigned)rows), dim3(64), lds, st, Z, d_out_pad, row_lp,
                     rows, desc, m0, n0, warm_pool);
  return hipGetLastError();
}

}  // namespace blpg
