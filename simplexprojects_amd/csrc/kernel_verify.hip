// kernel_verify.hip -- FP64 check of final bases (no reference counterpart: the reference
// returns only T[M-2][N-1], simplex.cu:380, and never checks it).
//
// One workgroup per LP, persistent over a work queue.  For an LP solved to OPTIMAL:
//   B x_B = b'  and  B^T y = c_B  by LU with partial pivoting (B in LDS when it fits,
//   else in an HBM slab), then
//   primal residual = max(-x_j, -s_i ('<='), |s_i| ('='), |b' - A'x - s|) / (1 + |b|_inf)
//   dual residual   = max over nonbasic, non-fixed variables of (c_j - y^T a'_j)_+ / (1 + |c|_inf)
// and the FP64 objective c^T x.  For UNBOUNDED: the ray of the entering variable v,
// d = B^-1 a'_v, must keep every basic variable feasible (d <= tol, |d| <= tol for fixed
// slacks) and improve the objective (c_v - y^T a'_v > tol).
// An FP32 result that fails (or an FP32 INFEASIBLE whose phase-I objective is not
// clearly negative, or an iteration-limit / numeric status) is appended to a fallback
// list that the FP64 kernel re-solves.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels.hpp"
#include "simplex_device.hpp"

namespace blpg {

namespace {

constexpr int VNT = 256;
constexpr int VNW = VNT / 64;

struct VerifyLayout {
  size_t B, piv, xb, y, d, inb, red, total;
  __host__ __device__ VerifyLayout(bool lds, int mcap, int ncap) {
    auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
    size_t o = 0;
    B = o; if (lds) o += al(sizeof(double) * (size_t)mcap * mcap);
    piv = o; o += al(sizeof(int) * mcap);
    xb = o; o += al(sizeof(double) * mcap);
    y = o; o += al(sizeof(double) * mcap);
    d = o; o += al(sizeof(double) * mcap);
    inb = o; o += al((size_t)mcap + ncap + 1);
    red = o; o += al(sizeof(RedSlot<double>) * 2 * VNW + sizeof(double) * 2 * VNW);
    total = o;
  }
};

__device__ double block_sum(double v, double* slots, int& phase) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  double* s = slots + phase * VNW;
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = s[0];
#pragma unroll
  for (int q = 1; q < VNW; ++q) r += s[q];
  phase ^= 1;
  return r;
}

__device__ double block_max(double v, double* slots, int& phase) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  double* s = slots + phase * VNW;
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = s[0];
#pragma unroll
  for (int q = 1; q < VNW; ++q) r = fmax(r, s[q]);
  phase ^= 1;
  return r;
}

}  // namespace

template <bool LDS>
__global__ __launch_bounds__(VNT) void k_verify(const LpDesc* __restrict__ desc,
                                                const double* __restrict__ pool,
                                                const int8_t* __restrict__ spool,
                                                const int32_t* __restrict__ idx, int32_t count,
                                                const int32_t* count_dev, Out out, Check chk,
                                                double tol, int fp32_source, double* ws,
                                                int64_t ws_stride, int mcap, int ncap,
                                                int32_t* queue, int32_t* fail_list,
                                                int32_t* fail_count) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_q, s_bad;
  const VerifyLayout L(LDS, mcap, ncap);
  double* Bm = LDS ? (double*)(smem + L.B) : ws + (int64_t)blockIdx.x * ws_stride;
  int* piv = (int*)(smem + L.piv);
  double* xb = (double*)(smem + L.xb);
  double* y = (double*)(smem + L.y);
  double* dv = (double*)(smem + L.d);
  uint8_t* inb = smem + L.inb;
  Reducer<double, VNW> red;
  red.init((RedSlot<double>*)(smem + L.red));
  double* fslots = (double*)(smem + L.red + sizeof(RedSlot<double>) * 2 * VNW);
  int fphase = 0;
  const int tid = threadIdx.x;
  if (count_dev) count = *count_dev;

  for (;;) {
    __syncthreads();
    if (tid == 0) s_q = atomicAdd(queue, 1);
    __syncthreads();
    const int q = s_q;
    if (q >= count) break;
    const int lp = idx ? idx[q] : q;
    const LpDesc D = desc[lp];
    const int m = D.m, n = D.n;
    const double* A = pool + D.a;
    const double* b = pool + D.b;
    const double* c = pool + D.c;
    const int8_t* sense = D.s >= 0 ? spool + D.s : nullptr;
    const int status = out.status[lp];
    const int32_t* basis = out.basis + D.basis;

    double bmax = 0, cmax = 0;
    for (int i = tid; i < m; i += VNT) bmax = fmax(bmax, fabs(b[i]));
    for (int j = tid; j < n; j += VNT) cmax = fmax(cmax, fabs(c[j]));
    bmax = block_max(bmax, fslots, fphase);
    cmax = block_max(cmax, fslots, fphase);

    int ok = 2;  // not checkable
    double rp = -1, rd = -1, zobj = out.obj[lp];
    const bool checkable = status == kStatusOptimal || status == kStatusUnbounded;
    if (!checkable) {
      if (fp32_source)
        ok = (status == kStatusInfeasible && out.wfin[lp] < -1e-3 * (1.0 + bmax)) ? 1 : 0;
    } else {
      // ---- basis matrix B (row-major m x m), columns of [A' | I] ----
      for (int v = tid; v < n + m; v += VNT) inb[v] = 0;
      if (tid == 0) s_bad = 0;
      __syncthreads();
      for (int k = tid; k < m; k += VNT) {
        const int v = basis[k];
        if (v < 0 || v >= n + m) s_bad = 1;
        else inb[v] = 1;
      }
      __syncthreads();
      int bad = s_bad;
      if (!bad) {
        for (int e = tid; e < m * m; e += VNT) {
          const int i = e / m, k = e - (e / m) * m;
          const int v = basis[k];
          const double sg = (sense && sense[i] == 2) ? -1.0 : 1.0;
          Bm[e] = v < n ? sg * A[(int64_t)i * n + v] : (v - n == i ? 1.0 : 0.0);
        }
        __syncthreads();
        double amax = 0;
        for (int e = tid; e < m * m; e += VNT) amax = fmax(amax, fabs(Bm[e]));
        amax = block_max(amax, fslots, fphase);
        // ---- LU with partial pivoting (lowest row index on ties) ----
        for (int k = 0; k < m && !bad; ++k) {
          double bv = DBL_MAX;
          int bk = INT_MAX, bp = -1;
          for (int i = k + tid; i < m; i += VNT) {
            const double v = -fabs(Bm[i * m + k]);
            if (better(v, i, bv, bk)) { bv = v; bk = i; bp = i; }
          }
          red.argmin(bv, bk, bp);
          if (!(-bv > 1e-13 * (amax > 0 ? amax : 1.0))) { bad = 1; break; }
          if (tid == 0) piv[k] = bp;
          if (bp != k)
            for (int j = tid; j < m; j += VNT) {
              const double t = Bm[k * m + j];
              Bm[k * m + j] = Bm[bp * m + j];
              Bm[bp * m + j] = t;
            }
          __syncthreads();
          const double dk = Bm[k * m + k];
          for (int i = k + 1 + tid; i < m; i += VNT) Bm[i * m + k] /= dk;
          __syncthreads();
          const int rem = m - k - 1;
          for (int e = tid; e < rem * rem; e += VNT) {
            const int i = k + 1 + e / rem, j = k + 1 + (e - (e / rem) * rem);
            Bm[i * m + j] = fma(-Bm[i * m + k], Bm[k * m + j], Bm[i * m + j]);
          }
          __syncthreads();
        }
      }
      if (bad) {
        ok = 0;
      } else {
        // ---- B x_B = b' ----
        for (int i = tid; i < m; i += VNT) xb[i] = (sense && sense[i] == 2) ? -b[i] : b[i];
        __syncthreads();
        if (tid == 0)
          for (int k = 0; k < m; ++k)
            if (piv[k] != k) { const double t = xb[k]; xb[k] = xb[piv[k]]; xb[piv[k]] = t; }
        __syncthreads();
        for (int j = 0; j < m; ++j) {
          const double xj = xb[j];
          for (int i = j + 1 + tid; i < m; i += VNT) xb[i] = fma(-Bm[i * m + j], xj, xb[i]);
          __syncthreads();
        }
        for (int j = m - 1; j >= 0; --j) {
          const double xj = xb[j] / Bm[j * m + j];
          __syncthreads();
          if (tid == 0) xb[j] = xj;
          for (int i = tid; i < j; i += VNT) xb[i] = fma(-Bm[i * m + j], xj, xb[i]);
          __syncthreads();
        }
        // ---- B^T y = c_B ----
        for (int k = tid; k < m; k += VNT) y[k] = basis[k] < n ? c[basis[k]] : 0.0;
        __syncthreads();
        for (int j = 0; j < m; ++j) {  // U^T z = c_B
          const double zj = y[j] / Bm[j * m + j];
          __syncthreads();
          if (tid == 0) y[j] = zj;
          for (int i = j + 1 + tid; i < m; i += VNT) y[i] = fma(-Bm[j * m + i], zj, y[i]);
          __syncthreads();
        }
        for (int j = m - 1; j >= 0; --j) {  // L^T w = z
          const double wj = y[j];
          for (int i = tid; i < j; i += VNT) y[i] = fma(-Bm[j * m + i], wj, y[i]);
          __syncthreads();
        }
        if (tid == 0)
          for (int k = m - 1; k >= 0; --k)
            if (piv[k] != k) { const double t = y[k]; y[k] = y[piv[k]]; y[piv[k]] = t; }
        __syncthreads();

        // ---- x (structural) and slacks ----
        double* x = chk.x + D.x;
        for (int j = tid; j < n; j += VNT) x[j] = 0.0;
        __syncthreads();
        for (int k = tid; k < m; k += VNT)
          if (basis[k] < n) x[basis[k]] = xb[k];
        __syncthreads();
        // primal: bounds of basic variables, then row residuals (one wave per row)
        double p = 0;
        for (int k = tid; k < m; k += VNT) {
          const int v = basis[k];
          const bool fixed = v >= n && sense && sense[v - n] == 1;
          p = fmax(p, fixed ? fabs(xb[k]) : -xb[k]);
        }
        // slack values (dv is free until the unbounded check)
        for (int i = tid; i < m; i += VNT) dv[i] = 0.0;
        __syncthreads();
        for (int k = tid; k < m; k += VNT)
          if (basis[k] >= n) dv[basis[k] - n] = xb[k];
        __syncthreads();
        const int lane = tid & 63, w = tid >> 6;
        for (int i = w; i < m; i += VNW) {
          double acc = 0;
          for (int j = lane; j < n; j += 64) acc = fma(A[(int64_t)i * n + j], x[j], acc);
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
          const double sg = (sense && sense[i] == 2) ? -1.0 : 1.0;
          if (lane == 0) p = fmax(p, fabs(sg * b[i] - dv[i] - sg * acc));
        }
        // scale-aware normalisation: |r| / (1 + |b|_inf + |A|_max |x|_inf)
        double xm = 0;
        for (int k = tid; k < m; k += VNT) xm = fmax(xm, fabs(xb[k]));
        xm = block_max(xm, fslots, fphase);
        double ym = 0;
        for (int k = tid; k < m; k += VNT) ym = fmax(ym, fabs(y[k]));
        ym = block_max(ym, fslots, fphase);
        double am = 0;
        for (int e = tid; e < m * n; e += VNT) am = fmax(am, fabs(A[e]));
        am = block_max(am, fslots, fphase);
        rp = block_max(p, fslots, fphase) / (1.0 + bmax + am * xm);
        // dual: reduced costs of nonbasic variables (one thread per column)
        double dmax = 0;
        for (int j = tid; j < n; j += VNT) {
          double r = c[j];
          for (int i = 0; i < m; ++i) {
            const double sg = (sense && sense[i] == 2) ? -1.0 : 1.0;
            r = fma(-y[i], sg * A[(int64_t)i * n + j], r);
          }
          if (!inb[j]) dmax = fmax(dmax, r);
        }
        for (int i = tid; i < m; i += VNT)
          if (!inb[n + i] && !(sense && sense[i] == 1)) dmax = fmax(dmax, -y[i]);
        rd = block_max(dmax, fslots, fphase) / (1.0 + cmax + am * ym);
        double zp = 0;
        for (int j = tid; j < n; j += VNT) zp = fma(c[j], x[j], zp);
        const double zsum = block_sum(zp, fslots, fphase);

        if (status == kStatusOptimal) {
          zobj = zsum;
          ok = (rp <= tol && rd <= tol) ? 1 : 0;
        } else {
          // ---- unbounded: ray of the entering variable ----
          const int v = out.unb[lp];
          __syncthreads();
          bool good = v >= 0 && v < n + m && !inb[v] && rp <= tol;
          if (good) {
            for (int i = tid; i < m; i += VNT) {
              const double sg = (sense && sense[i] == 2) ? -1.0 : 1.0;
              dv[i] = v < n ? sg * A[(int64_t)i * n + v] : (v - n == i ? 1.0 : 0.0);
            }
            __syncthreads();
            if (tid == 0)
              for (int k = 0; k < m; ++k)
                if (piv[k] != k) { const double t = dv[k]; dv[k] = dv[piv[k]]; dv[piv[k]] = t; }
            __syncthreads();
            for (int j = 0; j < m; ++j) {
              const double dj = dv[j];
              for (int i = j + 1 + tid; i < m; i += VNT) dv[i] = fma(-Bm[i * m + j], dj, dv[i]);
              __syncthreads();
            }
            for (int j = m - 1; j >= 0; --j) {
              const double dj = dv[j] / Bm[j * m + j];
              __syncthreads();
              if (tid == 0) dv[j] = dj;
              for (int i = tid; i < j; i += VNT) dv[i] = fma(-Bm[i * m + j], dj, dv[i]);
              __syncthreads();
            }
            // reduced cost of v
            double rv = 0;
            if (v < n) {
              double part = 0;
              for (int i = tid; i < m; i += VNT) {
                const double sg = (sense && sense[i] == 2) ? -1.0 : 1.0;
                part = fma(y[i], sg * A[(int64_t)i * n + v], part);
              }
              rv = c[v] - block_sum(part, fslots, fphase);
            } else {
              rv = -y[v - n];
            }
            double viol = 0;
            const double dtol = tol * (1.0 + bmax) * 1e3;
            for (int k = tid; k < m; k += VNT) {
              const int bv2 = basis[k];
              const bool fixed = bv2 >= n && sense && sense[bv2 - n] == 1;
              viol = fmax(viol, fixed ? fabs(dv[k]) : dv[k]);
            }
            viol = block_max(viol, fslots, fphase);
            good = viol <= dtol && rv > tol * (1.0 + cmax);
          }
          ok = good ? 1 : 0;
          rd = -1;  // no dual solution for an unbounded LP
        }
      }
    }
    __syncthreads();
    if (tid == 0) {
      chk.ok[lp] = ok;
      chk.obj[lp] = zobj;
      chk.kkt_p[lp] = rp;
      chk.kkt_d[lp] = rd;
      if (fp32_source && ok == 0 && fail_list) {
        const int pos = atomicAdd(fail_count, 1);
        fail_list[pos] = lp;
        chk.fb[lp] = 1;
      }
    }
  }
}

size_t verify_lds_bytes(int mcap, int ncap, bool lds) { return VerifyLayout(lds, mcap, ncap).total; }

int verify_max_blocks_per_cu(int mcap, int ncap, bool lds) {
  int nb = 0;
  const void* fn = lds ? reinterpret_cast<const void*>(&k_verify<true>)
                       : reinterpret_cast<const void*>(&k_verify<false>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, VNT, verify_lds_bytes(mcap, ncap, lds)) !=
      hipSuccess)
    return 0;
  return nb;
}

hipError_t launch_verify(const VerifyLaunch& V, hipStream_t st) {
  const bool lds = V.ws == nullptr;
  const size_t bytes = verify_lds_bytes(V.mcap, V.ncap, lds);
  const void* fn = lds ? reinterpret_cast<const void*>(&k_verify<true>)
                       : reinterpret_cast<const void*>(&k_verify<false>);
  void* args[] = {(void*)&V.desc, (void*)&V.pool, (void*)&V.spool, (void*)&V.idx,
                  (void*)&V.count, (void*)&V.count_dev, (void*)&V.out, (void*)&V.chk,
                  (void*)&V.tol, (void*)&V.fp32_source, (void*)&V.ws, (void*)&V.ws_stride,
                  (void*)&V.mcap, (void*)&V.ncap, (void*)&V.queue, (void*)&V.fail_list,
                  (void*)&V.fail_count};
  return hipLaunchKernel(fn, dim3(V.grid), dim3(VNT), args, bytes, st);
}

}  // namespace blpg
