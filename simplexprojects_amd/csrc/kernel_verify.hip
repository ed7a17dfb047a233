// kernel_verify.hip -- FP64 check of final bases by iterative refinement (no reference
// counterpart: the reference returns only T[M-2][N-1], simplex.cu:380, and never checks it).
//
// One workgroup (256 threads) per LP, persistent over a work queue.  The solve kernel
// dumps its final tableau (solve precision) and the nonbasic variable of every column.
// With A', b' the LP with '>=' rows negated and B the basis matrix of [A' I], the tableau
// holds B^-1 column by column: the column of a nonbasic slack s_p is B^-1 e_p, and a basic
// slack's column of B^-1 is the unit vector of its row.  So no factorisation is needed:
//   x_B <- T[.,0];  repeat: r = b' - B x_B (FP64, from A),  x_B += B~^-1 r
//   y_P <- T[0, slack columns];  repeat: rho = c_B - B^T y,  y += B~^-T rho
// where B~^-1 is the tableau's (rounded) inverse: each sweep shrinks the residual by about
// cond(B) u, u the tableau's unit roundoff.  A sweep stops at the rounding floor, after
// kMaxSweeps, or when the residual no longer shrinks.  Then, with
// s_P = 1 + |b|_inf + |A|_max |x|_inf and s_D = 1 + |c|_inf + |A|_max |y|_inf,
//   primal  max(|r|_inf, bound violations of x_B (basic '=' slacks at 0)) / s_P <= tol,
//   dual    max(|rho|_inf, c_j - y^T A'_j for nonbasic structurals, -y_p for nonbasic
//           slacks of '<=' rows ('=' slacks are free)) / s_D <= tol,
//   convergence  |r|_inf <= sqrt(tol) (1 + |b|_inf) and |rho|_inf <= sqrt(tol) (1 + |c|_inf):
//           a basis too ill-conditioned for the refinement stalls with a residual of the
//           data's own scale (and a huge x or y that would pass the relative tests);
//   OPTIMAL   objective c_S^T x_S, the verified x written to the LP's x;
//   UNBOUNDED the entering variable e's ray d = B^-1 a'_e (refined the same way) must
//             satisfy d <= tol (1 + |d|) (|d| on basic '=' slacks), residual <= tol (1 +
//             |A|_max |d|), and c_e - y^T a'_e > tol s_D.
// FP32 results that fail (or FP32 INFEASIBLE whose phase-I objective is not clearly
// negative, or iteration-limit / numeric statuses) are appended to the fallback list.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels.hpp"
#include "simplex_device.hpp"

namespace blpg {

namespace {

constexpr int kVT = 256;  // threads per LP
constexpr int kVW = kVT / 64;
constexpr int kMaxSweeps = 12;

// per-LP scratch (LDS when it fits, else a slab per workgroup in HBM)
struct VLayout {
  size_t x, r, y, rho, d, bs, rowof, colp, S, Srow, Pn, Pcol, total;
  __host__ __device__ VLayout(int mcap, int ncap) {
    auto al = [](size_t v) { return (v + 15) & ~(size_t)15; };
    size_t o = 0;
    x = o; o += al(8 * (size_t)mcap);      // basic values by row
    r = o; o += al(8 * (size_t)mcap);      // residual by constraint
    y = o; o += al(8 * (size_t)mcap);      // duals by constraint
    rho = o; o += al(8 * (size_t)mcap);    // dual residual by row
    d = o; o += al(8 * (size_t)mcap);      // ray by row
    bs = o; o += al(4 * (size_t)mcap);     // basic variable of each row
    rowof = o; o += al(4 * (size_t)(mcap + ncap + 1));  // row of each basic variable, -1
    colp = o; o += al(4 * (size_t)mcap);   // tableau column of each nonbasic slack, -1
    S = o; o += al(4 * (size_t)mcap);      // basic structurals (ascending)
    Srow = o; o += al(4 * (size_t)mcap);   // their rows
    Pn = o; o += al(4 * (size_t)mcap);     // nonbasic slacks (ascending constraint index)
    Pcol = o; o += al(4 * (size_t)mcap);   // their tableau columns
    total = o;
  }
};

__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// workgroup reductions through `red` (kVW doubles), uniform result
__device__ __forceinline__ double bmax(double v, double* red) {
  v = wmax(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < kVW; ++w) r = fmax(r, red[w]);
  return r;
}
__device__ __forceinline__ double bsum(double v, double* red) {
  v = wsum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < kVW; ++w) r += red[w];
  return r;
}

// compacts {i < cnt : pred(i)} in ascending order into out (and key(i) into out2);
// returns the count (uniform).  Ballot prefix sums over 256-wide chunks.
template <class Pred, class Key>
__device__ int compact(int cnt, Pred pred, Key key, int* out, int* out2, int* red_i) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int base = 0;
  for (int c0 = 0; c0 < cnt; c0 += kVT) {
    const int i = c0 + tid;
    const bool p = i < cnt && pred(i);
    const unsigned long long bal = __ballot(p);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    __syncthreads();
    if (lane == 0) red_i[w] = __popcll(bal);
    __syncthreads();
    int off = base;
    for (int q = 0; q < w; ++q) off += red_i[q];
    if (p) {
      out[off + before] = i;
      out2[off + before] = key(i);
    }
    int tot = 0;
    for (int q = 0; q < kVW; ++q) tot += red_i[q];
    base += tot;
  }
  __syncthreads();
  return base;
}

}  // namespace

template <typename TT>
__device__ void verify_lp(int lp, const LpDesc& D, const double* __restrict__ pool,
                          const int8_t* __restrict__ spool, Out out, Check chk, double tol,
                          int fp32_source, char* wsp, double* red, int* red_i,
                          int32_t* fail_list, int32_t* fail_count) {
  const int tid = threadIdx.x;
  const int m = D.m, n = D.n, R = m + 2, C = n + 2;
  const double* A = pool + D.a;
  const double* b = pool + D.b;
  const double* c = pool + D.c;
  const int8_t* sense = D.s >= 0 ? spool + D.s : nullptr;
  const int status = out.status[lp];
  const int32_t* basis = out.basis + D.basis;
  const TT* T = (const TT*)(out.tab + 8 * D.t);
  const int32_t* nbv = (const int32_t*)(out.tab + 8 * (D.t + (int64_t)R * C));
  auto sg = [&](int i) { return (sense && sense[i] == 2) ? -1.0 : 1.0; };
  auto is_eq = [&](int i) { return sense && sense[i] == 1; };
  const VLayout L(m, n);
  double* x = (double*)(wsp + L.x);
  double* r = (double*)(wsp + L.r);
  double* y = (double*)(wsp + L.y);
  double* rho = (double*)(wsp + L.rho);
  double* dv = (double*)(wsp + L.d);
  int* bs = (int*)(wsp + L.bs);
  int* rowof = (int*)(wsp + L.rowof);
  int* colp = (int*)(wsp + L.colp);
  int* S = (int*)(wsp + L.S);
  int* Srow = (int*)(wsp + L.Srow);
  int* Pn = (int*)(wsp + L.Pn);
  int* Pcol = (int*)(wsp + L.Pcol);

  double bm = 0, cm = 0;
  for (int i = tid; i < m; i += kVT) bm = fmax(bm, fabs(b[i]));
  for (int j = tid; j < n; j += kVT) cm = fmax(cm, fabs(c[j]));
  bm = bmax(bm, red);
  cm = bmax(cm, red);
  const double am = out.amax[lp];

  const double sq_tol = sqrt(tol);
  int ok = 2;
  double rp = -1, rd = -1, zobj = out.obj[lp];
  const bool checkable = (status == kStatusOptimal || status == kStatusUnbounded) && out.tab;
  int kx = -1;
  if (!checkable) {
    if (fp32_source)
      ok = (status == kStatusInfeasible && out.wfin[lp] < -1e-3 * (1.0 + bm)) ? 1 : 0;
  } else {
    // ---- basis bookkeeping ----
    for (int v = tid; v < n + m + 1; v += kVT) rowof[v] = -1;
    for (int p = tid; p < m; p += kVT) colp[p] = -1;
    __syncthreads();
    int bad = 0;
    for (int i = tid; i < m; i += kVT) {
      const int v = basis[i];
      bs[i] = v;
      if (v < 0 || v >= n + m) bad = 1;
      else rowof[v] = i;
    }
    for (int j = 1 + tid; j < C; j += kVT) {
      const int v = nbv[j];
      if (v >= n && v < n + m) colp[v - n] = j;
    }
    bad = (int)bmax((double)bad, red);
    __syncthreads();
    const int k = compact(n, [&](int j) { return rowof[j] >= 0; }, [&](int j) { return rowof[j]; },
                          S, Srow, red_i);
    const int kp = compact(m, [&](int p) { return colp[p] >= 0; }, [&](int p) { return colp[p]; },
                           Pn, Pcol, red_i);
    if (bad || k != kp) {
      ok = 0;
    } else {
      const int lane = tid & 63, wave = tid >> 6;
      // r = rhs - B v for the basis matrix B of [A' I]: one wavefront per constraint row,
      // lanes over the basic structurals (ascending, so a row's loads stay together).
      // rhs is b' (ray == false) or the entering column a'_e; returns the wave's max |r|.
      auto resid = [&](const double* vb, bool ray, int e) {
        double rl = 0;
        for (int p = wave; p < m; p += kVW) {
          const double* Ap = A + (int64_t)p * n;
          double s = 0;
          for (int a = lane; a < k; a += 64) s = fma(Ap[S[a]], vb[Srow[a]], s);
          s = wsum(s);
          const int q = rowof[n + p];
          const double rhs = !ray ? b[p] : (e < n ? Ap[e] : (e - n == p ? sg(p) : 0.0));
          const double val = sg(p) * (rhs - s) - (q >= 0 ? vb[q] : 0.0);
          if (lane == 0) r[p] = val;
          rl = fmax(rl, fabs(val));
        }
        return rl;
      };
      // vb += B~^-1 r: row i of B~^-1 is the tableau row's slack columns, plus 1 at the
      // row's own basic slack; one wavefront per row, lanes over the nonbasic slacks.
      auto correct = [&](double* vb) {
        for (int i = wave; i < m; i += kVW) {
          const TT* Ti = T + (int64_t)(2 + i) * C;
          double s = 0;
          for (int a = lane; a < k; a += 64) s = fma((double)Ti[Pcol[a]], r[Pn[a]], s);
          s = wsum(s);
          const int v = bs[i];
          if (lane == 0) vb[i] += s + (v >= n ? r[v - n] : 0.0);
        }
        __syncthreads();
      };
      auto vmax = [&](const double* vb) {
        double l = 0;
        for (int i = tid; i < m; i += kVT) l = fmax(l, fabs(vb[i]));
        return bmax(l, red);
      };
      // refinement: stops at the rounding floor, after kMaxSweeps, or when a sweep no
      // longer shrinks the residual (a basis too ill-conditioned for the tableau's inverse)
      auto refine = [&](double* vb, bool ray, int e, double scale0, double& res, double& vm) {
        double prev = DBL_MAX;
        for (int sweep = 0;; ++sweep) {
          res = bmax(resid(vb, ray, e), red);
          vm = vmax(vb);
          if (sweep == kMaxSweeps || res <= 1e-14 * (scale0 + am * vm) || res > 0.9 * prev) break;
          prev = res;
          correct(vb);
        }
      };

      // ---- x_B ----
      for (int i = tid; i < m; i += kVT) x[i] = (double)T[(int64_t)(2 + i) * C];
      __syncthreads();
      double xm = 0, rres = 0;
      refine(x, false, -1, 1.0 + bm, rres, xm);
      double pv = 0;
      for (int i = tid; i < m; i += kVT) {
        const int v = bs[i];
        const double xv = x[i];
        if (v < n) pv = fmax(pv, -xv);
        else pv = fmax(pv, is_eq(v - n) ? fabs(xv) : -xv);
      }
      pv = bmax(pv, red);
      rp = fmax(pv, rres) / (1.0 + bm + am * xm);
      const bool pconv = rres <= sq_tol * (1.0 + bm);

      // ---- y: rho = c_B - B^T y (basic structurals in S order, lanes along A's rows),
      //      y += B~^-T rho (nonbasic slack p: its tableau column against rho) ----
      for (int p = tid; p < m; p += kVT) y[p] = colp[p] >= 0 ? (double)T[colp[p]] : 0.0;
      __syncthreads();
      double ym = 0, rres_d = 0, prev = DBL_MAX;
      for (int sweep = 0;; ++sweep) {
        double rl = 0;
        for (int a = tid; a < k; a += kVT) {
          const int v = S[a];
          double s0 = 0, s1 = 0;
          int p = 0;
          for (; p + 2 <= m; p += 2) {
            s0 = fma(sg(p) * A[(int64_t)p * n + v], y[p], s0);
            s1 = fma(sg(p + 1) * A[(int64_t)(p + 1) * n + v], y[p + 1], s1);
          }
          if (p < m) s0 = fma(sg(p) * A[(int64_t)p * n + v], y[p], s0);
          const double val = c[v] - (s0 + s1);
          rho[Srow[a]] = val;
          rl = fmax(rl, fabs(val));
        }
        for (int p = tid; p < m; p += kVT) {
          const int q = rowof[n + p];
          if (q >= 0) {
            rho[q] = -y[p];
            rl = fmax(rl, fabs(y[p]));
          }
        }
        rres_d = bmax(rl, red);
        ym = vmax(y);
        if (sweep == kMaxSweeps || rres_d <= 1e-14 * (1.0 + cm + am * ym) || rres_d > 0.9 * prev) break;
        prev = rres_d;
        for (int a = tid; a < k; a += kVT) {
          const int j = Pcol[a];
          double s0 = 0, s1 = 0;
          int i = 0;
          for (; i + 2 <= m; i += 2) {
            s0 = fma((double)T[(int64_t)(2 + i) * C + j], rho[i], s0);
            s1 = fma((double)T[(int64_t)(3 + i) * C + j], rho[i + 1], s1);
          }
          if (i < m) s0 = fma((double)T[(int64_t)(2 + i) * C + j], rho[i], s0);
          y[Pn[a]] += s0 + s1;
        }
        for (int p = tid; p < m; p += kVT) {
          const int q = rowof[n + p];
          if (q >= 0) y[p] += rho[q];
        }
        __syncthreads();
      }
      // dual feasibility: nonbasic structurals and nonbasic '<=' slacks
      double dvv = 0;
      for (int j = tid; j < n; j += kVT) {
        if (rowof[j] >= 0) continue;
        double s0 = 0, s1 = 0;
        int p = 0;
        for (; p + 2 <= m; p += 2) {
          s0 = fma(sg(p) * A[(int64_t)p * n + j], y[p], s0);
          s1 = fma(sg(p + 1) * A[(int64_t)(p + 1) * n + j], y[p + 1], s1);
        }
        if (p < m) s0 = fma(sg(p) * A[(int64_t)p * n + j], y[p], s0);
        dvv = fmax(dvv, c[j] - (s0 + s1));
      }
      for (int p = tid; p < m; p += kVT)
        if (colp[p] >= 0 && !is_eq(p)) dvv = fmax(dvv, -y[p]);
      rd = fmax(bmax(dvv, red), rres_d) / (1.0 + cm + am * ym);
      const bool dconv = rres_d <= sq_tol * (1.0 + cm);

      // objective of the FP64 basic solution
      double zp = 0;
      for (int a = tid; a < k; a += kVT) zp = fma(c[S[a]], x[Srow[a]], zp);
      const double zsum = bsum(zp, red);
      kx = k;
      if (status == kStatusOptimal) {
        zobj = zsum;
        ok = (rp <= tol && rd <= tol && pconv && dconv) ? 1 : 0;
      } else {
        // ---- unbounded: ray of the entering variable e, d = B^-1 a'_e by refinement ----
        const int e = out.unb[lp];
        int je = -1;
        for (int j = 1 + tid; j < C; j += kVT)
          if (nbv[j] == e) je = j;
        je = (int)bmax((double)je, red);
        bool good = e >= 0 && e < n + m && rowof[e] < 0 && je > 0 && rp <= tol && pconv;
        if (good) {
          for (int i = tid; i < m; i += kVT) dv[i] = (double)T[(int64_t)(2 + i) * C + je];
          __syncthreads();
          double dmax = 0, rres2 = 0;
          refine(dv, true, e, 1.0, rres2, dmax);
          double viol = 0;
          for (int i = tid; i < m; i += kVT) {
            const int bv = bs[i];
            viol = fmax(viol, (bv >= n && is_eq(bv - n)) ? fabs(dv[i]) : dv[i]);
          }
          viol = bmax(viol, red);
          double part = 0;
          if (e < n) {
            for (int p = tid; p < m; p += kVT) part = fma(y[p], sg(p) * A[(int64_t)p * n + e], part);
          } else {
            for (int p = tid; p < m; p += kVT) if (p == e - n) part = y[p];
          }
          const double re = (e < n ? c[e] : 0.0) - bsum(part, red);
          good = viol <= tol * (1.0 + dmax) && rres2 <= tol * (1.0 + am * dmax) &&
                 re > tol * (1.0 + cm + am * ym) && dconv;
        }
        ok = good ? 1 : 0;
        rd = -1;
      }
    }
  }
  __syncthreads();
  // a verified result reports the FP64 basic solution: x_S from the refinement, 0 elsewhere
  if (ok == 1 && kx >= 0) {
    double* xo = out.x + D.x;
    for (int j = tid; j < n; j += kVT) xo[j] = 0.0;
    __syncthreads();
    for (int a = tid; a < kx; a += kVT) xo[S[a]] = x[Srow[a]];
  }
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
  // The fallback is warm-started from the FP32 basis only when that basis was
  // (nearly) optimal; other failed results restart cold (crash list cleared).
  if (fp32_source && ok == 0 && fail_list && status != kStatusOptimal) {
    int32_t* bw = out.basis + D.basis;
    for (int i = tid; i < m; i += kVT) bw[i] = -1;
  }
}

__global__ __launch_bounds__(kVT) void k_verify(const LpDesc* __restrict__ desc,
                                                const double* __restrict__ pool,
                                                const int8_t* __restrict__ spool,
                                                const int32_t* __restrict__ idx, int32_t count,
                                                const int32_t* count_dev, Out out, Check chk,
                                                double tol, int fp32_source, char* ws,
                                                int64_t ws_stride, int lds_scratch,
                                                int32_t* queue, int32_t* fail_list,
                                                int32_t* fail_count) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_q;
  __shared__ double red[kVW];
  __shared__ int red_i[kVW];
  char* wsp = lds_scratch ? (char*)smem : ws + (int64_t)blockIdx.x * ws_stride;
  if (count_dev) count = *count_dev;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
    __syncthreads();
    const int q = uniform(s_q);
    if (q >= count) break;
    const int lp = idx ? idx[q] : q;
    const LpDesc D = desc[lp];
    // the dump's precision: FP32 tableau, or FP64 for FP64 solves and FP64 fallbacks
    if (fp32_source)
      verify_lp<float>(lp, D, pool, spool, out, chk, tol, 1, wsp, red, red_i, fail_list, fail_count);
    else
      verify_lp<double>(lp, D, pool, spool, out, chk, tol, 0, wsp, red, red_i, fail_list, fail_count);
  }
}

size_t verify_scratch_bytes(int mcap, int ncap) { return VLayout(mcap, ncap).total; }

static bool verify_in_lds(int mcap, int ncap) { return verify_scratch_bytes(mcap, ncap) <= 48 * 1024; }

int verify_max_blocks_per_cu(int mcap, int ncap) {
  int nb = 0;
  const void* fn = reinterpret_cast<const void*>(&k_verify);
  const size_t lds = verify_in_lds(mcap, ncap) ? verify_scratch_bytes(mcap, ncap) : 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kVT, lds) != hipSuccess) return 0;
  return nb;
}

// HBM scratch per workgroup when the LP's vectors exceed the LDS share (0: LDS)
size_t verify_ws_stride(int mcap, int ncap) {
  return verify_in_lds(mcap, ncap) ? 0 : (verify_scratch_bytes(mcap, ncap) + 255) & ~(size_t)255;
}

hipError_t launch_verify(const VerifyLaunch& V, hipStream_t st) {
  const bool lds = verify_in_lds(V.mcap, V.ncap);
  const size_t bytes = lds ? verify_scratch_bytes(V.mcap, V.ncap) : 0;
  const int lds_scratch = lds ? 1 : 0;
  char* ws = (char*)V.ws;
  const int64_t stride = (int64_t)verify_ws_stride(V.mcap, V.ncap);
  const void* fn = reinterpret_cast<const void*>(&k_verify);
  void* args[] = {(void*)&V.desc, (void*)&V.pool, (void*)&V.spool, (void*)&V.idx,
                  (void*)&V.count, (void*)&V.count_dev, (void*)&V.out, (void*)&V.chk,
                  (void*)&V.tol, (void*)&V.fp32_source, (void*)&ws, (void*)&stride,
                  (void*)&lds_scratch, (void*)&V.queue, (void*)&V.fail_list,
                  (void*)&V.fail_count};
  return hipLaunchKernel(fn, dim3(V.grid), dim3(kVT), args, bytes, st);
}

}  // namespace blpg
