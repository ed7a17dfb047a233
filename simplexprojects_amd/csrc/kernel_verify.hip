// kernel_verify.hip -- FP64 check of final bases (no reference counterpart: the reference
// returns only T[M-2][N-1], simplex.cu:380, and never checks it).
//
// One wavefront per LP, persistent over a work queue, FP64 throughout.  A basis of m
// variable ids splits into basic structurals S and basic slacks; the rows P whose slack is
// nonbasic satisfy |P| = |S| = k.  With A', b' the LP with '>=' rows negated and
// M = A'[P, S] (k x k), the basic solution and duals are
//   x_S = M^-1 b'_P,   s_Q = b'_Q - A'[Q, S] x_S   (Q: rows with a basic slack),
//   y_P = M^-T c_S,    y_Q = 0,
// so only a k x k LU (partial pivoting, lowest row on ties) is needed, not an m x m one;
// M sits in LDS when k <= klds, else in an HBM slab owned by the wavefront.  The solve
// path runs a first pass with a small klds (more wavefronts per CU) that defers LPs with
// k > klds to a list, and a second pass with klds = 64 over that list.
// Checks, relative to (1 + |b|_inf + |A|_max |x|_inf) and (1 + |c|_inf + |A|_max |y|_inf):
//   primal  x_S >= 0, s_Q >= 0 ('<=') / s_Q = 0 ('='), |b'_P - A'[P,S] x_S|
//   dual    c_j - y_P^T A'[P, j] <= 0 for nonbasic structurals, -y_i <= 0 for nonbasic
//           slacks of '<=' rows (fixed '=' slacks are free)
//   OPTIMAL objective c_S^T x_S;  UNBOUNDED: the entering variable v's ray
//           d = B^-1 a'_v must satisfy d <= tol (|d| <= tol on basic '=' slacks) and
//           c_v - y^T a'_v > tol.
// FP32 results that fail (or FP32 INFEASIBLE whose phase-I objective is not clearly
// negative, or iteration-limit / numeric statuses) are appended to the fallback list.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels.hpp"
#include "simplex_device.hpp"

namespace blpg {

namespace {

struct VerifyLayout {
  size_t M, S, P, pos, rhs, y, d, inb, total;
  __host__ __device__ VerifyLayout(int klds, int mcap, int ncap) {
    auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
    size_t o = 0;
    M = o; o += al(sizeof(double) * (size_t)klds * klds);
    S = o; o += al(sizeof(int) * mcap);      // basic structural ids (column order of M)
    P = o; o += al(sizeof(int) * mcap);      // rows with nonbasic slack (row order of M), then Q
    pos = o; o += al(sizeof(int) * mcap);    // LU row permutation
    rhs = o; o += al(sizeof(double) * mcap); // x_S
    y = o; o += al(sizeof(double) * mcap);   // y_P
    d = o; o += al(sizeof(double) * mcap);   // ray
    inb = o; o += al((size_t)mcap + ncap + 16);
    total = o;
  }
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return uniform(v);
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return uniform(v);
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return uniform(v);
}

// M (k x k, row-major, stride k) := P L U in place; perm[c] = row swapped into c.
// Returns false when a pivot is below 1e-13 * max|M|.
__device__ bool lu64(double* Mx, int k, int* perm, double amax) {
  const int lane = threadIdx.x;
  for (int c = 0; c < k; ++c) {
    double bv = DBL_MAX;
    int bk = INT_MAX, bp = -1;
    for (int i = c + lane; i < k; i += 64) {
      const double v = -fabs(Mx[i * k + c]);
      if (better(v, i, bv, bk)) { bv = v; bk = i; bp = i; }
    }
    wave_argmin(bv, bk, bp);
    if (!(-bv > 1e-13 * (amax > 0 ? amax : 1.0))) return false;
    if (lane == 0) perm[c] = bp;
    if (bp != c)
      for (int j = lane; j < k; j += 64) {
        const double t = Mx[c * k + j];
        Mx[c * k + j] = Mx[bp * k + j];
        Mx[bp * k + j] = t;
      }
    __syncthreads();
    const double dc = uniform(Mx[c * k + c]);
    for (int i = c + 1 + lane; i < k; i += 64) Mx[i * k + c] /= dc;
    __syncthreads();
    // trailing update, lanes tiled 8 x 8 over (row, column)
    const int li = lane >> 3, lj = lane & 7;
    for (int i = c + 1 + li; i < k; i += 8) {
      const double l = Mx[i * k + c];
      if (l == 0.0) continue;
      for (int j = c + 1 + lj; j < k; j += 8) Mx[i * k + j] = fma(-l, Mx[c * k + j], Mx[i * k + j]);
    }
    __syncthreads();
  }
  return true;
}

// z := M^-1 z  (z in LDS, k entries, row order of M)
__device__ void lu_solve64(const double* Mx, int k, const int* perm, double* z) {
  const int lane = threadIdx.x;
  if (lane == 0)
    for (int c = 0; c < k; ++c)
      if (perm[c] != c) { const double t = z[c]; z[c] = z[perm[c]]; z[perm[c]] = t; }
  __syncthreads();
  for (int c = 0; c < k; ++c) {
    const double zc = uniform(z[c]);
    for (int i = c + 1 + lane; i < k; i += 64) z[i] = fma(-Mx[i * k + c], zc, z[i]);
    __syncthreads();
  }
  for (int c = k - 1; c >= 0; --c) {
    const double zc = uniform(z[c]) / uniform(Mx[c * k + c]);
    __syncthreads();
    if (lane == 0) z[c] = zc;
    for (int i = lane; i < c; i += 64) z[i] = fma(-Mx[i * k + c], zc, z[i]);
    __syncthreads();
  }
}

// z := M^-T z
__device__ void lu_solve_t64(const double* Mx, int k, const int* perm, double* z) {
  const int lane = threadIdx.x;
  for (int c = 0; c < k; ++c) {  // U^T w = z
    const double zc = uniform(z[c]) / uniform(Mx[c * k + c]);
    __syncthreads();
    if (lane == 0) z[c] = zc;
    for (int i = c + 1 + lane; i < k; i += 64) z[i] = fma(-Mx[c * k + i], zc, z[i]);
    __syncthreads();
  }
  for (int c = k - 1; c >= 0; --c) {  // L^T u = w
    const double zc = uniform(z[c]);
    for (int i = lane; i < c; i += 64) z[i] = fma(-Mx[c * k + i], zc, z[i]);
    __syncthreads();
  }
  if (lane == 0)
    for (int c = k - 1; c >= 0; --c)
      if (perm[c] != c) { const double t = z[c]; z[c] = z[perm[c]]; z[perm[c]] = t; }
  __syncthreads();
}

}  // namespace

__global__ __launch_bounds__(64) void k_verify(const LpDesc* __restrict__ desc,
                                               const double* __restrict__ pool,
                                               const int8_t* __restrict__ spool,
                                               const int32_t* __restrict__ idx, int32_t count,
                                               const int32_t* count_dev, Out out, Check chk,
                                               double tol, int fp32_source, double* ws,
                                               int64_t ws_stride, int mcap, int ncap,
                                               int32_t* queue, int32_t* fail_list,
                                               int32_t* fail_count, int klds,
                                               int32_t* big_list, int32_t* big_count) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_q, s_k;
  const VerifyLayout L(klds, mcap, ncap);
  double* Mlds = (double*)(smem + L.M);
  double* Mws = ws + (int64_t)blockIdx.x * ws_stride;
  int* Sl = (int*)(smem + L.S);
  int* Pl = (int*)(smem + L.P);
  int* perm = (int*)(smem + L.pos);
  double* xs = (double*)(smem + L.rhs);
  double* yp = (double*)(smem + L.y);
  double* dv = (double*)(smem + L.d);
  uint8_t* inb = smem + L.inb;
  const int lane = threadIdx.x;
  if (count_dev) count = *count_dev;

  for (;;) {
    __syncthreads();
    if (lane == 0) s_q = atomicAdd(queue, 1);
    __syncthreads();
    const int q = uniform(s_q);
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
    auto sg = [&](int i) { return (sense && sense[i] == 2) ? -1.0 : 1.0; };
    auto is_eq = [&](int i) { return sense && sense[i] == 1; };

    double bmax = 0, cmax = 0;
    for (int i = lane; i < m; i += 64) bmax = fmax(bmax, fabs(b[i]));
    for (int j = lane; j < n; j += 64) cmax = fmax(cmax, fabs(c[j]));
    bmax = wave_max(bmax);
    cmax = wave_max(cmax);
    const double am = out.amax[lp];
    bool defer = false;
    int kx = -1;  // size of the FP64 basic solution in xs, once computed

    int ok = 2;
    double rp = -1, rd = -1, zobj = out.obj[lp];
    const bool checkable = status == kStatusOptimal || status == kStatusUnbounded;
    if (!checkable) {
      if (fp32_source)
        ok = (status == kStatusInfeasible && out.wfin[lp] < -1e-3 * (1.0 + bmax)) ? 1 : 0;
    } else {
      // ---- split the basis: S (structurals), P (rows with nonbasic slack), Q ----
      for (int v = lane; v < n + m; v += 64) inb[v] = 0;
      __syncthreads();
      int bad = 0;
      for (int kk = lane; kk < m; kk += 64) {
        const int v = basis[kk];
        if (v < 0 || v >= n + m) bad = 1;
        else inb[v] = 1;
      }
      bad = wave_sum_i(bad);
      __syncthreads();
      if (lane == 0) {
        int ns = 0, np = 0;
        for (int j = 0; j < n; ++j) if (inb[j]) Sl[ns++] = j;
        for (int i = 0; i < m; ++i) if (!inb[n + i]) Pl[np++] = i;
        int nq = np;
        for (int i = 0; i < m; ++i) if (inb[n + i]) Pl[nq++] = i;
        s_k = (ns == np) ? ns : -1;
      }
      __syncthreads();
      const int k = uniform(s_k);
      if (bad || k < 0) {
        ok = 0;
      } else if (k > klds && big_list) {
        defer = true;  // checked by the second pass with a larger LDS block
      } else {
        auto check = [&](double* Mx) {
        // ---- M = A'[P, S] and its LU ----
        double mmax = 0;
        for (int e = lane; e < k * k; e += 64) {
          const int a = e / k, bcol = e - (e / k) * k;
          const int i = Pl[a];
          const double v = sg(i) * A[(int64_t)i * n + Sl[bcol]];
          Mx[e] = v;
          mmax = fmax(mmax, fabs(v));
        }
        mmax = wave_max(mmax);
        __syncthreads();
        if (k > 0 && !lu64(Mx, k, perm, mmax)) {
          ok = 0;
        } else {
          // x_S = M^-1 b'_P
          for (int a = lane; a < k; a += 64) xs[a] = sg(Pl[a]) * b[Pl[a]];
          __syncthreads();
          if (k > 0) lu_solve64(Mx, k, perm, xs);
          // y_P = M^-T c_S
          for (int a = lane; a < k; a += 64) yp[a] = c[Sl[a]];
          __syncthreads();
          if (k > 0) lu_solve_t64(Mx, k, perm, yp);
          double xm = 0, ym = 0;
          for (int a = lane; a < k; a += 64) { xm = fmax(xm, fabs(xs[a])); ym = fmax(ym, fabs(yp[a])); }
          // primal: bounds of x_S, basic slacks s_Q, residual rows P
          double pv = 0;
          for (int a = lane; a < k; a += 64) pv = fmax(pv, -xs[a]);
          // residual rows: four independent partial sums keep several loads in flight
          auto row_dot = [&](int i) {
            const double* Ai = A + (int64_t)i * n;
            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            int a = 0;
            for (; a + 4 <= k; a += 4) {
              s0 = fma(Ai[Sl[a]], xs[a], s0);
              s1 = fma(Ai[Sl[a + 1]], xs[a + 1], s1);
              s2 = fma(Ai[Sl[a + 2]], xs[a + 2], s2);
              s3 = fma(Ai[Sl[a + 3]], xs[a + 3], s3);
            }
            for (; a < k; ++a) s0 = fma(Ai[Sl[a]], xs[a], s0);
            return (s0 + s1) + (s2 + s3);
          };
          for (int qq = k + lane; qq < m; qq += 64) {
            const int i = Pl[qq];
            const double s = sg(i) * (b[i] - row_dot(i));
            xm = fmax(xm, fabs(s));
            pv = fmax(pv, is_eq(i) ? fabs(s) : -s);
          }
          for (int a = lane; a < k; a += 64) {
            const int i = Pl[a];
            const double r = sg(i) * (b[i] - row_dot(i));
            pv = fmax(pv, fabs(r));
          }
          xm = wave_max(xm);
          ym = wave_max(ym);
          rp = wave_max(pv) / (1.0 + bmax + am * xm);
          // dual: reduced costs of nonbasic structurals and '<=' slacks
          double dvv = 0;
          for (int j = lane; j < n; j += 64) {
            if (inb[j]) continue;
            double r0 = 0, r1 = 0, r2 = 0, r3 = 0;
            int a = 0;
            for (; a + 4 <= k; a += 4) {
              r0 = fma(yp[a] * sg(Pl[a]), A[(int64_t)Pl[a] * n + j], r0);
              r1 = fma(yp[a + 1] * sg(Pl[a + 1]), A[(int64_t)Pl[a + 1] * n + j], r1);
              r2 = fma(yp[a + 2] * sg(Pl[a + 2]), A[(int64_t)Pl[a + 2] * n + j], r2);
              r3 = fma(yp[a + 3] * sg(Pl[a + 3]), A[(int64_t)Pl[a + 3] * n + j], r3);
            }
            for (; a < k; ++a) r0 = fma(yp[a] * sg(Pl[a]), A[(int64_t)Pl[a] * n + j], r0);
            dvv = fmax(dvv, c[j] - ((r0 + r1) + (r2 + r3)));
          }
          for (int a = lane; a < k; a += 64)
            if (!is_eq(Pl[a])) dvv = fmax(dvv, -yp[a]);
          rd = wave_max(dvv) / (1.0 + cmax + am * ym);
          // objective of the FP64 basic solution (x itself is written once the check passed)
          kx = k;
          double zp = 0;
          for (int a = lane; a < k; a += 64) zp = fma(c[Sl[a]], xs[a], zp);
          const double zsum = wave_sum(zp);
          if (status == kStatusOptimal) {
            zobj = zsum;
            ok = (rp <= tol && rd <= tol) ? 1 : 0;
          } else {
            // ray of the entering variable v: d = B^-1 a'_v
            const int v = out.unb[lp];
            bool good = v >= 0 && v < n + m && !inb[v] && rp <= tol;
            if (good) {
              for (int a = lane; a < k; a += 64) {
                const int i = Pl[a];
                dv[a] = v < n ? sg(i) * A[(int64_t)i * n + v] : (v - n == i ? 1.0 : 0.0);
              }
              __syncthreads();
              if (k > 0) lu_solve64(Mx, k, perm, dv);
              double viol = 0, dmax = 0;
              for (int a = lane; a < k; a += 64) { viol = fmax(viol, dv[a]); dmax = fmax(dmax, fabs(dv[a])); }
              for (int qq = k + lane; qq < m; qq += 64) {
                const int i = Pl[qq];
                double s = v < n ? sg(i) * A[(int64_t)i * n + v] : (v - n == i ? 1.0 : 0.0);
                for (int a = 0; a < k; ++a) s = fma(-sg(i) * A[(int64_t)i * n + Sl[a]], dv[a], s);
                viol = fmax(viol, is_eq(i) ? fabs(s) : s);
                dmax = fmax(dmax, fabs(s));
              }
              viol = wave_max(viol);
              dmax = wave_max(dmax);
              double part = 0;
              if (v < n) {
                for (int a = lane; a < k; a += 64) part = fma(yp[a], sg(Pl[a]) * A[(int64_t)Pl[a] * n + v], part);
              } else {
                for (int a = lane; a < k; a += 64) if (Pl[a] == v - n) part = yp[a];
              }
              const double rv = (v < n ? c[v] : 0.0) - wave_sum(part);
              good = viol <= tol * (1.0 + dmax) && rv > tol * (1.0 + cmax + am * ym);
            }
            ok = good ? 1 : 0;
            rd = -1;
          }
        }
        };
        if (k <= klds) check(Mlds);
        else check(Mws);
      }
    }
    __syncthreads();
    if (defer) {
      if (lane == 0) big_list[atomicAdd(big_count, 1)] = lp;
      continue;
    }
    // a verified result reports the FP64 basic solution: x_S from the LU solve, 0 elsewhere
    if (ok == 1 && kx >= 0) {
      double* x = out.x + D.x;
      for (int j = lane; j < n; j += 64) x[j] = 0.0;
      __syncthreads();
      for (int a = lane; a < kx; a += 64) x[Sl[a]] = xs[a];
    }
    if (lane == 0) {
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
      for (int i = lane; i < m; i += 64) bw[i] = -1;
    }
  }
}

size_t verify_lds_bytes(int mcap, int ncap, int klds) { return VerifyLayout(klds, mcap, ncap).total; }

int verify_max_blocks_per_cu(int mcap, int ncap, int klds) {
  int nb = 0;
  const void* fn = reinterpret_cast<const void*>(&k_verify);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 64, verify_lds_bytes(mcap, ncap, klds)) !=
      hipSuccess)
    return 0;
  return nb;
}

hipError_t launch_verify(const VerifyLaunch& V, hipStream_t st) {
  const size_t bytes = verify_lds_bytes(V.mcap, V.ncap, V.klds);
  const void* fn = reinterpret_cast<const void*>(&k_verify);
  void* args[] = {(void*)&V.desc, (void*)&V.pool, (void*)&V.spool, (void*)&V.idx,
                  (void*)&V.count, (void*)&V.count_dev, (void*)&V.out, (void*)&V.chk,
                  (void*)&V.tol, (void*)&V.fp32_source, (void*)&V.ws, (void*)&V.ws_stride,
                  (void*)&V.mcap, (void*)&V.ncap, (void*)&V.queue, (void*)&V.fail_list,
                  (void*)&V.fail_count, (void*)&V.klds, (void*)&V.big_list, (void*)&V.big_count};
  return hipLaunchKernel(fn, dim3(V.grid), dim3(64), args, bytes, st);
}

}  // namespace blpg
