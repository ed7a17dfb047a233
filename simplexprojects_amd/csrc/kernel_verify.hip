// kernel_verify.hip -- FP64 check of final bases by iterative refinement (no reference
// counterpart: the reference returns only T[M-2][N-1], simplex.cu:380, and never checks it).
//
// The solve kernel dumps its final tableau (solve precision) and the nonbasic variable of
// every column.  With A', b' the LP with '>=' rows negated and B the basis matrix of
// [A' I], the tableau holds B^-1 column by column: the column of a nonbasic slack s_p is
// B^-1 e_p, and a basic slack's column of B^-1 is the unit vector of its row.  So no
// factorisation is needed:
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
//   UNBOUNDED a feasible point (the origin when b' >= 0 on inequality rows and b = 0 on
//             '=' rows, else the final basis, which must pass the primal test) and the
//             entering variable e's ray d = B^-1 a'_e (refined the same way) with
//             d <= tol (1 + |d|) (|d| on basic '=' slacks), residual <= tol (1 +
//             |A|_max |d|), and gain c_e - c_B^T d > tol (1 + |c|_inf (1 + |d|_inf)).
//   INFEASIBLE (FP32, phase I ended with x0 basic) the phase-I duals y, read from row 1 at
//             the slack columns and refined like y above against the auxiliary problem's
//             costs (0 everywhere but x0, whose own row is left out), form a Farkas
//             certificate: y >= 0 on inequality rows (clamped), y^T A'_j >= -tol |A|_max
//             |y|_inf for every structural j and y^T b' <= -sqrt(tol) (1 + |b|_inf) |y|_inf;
//             infeasibility found by the equality stage (a row with no usable pivot) has
//             no such certificate and is re-solved.
// FP32 results that fail (or have iteration-limit / numeric statuses) are appended to the
// fallback list.
//
// Every sweep touches only the k basic structural columns of A' and the k nonbasic slack
// columns of the tableau (k = number of basic structurals).  Both m x k blocks are staged
// once into LDS (row-major, odd row stride: row sweeps and column sweeps are both free of
// bank conflicts), so the sweeps run from LDS.  Up to three tiers run in sequence on a
// bin, each with a larger LDS budget per LP (fewer LPs per CU) and each taking the LPs the
// previous one deferred: k_verify<128> with 31 KB, k_verify<128> with 78 KB, and
// k_verify<256> with 150 KB, where an LP that still does not fit runs its sweeps from
// global memory (rows batched per wavefront).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels.hpp"
#include "simplex_device.hpp"

namespace blpg {

namespace {

constexpr int kMaxSweeps = 12;
constexpr int kRB = 8;  // rows per wavefront pass of the global-memory row sweeps

__host__ __device__ inline int part_cap(int mcap, int nt) { return mcap + 1 > nt ? mcap + 1 : nt; }

// per-LP vectors (LDS when they fit, else a slab per workgroup in HBM) for NT threads.
// The double arrays are 16-byte aligned, the int arrays (read one element at a time) only
// 4-byte aligned, and the total is rounded to 16 for the staged blocks that follow: at
// 105x103 the 16-byte padding of the int arrays alone kept LPs with 16 or 17 basic
// structurals out of the 31 KB tier.
struct VLayout {
  size_t x, r, y, rho, d, w, bp, ae, cS, part, bs, rowof, colp, S, Srow, Pn, Pcol, total;
  __host__ __device__ VLayout(int mcap, int ncap, int nt) {
    auto al = [](size_t v) { return (v + 15) & ~(size_t)15; };
    auto al4 = [](size_t v) { return (v + 3) & ~(size_t)3; };
    size_t o = 0;
    x = o; o += al(8 * (size_t)mcap);      // basic values by row
    r = o; o += al(8 * (size_t)mcap);      // residual by constraint
    y = o; o += al(8 * (size_t)mcap);      // duals by constraint
    rho = o; o += al(8 * (size_t)mcap);    // dual residual by row
    d = rho;                               // ray by row (after the duals are done)
    w = o; o += al(8 * (size_t)mcap);      // gathered vector of the current sweep
    bp = o; o += al(8 * (size_t)mcap);     // b' (signed rhs)
    ae = bp;                               // a'_e, the entering column (after x is done)
    cS = o; o += al(8 * (size_t)mcap);     // c of the basic structurals (S order)
    part = o;                              // column-sum partials (part_cap doubles)
    o += al(8 * (size_t)part_cap(mcap, nt));
    bs = o; o += al4(4 * (size_t)mcap);    // basic variable of each row
    rowof = o; o += al4(4 * (size_t)(mcap + ncap + 1));  // row of each basic variable, -1
    colp = o; o += al4(4 * (size_t)mcap);  // tableau column of each nonbasic slack, -1
    S = o; o += al4(4 * (size_t)mcap);     // basic structurals (ascending)
    Srow = o; o += al4(4 * (size_t)mcap);  // their rows
    Pn = o; o += al4(4 * (size_t)mcap);    // nonbasic slacks (ascending constraint index)
    Pcol = o; o += al4(4 * (size_t)mcap);  // their tableau columns
    total = al(o);
  }
};

// LDS bytes of the staged blocks: A'[:, S] (FP64) and T[2.., Pcol] (tableau precision)
__host__ __device__ inline size_t stage_bytes(int m, int k, int tsz) {
  return (size_t)m * (size_t)(k | 1) * (size_t)(8 + tsz);
}

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

// workgroup reductions through `red` (NT/64 doubles), uniform result
template <int NT>
__device__ __forceinline__ double bmax(double v, double* red) {
  v = wmax(v);
  __syncthreads();
  if constexpr (NT == 64) {
    return v;
  } else {
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) r = fmax(r, red[w]);
    return r;
  }
}
// two maxima in one reduction (red: 2 * NT/64 doubles)
template <int NT>
__device__ __forceinline__ void bmax2(double& a, double& b, double* red) {
  a = wmax(a);
  b = wmax(b);
  __syncthreads();
  if constexpr (NT != 64) {
    if ((threadIdx.x & 63) == 0) {
      red[threadIdx.x >> 6] = a;
      red[NT / 64 + (threadIdx.x >> 6)] = b;
    }
    __syncthreads();
    double ra = red[0], rb = red[NT / 64];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) {
      ra = fmax(ra, red[w]);
      rb = fmax(rb, red[NT / 64 + w]);
    }
    a = ra;
    b = rb;
  }
}
template <int NT>
__device__ __forceinline__ double bsum(double v, double* red) {
  v = wsum(v);
  __syncthreads();
  if constexpr (NT == 64) {
    return v;
  } else {
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) r += red[w];
    return r;
  }
}

// compacts {i < cnt : pred(i)} in ascending order into out (and key(i) into out2);
// returns the count (uniform).  Ballot prefix sums over NT-wide chunks.
template <int NT, class Pred, class Key>
__device__ int compact(int cnt, Pred pred, Key key, int* out, int* out2, int* red_i) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int base = 0;
  for (int c0 = 0; c0 < cnt; c0 += NT) {
    const int i = c0 + tid;
    const bool p = i < cnt && pred(i);
    const unsigned long long bal = __ballot(p);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    int off = base, tot = 0;
    if constexpr (NT == 64) {
      tot = __popcll(bal);
    } else {
      __syncthreads();
      if (lane == 0) red_i[w] = __popcll(bal);
      __syncthreads();
      for (int q = 0; q < w; ++q) off += red_i[q];
      for (int q = 0; q < NT / 64; ++q) tot += red_i[q];
    }
    if (p) {
      out[off + before] = i;
      out2[off + before] = key(i);
    }
    base += tot;
  }
  __syncthreads();
  return base;
}

}  // namespace

// Checks one LP with NT threads.  wsp: the per-LP vectors; stg: `stage_cap` bytes of LDS
// for the staged blocks.  Returns false (and writes nothing) when the blocks do not fit
// and `defer` is set or NT < 256: the caller hands the LP to the four-wavefront kernel.
template <typename TT, int NT>
__device__ bool verify_lp(int lp, const LpDesc& D, const double* __restrict__ pool,
                          const int8_t* __restrict__ spool, Out out, Check chk, double tol,
                          int fp32_source, char* wsp, char* stg, int64_t stage_cap, bool defer,
                          double* red, int* red_i, int32_t* fail_list, int32_t* fail_count) {
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
  const VLayout L(m, n, NT);
  double* x = (double*)(wsp + L.x);
  double* r = (double*)(wsp + L.r);
  double* y = (double*)(wsp + L.y);
  double* rho = (double*)(wsp + L.rho);
  double* dv = (double*)(wsp + L.d);
  double* w = (double*)(wsp + L.w);
  double* bp = (double*)(wsp + L.bp);
  double* ae = (double*)(wsp + L.ae);
  double* cS = (double*)(wsp + L.cS);
  double* part = (double*)(wsp + L.part);
  int* bs = (int*)(wsp + L.bs);
  int* rowof = (int*)(wsp + L.rowof);
  int* colp = (int*)(wsp + L.colp);
  int* S = (int*)(wsp + L.S);
  int* Srow = (int*)(wsp + L.Srow);
  int* Pn = (int*)(wsp + L.Pn);
  int* Pcol = (int*)(wsp + L.Pcol);

  // an FP32 INFEASIBLE result is certified by the phase-I duals (Farkas), see below
  const bool infeas = status == kStatusInfeasible && fp32_source;
  const bool checkable =
      (status == kStatusOptimal || status == kStatusUnbounded || infeas) && out.tab;
  int k = 0, kq = 0, r0 = -1, bad = 0;
  if (checkable) {
    // ---- basis bookkeeping (the auxiliary x0, id n+m, is basic at the end of an
    //      infeasible phase I) ----
    int kl = 0, xl = 0;
    for (int i = tid; i < m; i += NT) {
      const int v = basis[i];
      bs[i] = v;
      kl += (v >= 0 && v < n) ? 1 : 0;
      xl |= v == n + m ? 1 : 0;
    }
    if (defer || NT < 256) {
      // an LP this tier hands on is decided from its basis alone (the size of its staged
      // blocks), before the rest of the bookkeeping
      const int ke = (int)bsum<NT>((double)kl, red) + ((infeas && bmax<NT>((double)xl, red) > 0) ? 1 : 0);
      if (stage_bytes(m, ke, (int)sizeof(TT)) > (size_t)stage_cap) return false;
    }
    for (int v = tid; v < n + m + 1; v += NT) rowof[v] = -1;
    for (int p = tid; p < m; p += NT) colp[p] = -1;
    __syncthreads();
    for (int i = tid; i < m; i += NT) {
      const int v = bs[i];
      if (v < 0 || v >= n + m + (infeas ? 1 : 0)) bad = 1;
      else rowof[v] = i;
    }
    for (int j = 1 + tid; j < C; j += NT) {
      const int v = nbv[j];
      if (v >= n && v < n + m) colp[v - n] = j;
    }
    bad = (int)bmax<NT>((double)bad, red);
    k = compact<NT>(n, [&](int j) { return rowof[j] >= 0; }, [&](int j) { return rowof[j]; },
                    S, Srow, red_i);
    kq = compact<NT>(m, [&](int p) { return colp[p] >= 0; },
                     [&](int p) { return colp[p]; }, Pn, Pcol, red_i);
    if (infeas) {
      r0 = rowof[n + m];
      // (infeasibility found by the equality stage leaves x0 out: no certificate here)
      if (r0 < 0) bad = 1;
    }
    if (kq != k + (r0 >= 0 ? 1 : 0)) bad = 1;
  }
  const int kk = kq > k ? kq : k;  // columns of the staged blocks
  const bool fits = checkable && !bad && stage_bytes(m, kk, (int)sizeof(TT)) <= (size_t)stage_cap;
  if (checkable && !bad && !fits && (defer || NT < 256)) return false;
  // the first-pass kernel only ever runs staged (its global-memory sweeps are not built)
  const bool staged = NT < 256 || fits;

  double bm = 0, cm = 0;
  for (int i = tid; i < m; i += NT) bm = fmax(bm, fabs(b[i]));
  for (int j = tid; j < n; j += NT) cm = fmax(cm, fabs(c[j]));
  bmax2<NT>(bm, cm, red);
  const double am = out.amax[lp];
  const double sq_tol = sqrt(tol);

  int ok = 2;
  double rp = -1, rd = -1, zobj = out.obj[lp];
  int kx = -1;
  // warm-start a fallback from this basis (the crash skips x0's id): an optimal one, a
  // phase-I basis that ended with x0 basic (3-4x fewer FP64 pivots than a cold re-solve
  // on the two-phase LPs whose certificate does not refine), or (below) an unbounded one
  // whose basic solution refined
  bool keep_basis = status == kStatusOptimal || (infeas && r0 >= 0);
  bool undo_last = false;  // the final basis did not refine: warm-start before its last pivot
  if (!checkable) {
    if (fp32_source) ok = 0;  // iteration limit, numeric trouble: re-solved in FP64
  } else if (bad) {
    ok = 0;
  } else {
    const int kpad = kk | 1;
    double* AB = (double*)stg;                         // m x kpad: A'[p][S[a]]
    TT* TB = (TT*)(stg + (size_t)m * kpad * 8);        // m x kpad: T[2+i][Pcol[a]]
    for (int p = tid; p < m; p += NT) bp[p] = sg(p) * b[p];
    for (int a = tid; a < k; a += NT) cS[a] = c[S[a]];
    if (staged && kk > 0) {
      // one pass of gathers; thread (tp, ta) copies rows tp, tp + per, ... of column ta,
      // eight rows per step with all loads of a step in flight together.  128 threads: the
      // last step's rows past m re-read row m - 1 and are not stored; 256 threads (whose
      // registers the clamped step would spill): the rows left after the full steps one
      // at a time.  (With x0 basic the T block has one column more than the A block; its
      // A column a = k is a copy of column k-1, or of column 0 of A, never read.)
      const int per = kk >= NT ? 1 : NT / kk;
      const int ta = kk >= NT ? tid : tid % kk, tp = kk >= NT ? 0 : tid / kk;
      if (tp < per) {
        for (int a = ta; a < kk; a += NT) {
          const int ca = a < k ? S[a] : (k > 0 ? S[k - 1] : 0), cp = Pcol[a];
          if constexpr (NT < 256) {
            for (int p = tp; p < m; p += 8 * per) {
              double av[8];
              TT tv[8];
              int8_t sv[8];
#pragma unroll
              for (int u = 0; u < 8; ++u) {
                const int pu = min(p + u * per, m - 1);
                av[u] = A[(int64_t)pu * n + ca];
                tv[u] = T[(int64_t)(2 + pu) * C + cp];
                sv[u] = sense ? sense[pu] : (int8_t)0;
              }
#pragma unroll
              for (int u = 0; u < 8; ++u) {
                const int pu = p + u * per;
                if (pu < m) {
                  AB[pu * kpad + a] = (sv[u] == 2 ? -1.0 : 1.0) * av[u];
                  TB[pu * kpad + a] = tv[u];
                }
              }
            }
          } else {
            int p = tp;
            for (; p + 7 * per < m; p += 8 * per) {
              double av[8];
              TT tv[8];
#pragma unroll
              for (int u = 0; u < 8; ++u) {
                av[u] = A[(int64_t)(p + u * per) * n + ca];
                tv[u] = T[(int64_t)(2 + p + u * per) * C + cp];
              }
#pragma unroll
              for (int u = 0; u < 8; ++u) {
                AB[(p + u * per) * kpad + a] = sg(p + u * per) * av[u];
                TB[(p + u * per) * kpad + a] = tv[u];
              }
            }
            for (; p < m; p += per) {
              AB[p * kpad + a] = sg(p) * A[(int64_t)p * n + ca];
              TB[p * kpad + a] = T[(int64_t)(2 + p) * C + cp];
            }
          }
        }
      }
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;

    // r = rhs - B v for the basis matrix B of [A' I]; rhs is b' (ray == false) or the
    // entering column a'_e.  Returns the thread's max |r| (reduced by the caller).
    auto resid = [&](const double* vb, bool ray, int e) {
      double rl = 0;
      if (staged) {
        // one thread per constraint row over the staged block
        for (int a = tid; a < k; a += NT) w[a] = vb[Srow[a]];
        __syncthreads();
        for (int p = tid; p < m; p += NT) {
          const double* ABp = AB + p * kpad;
          double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
          int a = 0;
          for (; a + 4 <= k; a += 4) {
            s0 = fma(ABp[a], w[a], s0);
            s1 = fma(ABp[a + 1], w[a + 1], s1);
            s2 = fma(ABp[a + 2], w[a + 2], s2);
            s3 = fma(ABp[a + 3], w[a + 3], s3);
          }
          for (; a < k; ++a) s0 = fma(ABp[a], w[a], s0);
          const int q = rowof[n + p];
          const double rhs = ray ? ae[p] : bp[p];
          const double val = rhs - ((s0 + s1) + (s2 + s3)) - (q >= 0 ? vb[q] : 0.0);
          r[p] = val;
          rl = fmax(rl, fabs(val));
        }
      } else {
        // from global memory: each wavefront takes kRB rows at a time (kRB independent
        // loads in flight per lane), lanes over the basic structurals (ascending, so a
        // row's loads stay within a few cache lines)
        for (int p0 = wave * kRB; p0 < m; p0 += (NT / 64) * kRB) {
          double s[kRB];
#pragma unroll
          for (int u = 0; u < kRB; ++u) s[u] = 0;
          for (int a = lane; a < k; a += 64) {
            const int col = S[a];
            const double xv = vb[Srow[a]];
#pragma unroll
            for (int u = 0; u < kRB; ++u) {
              const int p = min(p0 + u, m - 1);
              s[u] = fma(A[(int64_t)p * n + col], xv, s[u]);
            }
          }
#pragma unroll
          for (int u = 0; u < kRB; ++u) {
            const double tot = wsum(s[u]);
            const int p = p0 + u;
            if (p < m) {
              const int q = rowof[n + p];
              const double rhs = ray ? ae[p] : bp[p];
              const double val = rhs - sg(p) * tot - (q >= 0 ? vb[q] : 0.0);
              if (lane == 0) r[p] = val;
              rl = fmax(rl, fabs(val));
            }
          }
        }
      }
      return rl;
    };
    // vb += B~^-1 r: row i of B~^-1 is the tableau row's slack columns, plus 1 at the
    // row's own basic slack
    auto correct = [&](double* vb) {
      if (staged) {
        for (int a = tid; a < k; a += NT) w[a] = r[Pn[a]];
        __syncthreads();
        for (int i = tid; i < m; i += NT) {
          const TT* TBi = TB + i * kpad;
          double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
          int a = 0;
          for (; a + 4 <= k; a += 4) {
            s0 = fma((double)TBi[a], w[a], s0);
            s1 = fma((double)TBi[a + 1], w[a + 1], s1);
            s2 = fma((double)TBi[a + 2], w[a + 2], s2);
            s3 = fma((double)TBi[a + 3], w[a + 3], s3);
          }
          for (; a < k; ++a) s0 = fma((double)TBi[a], w[a], s0);
          const int v = bs[i];
          vb[i] += ((s0 + s1) + (s2 + s3)) + (v >= n ? r[v - n] : 0.0);
        }
      } else {
        for (int i0 = wave * kRB; i0 < m; i0 += (NT / 64) * kRB) {
          double s[kRB];
#pragma unroll
          for (int u = 0; u < kRB; ++u) s[u] = 0;
          for (int a = lane; a < k; a += 64) {
            const int col = Pcol[a];
            const double rv = r[Pn[a]];
#pragma unroll
            for (int u = 0; u < kRB; ++u) {
              const int i = min(i0 + u, m - 1);
              s[u] = fma((double)T[(int64_t)(2 + i) * C + col], rv, s[u]);
            }
          }
#pragma unroll
          for (int u = 0; u < kRB; ++u) {
            const double tot = wsum(s[u]);
            const int i = i0 + u;
            if (i < m && lane == 0) {
              const int v = bs[i];
              vb[i] += tot + (v >= n ? r[v - n] : 0.0);
            }
          }
        }
      }
      __syncthreads();
    };
    // column sums over a staged block: emit(a, sum_p M[p][a] v[p]) for a < cnt.  Thread
    // (tp, ta) sums rows tp, tp + per, ... of column ta (all NT threads busy when cnt < NT);
    // the per partials of a column are then added in order.
    auto colsum = [&](const auto* M, const double* v, int cnt, auto&& emit) {
      const int pc = part_cap(m, NT) / kpad;
      const int per = cnt >= NT || cnt == 0 ? 1 : (NT / cnt < pc ? NT / cnt : pc);
      const int ta = cnt >= NT || cnt == 0 ? tid : tid % cnt,
                tp = cnt >= NT || cnt == 0 ? 0 : tid / cnt;
      if (cnt > 0 && tp < per) {
        for (int a = ta; a < cnt; a += NT) {
          double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
          int p = tp;
          for (; p + 3 * per < m; p += 4 * per) {
            s0 = fma((double)M[p * kpad + a], v[p], s0);
            s1 = fma((double)M[(p + per) * kpad + a], v[p + per], s1);
            s2 = fma((double)M[(p + 2 * per) * kpad + a], v[p + 2 * per], s2);
            s3 = fma((double)M[(p + 3 * per) * kpad + a], v[p + 3 * per], s3);
          }
          for (; p < m; p += per) s0 = fma((double)M[p * kpad + a], v[p], s0);
          part[tp * kpad + a] = (s0 + s1) + (s2 + s3);
        }
      }
      __syncthreads();
      for (int a = tid; a < cnt; a += NT) {
        double sum = 0;
        for (int q = 0; q < per; ++q) sum += part[q * kpad + a];
        emit(a, sum);
      }
    };
    // refinement: stops at the rounding floor, after kMaxSweeps, or when a sweep no
    // longer shrinks the residual (a basis too ill-conditioned for the tableau's inverse)
    auto refine = [&](double* vb, bool ray, int e, double scale0, double& res, double& vm) {
      double prev = DBL_MAX;
      for (int sweep = 0;; ++sweep) {
        res = resid(vb, ray, e);
        vm = 0;
        for (int i = tid; i < m; i += NT) vm = fmax(vm, fabs(vb[i]));
        bmax2<NT>(res, vm, red);
        if (sweep == kMaxSweeps || res <= 1e-14 * (scale0 + am * vm) || res > 0.9 * prev) break;
        prev = res;
        correct(vb);
      }
    };

    if (infeas) {
      // ---- infeasible: the phase-I duals y (row 1 at the nonbasic slack columns, 0 at
      //      basic slacks) refined against B^T y = c_B of the auxiliary problem, whose
      //      costs are 0 but at x0 (x0's column is not kept, so its row's residual is left
      //      out: the sweeps fix every other row and keep y's scale); then the Farkas test
      //      in FP64 on y clamped to >= 0 on inequality rows:
      //        (y^T A')_j >= -tol |A|_max |y|_inf for every structural j,
      //        y^T b' <= -sqrt(tol) (1 + |b|_inf) |y|_inf,
      //      which no x >= 0 with A'x <= b' (= on '=' rows) can meet, as
      //      0 <= y^T A' x <= y^T b' < 0. ----
      for (int p = tid; p < m; p += NT) y[p] = colp[p] >= 0 ? (double)T[C + colp[p]] : 0.0;
      __syncthreads();
      double ym = 0, rres_d = 0, prev = DBL_MAX;
      for (int sweep = 0;; ++sweep) {
        double rl = 0;
        if (staged) {
          colsum(AB, y, k, [&](int a, double sum) {
            rho[Srow[a]] = -sum;
            rl = fmax(rl, fabs(sum));
          });
        } else {
          for (int a = tid; a < k; a += NT) {
            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            int p = 0;
            const int v = S[a];
            for (; p + 4 <= m; p += 4) {
              s0 = fma(sg(p) * A[(int64_t)p * n + v], y[p], s0);
              s1 = fma(sg(p + 1) * A[(int64_t)(p + 1) * n + v], y[p + 1], s1);
              s2 = fma(sg(p + 2) * A[(int64_t)(p + 2) * n + v], y[p + 2], s2);
              s3 = fma(sg(p + 3) * A[(int64_t)(p + 3) * n + v], y[p + 3], s3);
            }
            for (; p < m; ++p) s0 = fma(sg(p) * A[(int64_t)p * n + v], y[p], s0);
            const double sum = (s0 + s1) + (s2 + s3);
            rho[Srow[a]] = -sum;
            rl = fmax(rl, fabs(sum));
          }
        }
        for (int p = tid; p < m; p += NT) {
          const int q = rowof[n + p];
          if (q >= 0) {
            rho[q] = -y[p];
            rl = fmax(rl, fabs(y[p]));
          }
        }
        if (tid == 0) rho[r0] = 0.0;
        rres_d = rl;
        ym = 0;
        for (int p = tid; p < m; p += NT) ym = fmax(ym, fabs(y[p]));
        bmax2<NT>(rres_d, ym, red);
        if (sweep == kMaxSweeps || rres_d <= 1e-14 * am * ym || rres_d > 0.9 * prev) break;
        prev = rres_d;
        if (staged) {
          colsum(TB, rho, kq, [&](int a, double sum) { y[Pn[a]] += sum; });
        } else {
          for (int a = tid; a < kq; a += NT) {
            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            int i = 0;
            const int j = Pcol[a];
            for (; i + 4 <= m; i += 4) {
              s0 = fma((double)T[(int64_t)(2 + i) * C + j], rho[i], s0);
              s1 = fma((double)T[(int64_t)(3 + i) * C + j], rho[i + 1], s1);
              s2 = fma((double)T[(int64_t)(4 + i) * C + j], rho[i + 2], s2);
              s3 = fma((double)T[(int64_t)(5 + i) * C + j], rho[i + 3], s3);
            }
            for (; i < m; ++i) s0 = fma((double)T[(int64_t)(2 + i) * C + j], rho[i], s0);
            y[Pn[a]] += (s0 + s1) + (s2 + s3);
          }
        }
        for (int p = tid; p < m; p += NT) {
          const int q = rowof[n + p];
          if (q >= 0) y[p] += rho[q];
        }
        __syncthreads();
      }
      // y clamped to >= 0 on inequality rows (w = sign-applied, so A w = A'^T y per column)
      double yneg = 0;
      for (int p = tid; p < m; p += NT) {
        const double yp = y[p];
        if (!is_eq(p)) yneg = fmax(yneg, -yp);
        w[p] = sg(p) * (is_eq(p) ? yp : fmax(yp, 0.0));
      }
      __syncthreads();
      double vg = 0;
      for (int j = tid; j < n; j += NT) {
        // one column per thread, 16 independent row loads in flight per step
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        int p = 0;
        for (; p + 16 <= m; p += 16) {
          double av[16];
  #pragma unroll
          for (int u = 0; u < 16; ++u) av[u] = A[(int64_t)(p + u) * n + j];
  #pragma unroll
          for (int u = 0; u < 16; u += 4) {
            s0 = fma(av[u], w[p + u], s0);
            s1 = fma(av[u + 1], w[p + u + 1], s1);
            s2 = fma(av[u + 2], w[p + u + 2], s2);
            s3 = fma(av[u + 3], w[p + u + 3], s3);
          }
        }
        for (; p < m; ++p) s0 = fma(A[(int64_t)p * n + j], w[p], s0);
        vg = fmax(vg, -((s0 + s1) + (s2 + s3)));
      }
      double hp = 0;
      for (int p = tid; p < m; p += NT) hp = fma(b[p], w[p], hp);
      bmax2<NT>(vg, yneg, red);
      const double h = bsum<NT>(hp, red);
      const double gs = am * ym;
      rp = fmax(vg / fmax(gs, DBL_MIN), yneg / fmax(ym, DBL_MIN));
      rd = -h / ((1.0 + bm) * fmax(ym, DBL_MIN));
      ok = (ym > 0 && vg <= tol * gs && yneg <= tol * ym && -h >= sq_tol * (1.0 + bm) * ym) ? 1 : 0;
    } else {
      // ---- x_B ----
      for (int i = tid; i < m; i += NT) x[i] = (double)T[(int64_t)(2 + i) * C];
      __syncthreads();
      double xm = 0, rres = 0;
      refine(x, false, -1, 1.0 + bm, rres, xm);
      double pv = 0;
      for (int i = tid; i < m; i += NT) {
        const int v = bs[i];
        const double xv = x[i];
        if (v < n) pv = fmax(pv, -xv);
        else pv = fmax(pv, is_eq(v - n) ? fabs(xv) : -xv);
      }
      pv = bmax<NT>(pv, red);
      rp = fmax(pv, rres) / (1.0 + bm + am * xm);
      const bool pconv = rres <= sq_tol * (1.0 + bm);
      keep_basis = keep_basis || pconv;
      undo_last = !pconv;

      // The duals decide optimal results only: an unbounded result is certified by a
      // feasible point and its ray (below), so its dual sweeps and the pass over A are skipped.
      bool dconv = false;
      if (status == kStatusOptimal) {
        // ---- y: rho = c_B - B^T y (one thread per basic structural), y += B~^-T rho (one
        //      thread per nonbasic slack: its tableau column against rho) ----
        for (int p = tid; p < m; p += NT) y[p] = colp[p] >= 0 ? (double)T[colp[p]] : 0.0;
        __syncthreads();
        double ym = 0, rres_d = 0, prev = DBL_MAX;
        for (int sweep = 0;; ++sweep) {
          double rl = 0;
          if (staged) {
            colsum(AB, y, k, [&](int a, double sum) {
              const double val = cS[a] - sum;
              rho[Srow[a]] = val;
              rl = fmax(rl, fabs(val));
            });
          } else {
            for (int a = tid; a < k; a += NT) {
              double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
              int p = 0;
              const int v = S[a];
              for (; p + 4 <= m; p += 4) {
                s0 = fma(sg(p) * A[(int64_t)p * n + v], y[p], s0);
                s1 = fma(sg(p + 1) * A[(int64_t)(p + 1) * n + v], y[p + 1], s1);
                s2 = fma(sg(p + 2) * A[(int64_t)(p + 2) * n + v], y[p + 2], s2);
                s3 = fma(sg(p + 3) * A[(int64_t)(p + 3) * n + v], y[p + 3], s3);
              }
              for (; p < m; ++p) s0 = fma(sg(p) * A[(int64_t)p * n + v], y[p], s0);
              const double val = cS[a] - ((s0 + s1) + (s2 + s3));
              rho[Srow[a]] = val;
              rl = fmax(rl, fabs(val));
            }
          }
          for (int p = tid; p < m; p += NT) {
            const int q = rowof[n + p];
            if (q >= 0) {
              rho[q] = -y[p];
              rl = fmax(rl, fabs(y[p]));
            }
          }
          rres_d = rl;
          ym = 0;
          for (int p = tid; p < m; p += NT) ym = fmax(ym, fabs(y[p]));
          bmax2<NT>(rres_d, ym, red);
          if (sweep == kMaxSweeps || rres_d <= 1e-14 * (1.0 + cm + am * ym) || rres_d > 0.9 * prev) break;
          prev = rres_d;
          if (staged) {
            colsum(TB, rho, k, [&](int a, double sum) { y[Pn[a]] += sum; });
          } else {
            for (int a = tid; a < k; a += NT) {
              double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
              int i = 0;
              const int j = Pcol[a];
              for (; i + 4 <= m; i += 4) {
                s0 = fma((double)T[(int64_t)(2 + i) * C + j], rho[i], s0);
                s1 = fma((double)T[(int64_t)(3 + i) * C + j], rho[i + 1], s1);
                s2 = fma((double)T[(int64_t)(4 + i) * C + j], rho[i + 2], s2);
                s3 = fma((double)T[(int64_t)(5 + i) * C + j], rho[i + 3], s3);
              }
              for (; i < m; ++i) s0 = fma((double)T[(int64_t)(2 + i) * C + j], rho[i], s0);
              y[Pn[a]] += (s0 + s1) + (s2 + s3);
            }
          }
          for (int p = tid; p < m; p += NT) {
            const int q = rowof[n + p];
            if (q >= 0) y[p] += rho[q];
          }
          __syncthreads();
        }
        // dual feasibility: nonbasic structurals (one pass over A, coalesced across j) and
        // nonbasic '<=' slacks
        for (int p = tid; p < m; p += NT) w[p] = sg(p) * y[p];
        __syncthreads();
        double dvv = 0;
        for (int j = tid; j < n; j += NT) {
          // one column per thread, 16 independent row loads in flight per step
          double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
          int p = 0;
          for (; p + 16 <= m; p += 16) {
            double av[16];
    #pragma unroll
            for (int u = 0; u < 16; ++u) av[u] = A[(int64_t)(p + u) * n + j];
    #pragma unroll
            for (int u = 0; u < 16; u += 4) {
              s0 = fma(av[u], w[p + u], s0);
              s1 = fma(av[u + 1], w[p + u + 1], s1);
              s2 = fma(av[u + 2], w[p + u + 2], s2);
              s3 = fma(av[u + 3], w[p + u + 3], s3);
            }
          }
          for (; p < m; ++p) s0 = fma(A[(int64_t)p * n + j], w[p], s0);
          if (rowof[j] < 0) dvv = fmax(dvv, c[j] - ((s0 + s1) + (s2 + s3)));
        }
        for (int p = tid; p < m; p += NT)
          if (colp[p] >= 0 && !is_eq(p)) dvv = fmax(dvv, -y[p]);
        rd = fmax(bmax<NT>(dvv, red), rres_d) / (1.0 + cm + am * ym);
        dconv = rres_d <= sq_tol * (1.0 + cm);
      }

      // objective of the FP64 basic solution
      double zp = 0;
      for (int a = tid; a < k; a += NT) zp = fma(c[S[a]], x[Srow[a]], zp);
      const double zsum = bsum<NT>(zp, red);
      kx = k;
      if (status == kStatusOptimal) {
        zobj = zsum;
        ok = (rp <= tol && rd <= tol && pconv && dconv) ? 1 : 0;
      } else {
        // ---- unbounded: a feasible point and an improving ray.  The point is the origin
        //      when it is feasible (b' >= 0 on inequality rows, b = 0 on '=' rows), else the
        //      final basis, which must then pass the primal test.  The ray of the entering
        //      variable e, d = B^-1 a'_e by refinement, must keep every basic variable from
        //      decreasing and gain c_e - c_B^T d > 0. ----
        const int e = out.unb[lp];
        int je = -1;
        for (int j = 1 + tid; j < C; j += NT)
          if (nbv[j] == e) je = j;
        je = (int)bmax<NT>((double)je, red);
        double onot = 0;
        for (int p = tid; p < m; p += NT)
          if (is_eq(p) ? bp[p] != 0.0 : bp[p] < 0.0) onot = 1;
        const bool origin_ok = bmax<NT>(onot, red) == 0.0;
        bool good = e >= 0 && e < n + m && rowof[e] < 0 && je > 0 && (origin_ok || (rp <= tol && pconv));
        if (good) {
          for (int i = tid; i < m; i += NT) dv[i] = (double)T[(int64_t)(2 + i) * C + je];
          for (int p = tid; p < m; p += NT)
            ae[p] = e < n ? sg(p) * A[(int64_t)p * n + e] : (e - n == p ? 1.0 : 0.0);
          __syncthreads();
          double dmax = 0, rres2 = 0;
          refine(dv, true, e, 1.0, rres2, dmax);
          double viol = 0;
          for (int i = tid; i < m; i += NT) {
            const int bv = bs[i];
            viol = fmax(viol, (bv >= n && is_eq(bv - n)) ? fabs(dv[i]) : dv[i]);
          }
          viol = bmax<NT>(viol, red);
          double part = 0;
          for (int a = tid; a < k; a += NT) part = fma(c[S[a]], dv[Srow[a]], part);
          const double gain = (e < n ? c[e] : 0.0) - bsum<NT>(part, red);
          good = viol <= tol * (1.0 + dmax) && rres2 <= tol * (1.0 + am * dmax) &&
                 gain > tol * (1.0 + cm * (1.0 + dmax));
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
    for (int j = tid; j < n; j += NT) xo[j] = 0.0;
    __syncthreads();
    for (int a = tid; a < kx; a += NT) xo[S[a]] = x[Srow[a]];
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
  // The fallback is warm-started from the FP32 basis when that basis ended phase II as
  // optimal, phase I with x0 basic, or phase II as unbounded with a basic solution that
  // refined (only the ray failed).  A final basis whose basic solution does not refine
  // (numerically singular in FP64) is replaced by the basis before the FP32 solve's last
  // pivot, when that pivot is known: the usual cause is that pivot itself, taken on an
  // element that is tiny in exact arithmetic.  Its leaving variable is the final nonbasic
  // variable of the pivot column, and goes back to the pivot row.  Other failed results
  // restart cold (crash list cleared).
  if (fp32_source && ok == 0 && fail_list) {
    int32_t* bw = out.basis + D.basis;
    const int lpv = (undo_last && out.lastpiv) ? out.lastpiv[lp] : -1;
    const int pr = lpv & 0xFFFF, pc = lpv >> 16;
    const int lvar = (lpv >= 0 && pr >= 2 && pr < R && pc >= 1 && pc < C) ? nbv[pc] : -1;
    if (lvar >= 0 && lvar < n + m) {
      if (tid == 0) bw[pr - 2] = lvar;
    } else if (!keep_basis) {
      for (int i = tid; i < m; i += NT) bw[i] = -1;
    }
  }
  return true;
}

// Persistent over an atomic queue (one LP per take, `queue` zeroed before the launch): a
// workgroup that draws slow LPs takes fewer, so the launch ends with the LPs, not with the
// unluckiest fixed share.  Dynamic LDS: [vectors (if lds_scratch)][stage_cap bytes].  With
// a defer_list, LPs whose staged blocks exceed stage_cap go to it (the next tier takes
// them); NT < 256 always defers them.
// NT = 128: at most 168 VGPRs, three waves per SIMD (five LPs per CU, as its LDS allows)
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 128 ? 3 : 1))) void k_verify(const LpDesc* __restrict__ desc,
                                               const double* __restrict__ pool,
                                               const int8_t* __restrict__ spool,
                                               const int32_t* __restrict__ idx, int32_t count,
                                               const int32_t* count_dev, Out out, Check chk,
                                               double tol, int fp32_source, char* ws,
                                               int64_t ws_stride, int lds_scratch,
                                               int64_t vec_bytes, int64_t stage_cap,
                                               int32_t* queue, int32_t* fail_list,
                                               int32_t* fail_count, int32_t* defer_list,
                                               int32_t* defer_count) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_q;
  __shared__ double red[2 * (NT / 64)];
  __shared__ int red_i[NT / 64];
  char* wsp = lds_scratch ? (char*)smem : ws + (int64_t)blockIdx.x * ws_stride;
  char* stg = (char*)smem + (lds_scratch ? vec_bytes : 0);
  const bool defer = defer_list != nullptr;
  if (count_dev) count = *count_dev;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
    __syncthreads();
    const int q = s_q;
    if (q >= count) break;
    const int lp = idx ? idx[q] : q;
    const LpDesc D = desc[lp];
    // the dump's precision: FP32 tableau, or FP64 for FP64 solves and FP64 fallbacks
    const bool done =
        fp32_source
            ? verify_lp<float, NT>(lp, D, pool, spool, out, chk, tol, 1, wsp, stg, stage_cap,
                                   defer, red, red_i, fail_list, fail_count)
            : verify_lp<double, NT>(lp, D, pool, spool, out, chk, tol, 0, wsp, stg, stage_cap,
                                    defer, red, red_i, fail_list, fail_count);
    if (!done && threadIdx.x == 0 && defer_list) defer_list[atomicAdd(defer_count, 1)] = lp;
  }
}

namespace {
// Tiers of the check, run in sequence on a bin; each takes the LPs the previous one
// deferred.  0: 128 threads and 31 KB per LP (5 LPs per CU: at 32 KB the fifth one is not
// resident although the occupancy query counts it), for LPs with few basic
// structurals; 1: 128 threads, 78 KB (2 per CU; 256 threads would need 256 VGPRs and leave
// one); 2: 256 threads, 150 KB (1 per CU), which
// defers nothing (global-memory sweeps for what does not fit, vectors in HBM when even
// they exceed 48 KB).
constexpr int kTierNT[3] = {128, 128, 256};
constexpr size_t kTierLds[3] = {31 * 1024, 78 * 1024, 150 * 1024};
constexpr size_t kVecLdsMax = 48 * 1024;

size_t vec_bytes(int mcap, int ncap, int nt) { return VLayout(mcap, ncap, nt).total; }
bool vec_in_lds(int mcap, int ncap, int nt) { return vec_bytes(mcap, ncap, nt) <= kVecLdsMax; }

// (stage_cap, total dynamic LDS) of a launch
void lds_plan(int tier, int mcap, int ncap, int64_t& stage_cap, size_t& lds) {
  const int nt = kTierNT[tier];
  const size_t vb = vec_bytes(mcap, ncap, nt);
  if (!vec_in_lds(mcap, ncap, nt)) { stage_cap = 0; lds = 0; return; }
  const size_t budget = kTierLds[tier];
  stage_cap = budget > vb ? (int64_t)(budget - vb) : 0;
  // never more than the largest LP of the bin can use (FP64 tableau, k = min(m, n))
  const int64_t most = (int64_t)stage_bytes(mcap, mcap < ncap ? mcap : ncap, 8);
  if (stage_cap > most) stage_cap = most;
  lds = vb + (size_t)stage_cap;
}
const void* kernel_of(int tier) {
  return kTierNT[tier] == 128 ? reinterpret_cast<const void*>(&k_verify<128>)
                              : reinterpret_cast<const void*>(&k_verify<256>);
}
}  // namespace

bool verify_tier_used(int tier, int mcap, int ncap, int tsz) {
  int64_t sc;
  size_t lds;
  lds_plan(tier, mcap, ncap, sc, lds);
  const int kmax = mcap < ncap ? mcap : ncap;
  if (tier == 0)  // worth a first pass when LPs with a dozen basic structurals fit
    return vec_in_lds(mcap, ncap, kTierNT[0]) && sc >= (int64_t)stage_bytes(mcap, 12, tsz);
  if (tier == 1)  // worth it when a quarter of the largest blocks fit
    return vec_in_lds(mcap, ncap, kTierNT[1]) && sc >= (int64_t)stage_bytes(mcap, kmax / 4, tsz);
  return true;
}

bool verify_tier_covers(int tier, int mcap, int ncap, int tsz) {
  int64_t sc;
  size_t lds;
  lds_plan(tier, mcap, ncap, sc, lds);
  return vec_in_lds(mcap, ncap, kTierNT[tier]) &&
         sc >= (int64_t)stage_bytes(mcap, mcap < ncap ? mcap : ncap, tsz);
}

int verify_max_blocks_per_cu(int tier, int mcap, int ncap) {
  int nb = 0;
  int64_t sc;
  size_t lds;
  lds_plan(tier, mcap, ncap, sc, lds);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel_of(tier), kTierNT[tier], lds) !=
      hipSuccess)
    return 0;
  return nb;
}

// HBM vectors per workgroup of the last tier when they exceed the LDS share (0: LDS)
size_t verify_ws_stride(int mcap, int ncap) {
  return vec_in_lds(mcap, ncap, 256) ? 0 : (vec_bytes(mcap, ncap, 256) + 255) & ~(size_t)255;
}

hipError_t launch_verify(const VerifyLaunch& V, hipStream_t st) {
  if (V.tier < 0 || V.tier > 2 || !V.queue) return hipErrorInvalidValue;
  const int nt = kTierNT[V.tier];
  const bool defers = V.tier < 2 && V.defer_list;
  if (V.tier < 2 && !vec_in_lds(V.mcap, V.ncap, nt)) return hipErrorInvalidValue;
  // a tier that may defer must be given a list, unless every LP of the bin fits it
  if (V.tier < 2 && (!V.defer_list || !V.defer_count) &&
      !verify_tier_covers(V.tier, V.mcap, V.ncap, V.fp32_source ? 4 : 8))
    return hipErrorInvalidValue;
  int64_t stage_cap;
  size_t lds;
  lds_plan(V.tier, V.mcap, V.ncap, stage_cap, lds);
  const int lds_scratch = vec_in_lds(V.mcap, V.ncap, nt) ? 1 : 0;
  const int64_t vb = (int64_t)vec_bytes(V.mcap, V.ncap, nt);
  char* ws = (char*)V.ws;
  const int64_t stride = V.tier < 2 ? 0 : (int64_t)verify_ws_stride(V.mcap, V.ncap);
  int32_t* dl = defers ? V.defer_list : nullptr;
  int32_t* dc = defers ? V.defer_count : nullptr;
  void* args[] = {(void*)&V.desc, (void*)&V.pool, (void*)&V.spool, (void*)&V.idx,
                  (void*)&V.count, (void*)&V.count_dev, (void*)&V.out, (void*)&V.chk,
                  (void*)&V.tol, (void*)&V.fp32_source, (void*)&ws, (void*)&stride,
                  (void*)&lds_scratch, (void*)&vb, (void*)&stage_cap, (void*)&V.queue,
                  (void*)&V.fail_list, (void*)&V.fail_count, (void*)&dl, (void*)&dc};
  return hipLaunchKernel(kernel_of(V.tier), dim3(V.grid), dim3(nt), args, lds, st);
}

}  // namespace blpg
