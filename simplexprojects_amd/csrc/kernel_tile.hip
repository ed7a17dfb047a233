// kernel_tile.hip -- register-resident simplex kernel for gfx950.
//
// One workgroup (NW wavefronts) per LP, persistent over a work queue.  The condensed
// tableau never leaves the vector register file: thread t = gc + TC*gr owns rows
// gr + TR*k (k < RPT) and columns gc + TC*l (l < CPT), i.e. an RPT x CPT block held in
// `v[RPT][CPT]`.  A CU's 512 KB of VGPRs is 3.2x its LDS, so tableaux up to ~207x205 FP64
// stay on-chip, and the rank-1 update is RPT*CPT register FMAs per thread with no memory
// traffic at all.  Per pivot LDS carries only
//   colbuf  the pivot column (written by the TR owners of column s)
//   rhsbuf  the right-hand side column, kept current by its owners after every update
//   rowbuf  the scaled pivot row (written by the TC owners of row r)
// Division of labour per pivot: every (value, index) reduction -- pricing over row 0
// (whose owners are lanes 0..TC-1 of wavefront 0), ratio over the published column, and
// the phase-I / crash / dual selections -- runs on wavefront 0 alone and its result is
// broadcast through LDS; the other wavefronts wait at the barrier without issuing
// instructions, so the SIMDs they share with other workgroups stay free.  Only the
// rank-1 update uses every wavefront.  Indices of the pivot row / column are
// wave-uniform, so their extraction from the register block is a scalar branch over the
// unrolled k / l slots, not a dynamic register index.
//
// Arithmetic and rule are those of kernel_stream.hip and the oracle, operation for
// operation: results are bit-identical across kernels.  Replaces the reference's
// mykernel (simplex.cu:35-385), whose tableau lives in global memory.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "simplex_device.hpp"

namespace blpg {

// lane l's value (l wave-uniform)
__device__ __forceinline__ float lane_read(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double lane_read(double x, int l) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// wavefront 0's decision, broadcast to the workgroup
template <typename T>
struct Decision {
  T v;
  int k, p;
};

// BC: crash in blocks (crash_prepare).  Instantiated only for batches with warm lists, and
// only where its registers do not spill the tile; cold solves run the BC = false kernel.
template <typename T, int NW, int TC, int RPT, int CPT, bool BC>
struct TileTab {
  static constexpr int NT = NW * 64;
  static constexpr int TR = NT / TC;
  static constexpr int RCAP = TR * RPT;
  static constexpr int CCAP = TC * CPT;
  static_assert(TC <= 32 && 64 % TC == 0, "rows 0 and 1 must sit in wavefront 0");
  static_assert(CPT <= 32, "dead-column mask is 32 bits");

  T v[RPT][CPT];
  uint32_t dead;  // bit l: column gc + TC*l cannot enter
  int gr, gc, tid, lane, wv;
  int R, C, m, n;
  int col_ready;  // column whose entries are in colbuf (-1: none)
  int phase;      // broadcast slot
  T* colbuf;
  T* rhsbuf;
  T* rowbuf;
  T* scal;
  int* bs;
  int* nb;
  uint8_t* redund;
  uint8_t* target;  // target[var] = 1: var belongs to the warm-start basis
  int* wl;          // the warm-start list (m entries), read by the crash in order
  int* colof;       // column of each nonbasic variable, -1 while it is basic
  double* s_am;     // per-wave max |A_ij| (NW entries)
  Decision<T>* slots;  // 2 broadcast slots

  // Block crash: KB targets per block.  Their columns and pivot rows take at most 16 KB
  // of LDS, and wavefront 0 / each column thread holds KB values per row / column in
  // registers during the block's preparation (more than 8 spills the tile).
  static constexpr int KB0 = 16384 / ((RCAP + CCAP) * (int)sizeof(T));
  static constexpr int KB = !BC ? 1 : KB0 < 4 ? 4 : (KB0 > 8 ? 8 : KB0);
  static constexpr bool kBlockCrash = BC;
  static constexpr int QR = (RCAP + 63) / 64;  // rows per lane of wavefront 0
  T* cbM;    // KB x RCAP: column of target u, updated until its own step, then frozen
  T* cbP;    // KB x CCAP: row of pivot u, updated until its own step, then scaled
  T* cbRp;   // KB: 1 / pivot element of step u
  int* cbR;  // KB + 3: pivot row of step u (-1: none); [KB] steps, [KB + 1] next ci,
             // [KB + 2] pivots made
  int* cbS;  // KB: column of target u
  int* cbW;  // KB: the block's pivots in order, packed: row | column << 12 | slot << 24 |
             // (leaving variable is an equality slack) << 30
  int blk;   // >= 0: the next pivot() is the crash block's step in this slot

  // Called once per state-machine step: the per-thread indices become opaque values, so
  // the compiler recomputes index arithmetic where it is used instead of hoisting every
  // stage's invariants out of the loop (which holds registers for the whole kernel and
  // costs a wave per SIMD of occupancy).
  __device__ __forceinline__ void loop_fence() { asm volatile("" : "+v"(gr), "+v"(gc), "+v"(tid), "+v"(lane)); }

  __device__ __forceinline__ int row_of(int k) const { return gr + TR * k; }
  __device__ __forceinline__ int col_of(int l) const { return gc + TC * l; }

  // wavefront 0's (wave-uniform) reduction result to every wavefront
  __device__ __forceinline__ void bcast(T& bv, int& bk, int& bp) {
    Decision<T>* sl = slots + phase;
    if (tid == 0) { sl->v = bv; sl->k = bk; sl->p = bp; }
    __syncthreads();
    bv = uniform(sl->v);
    bk = uniform(sl->k);
    bp = uniform(sl->p);
    phase ^= 1;
  }

  __device__ __forceinline__ void publish_rhs() {
    if (gc == 0) {
#pragma unroll
      for (int k = 0; k < RPT; ++k) rhsbuf[row_of(k)] = v[k][0];
    }
  }

  // Loads the condensed tableau (row 0: -c, row 1: 0, rows 2..: A and b with '>=' rows
  // negated) and max |A_ij|.  Every thread loads each of its entries from a valid address
  // (a dummy one where the entry is structurally 0), so all loads are unconditional and
  // go out together instead of one predicated load at a time.
  __device__ __forceinline__ void init(const double* A, const double* b, const double* c, const int8_t* sense) {
    double am = 0.0;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const int i = row_of(k);
      const bool crow = i >= 2 && i < R;
      const int r = crow ? i - 2 : 0;
      const bool neg = crow && sense && sense[r] == 2;
      const double* Ar = A + (int64_t)r * n;
#pragma unroll
      for (int l = 0; l < CPT; ++l) {
        const int j = col_of(l);
        const bool acol = j >= 1 && j <= n;
        const bool isA = crow && acol, isB = crow && j == 0, isC = i == 0 && acol;
        const double* src = isA ? Ar + (j - 1) : isB ? b + r : isC ? c + (j - 1) : b;
        const double x = *src;
        const bool ng = isC || (neg && (isA || isB));
        v[k][l] = (isA || isB || isC) ? T(ng ? -x : x) : T(0);
        if (isA) am = fmax(am, fabs(x));
      }
    }
    dead = 0;
#pragma unroll
    for (int l = 0; l < CPT; ++l) {
      const int j = col_of(l);
      if (j == 0 || j >= C - 1) dead |= 1u << l;
    }
    // max |A_ij| in FP64 (the check scales its tolerances by it)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) am = fmax(am, __shfl_xor(am, off));
    if (lane == 0) s_am[tid >> 6] = am;
    for (int i = tid; i < RCAP; i += NT) { bs[i] = (i >= 2 && i < R) ? n + i - 2 : -1; redund[i] = 0; }
    for (int j = tid; j < CCAP; j += NT) nb[j] = (j == 0 || j >= C) ? -1 : (j <= n ? j - 1 : n + m);
    for (int x = tid; x < RCAP + CCAP; x += NT)
      colof[x] = x < n ? x + 1 : (x == n + m ? C - 1 : -1);
    publish_rhs();
    col_ready = -1;
    __syncthreads();
  }

  // marks the warm-start targets and keeps the list in LDS; false when the list holds no
  // valid variable id
  __device__ __forceinline__ bool set_target(const int32_t* warm) {
    for (int x = tid; x < RCAP + CCAP; x += NT) target[x] = 0;
    __syncthreads();
    int any = 0;
    if (warm)
      for (int i = tid; i < m; i += NT) {
        const int x = warm[i];
        wl[i] = x;
        if (x >= 0 && x < n + m) { target[x] = 1; any = 1; }
      }
    return __syncthreads_or(any) != 0;
  }

  // entry ci of the warm-start list (from LDS; uniform)
  __device__ __forceinline__ int warm_at(const int32_t*, int ci) const { return uniform(wl[ci]); }

  // basic variable of row i when no pivot has happened since the last barrier
  __device__ __forceinline__ int bs_now(int i) const { return uniform(bs[i]); }

  // the TR owners of column s store it in dst (RCAP entries)
  __device__ __forceinline__ void publish_col(int s, T* dst) {
    const int ls = s / TC, gcs = s - ls * TC;
#pragma unroll
    for (int l = 0; l < CPT; ++l) {
      if (l == ls) {
        if (gc == gcs) {
#pragma unroll
          for (int k = 0; k < RPT; ++k) dst[row_of(k)] = v[k][l];
        }
      }
    }
  }
  __device__ __forceinline__ void publish_col(int s) { publish_col(s, colbuf); }

  // the TC owners of row r store it (unscaled) in dst (CCAP entries)
  __device__ __forceinline__ void publish_row(int r, T* dst) {
    const int kr = r / TR, grr = r - kr * TR;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      if (k == kr && gr == grr) {
#pragma unroll
        for (int l = 0; l < CPT; ++l) dst[col_of(l)] = v[k][l];
      }
    }
  }

  // row of largest |T[i][s]| among constraint rows whose basic variable is not a target
  __device__ __forceinline__ int absmax_row(int s, T eps_p) {
    __syncthreads();
    publish_col(s);
    __syncthreads();
    col_ready = s;
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    if (wv == 0) {
      ArgMin<T> am;
      for (int i = 2 + lane; i < R; i += 64) {
        const int x = bs[i];
        const bool ok = !redund[i] && x >= 0 && !target[x < 0 ? 0 : x];
        am.add(ok, -tabs(colbuf[i]), i, i);
      }
      wave_finish(am, bv, bk, bp);
    }
    bcast(bv, bk, bp);
    if (bp < 0 || !(-bv > eps_p)) return -1;
    return bp;
  }

  __device__ __forceinline__ int bs_of(int i) {
    __syncthreads();
    return uniform(bs[i]);
  }

  // Dantzig pricing on objective row `orow` (lanes orow*TC .. orow*TC+TC-1 of wavefront 0)
  __device__ __forceinline__ int pricing(int orow, bool bland, T eps_d) {
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    if (wv == 0) {
      ArgMin<T> am;
      if (gr == orow) {
        if (bland) {
#pragma unroll
          for (int l = 0; l < CPT; ++l) {
            const int j = col_of(l);
            am.add(!((dead >> l) & 1u) && v[0][l] < -eps_d, T(nb[j]), j, j);
          }
        } else {
#pragma unroll
          for (int l = 0; l < CPT; ++l) {
            const int j = col_of(l);
            am.add(!((dead >> l) & 1u) && v[0][l] < -eps_d, v[0][l], j, j);
          }
        }
      }
      wave_finish(am, bv, bk, bp);
    }
    bcast(bv, bk, bp);
    return bp;
  }

  __device__ __forceinline__ int ratio(int s, bool bland, T eps_p, T& theta) {
    publish_col(s);
    __syncthreads();
    col_ready = s;
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    if (wv == 0) {
      ArgMin<T> am;
      for (int i = 2 + lane; i < R; i += 64) {
        const T a = colbuf[i];
        const T beta = rhsbuf[i];
        const T th = (beta > T(0) ? beta : T(0)) / (a > eps_p ? a : T(1));
        const T key = th > T(0) ? th : -a;  // degenerate rows: largest pivot first
        am.add(a > eps_p, key, bland ? bs[i] : i, i);
      }
      wave_finish(am, bv, bk, bp);
    }
    bcast(bv, bk, bp);
    theta = bv > T(0) ? bv : T(0);
    return bp;
  }

  // One pivot.  With blk >= 0 it is the crash block's step in slot blk, whose column,
  // scaled row and basis change crash_prepare made: no barrier, every thread reads only
  // LDS that no one writes until the next barrier.  colbuf / rowbuf are slot 0 of the
  // block arrays, so both cases read the column and row at slot * RCAP / CCAP.
  __device__ __forceinline__ void pivot(int r, int s) {
    const int kr = r / TR, grr = r - kr * TR;
    const int ls = s / TC, gcs = s - ls * TC;
    int slot = 0;
    bool prepared = false;
    if constexpr (BC) {
      prepared = blk >= 0;
      if (prepared) slot = blk;
    }
    if (!prepared) {
      if (col_ready != s) {
        __syncthreads();
        publish_col(s);
        __syncthreads();
      }
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        if (k == kr) {
          if (gr == grr) {
            const T rp = T(1) / colbuf[r];
#pragma unroll
            for (int l = 0; l < CPT; ++l) {
              const int j = col_of(l);
              rowbuf[j] = (j == s) ? rp : v[k][l] * rp;
            }
          }
        }
      }
      if (tid == 0) {
        const int t = bs[r], e = nb[s];
        bs[r] = e;
        nb[s] = t;
        colof[e] = -1;
        colof[t] = s;
      }
      __syncthreads();
    }
    T col[RPT], prow[CPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) col[k] = -colbuf[slot * RCAP + row_of(k)];
#pragma unroll
    for (int l = 0; l < CPT; ++l) prow[l] = rowbuf[slot * CCAP + col_of(l)];
#pragma unroll
    for (int l = 0; l < CPT; ++l) {
      if (l == ls) {
        const bool own = gc == gcs;
#pragma unroll
        for (int k = 0; k < RPT; ++k) v[k][l] = own ? T(0) : v[k][l];
      }
    }
#pragma unroll
    for (int k = 0; k < RPT; ++k)
#pragma unroll
      for (int l = 0; l < CPT; ++l) v[k][l] = fma_t(col[k], prow[l], v[k][l]);
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      if (k == kr) {
        const bool own = gr == grr;
#pragma unroll
        for (int l = 0; l < CPT; ++l) v[k][l] = own ? prow[l] : v[k][l];
      }
    }
    publish_rhs();
    col_ready = -1;
    blk = -1;
  }

  // column of largest |T[r][j]| among live columns; -1 if all <= eps_p
  __device__ __forceinline__ int absmax_col(int r, T eps_p) {
    __syncthreads();
    publish_row(r, rowbuf);
    __syncthreads();
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    if (wv == 0) {
      ArgMin<T> am;
      if (gr == 0) {
#pragma unroll
        for (int l = 0; l < CPT; ++l) {
          const int j = col_of(l);
          am.add(!((dead >> l) & 1u), -tabs(rowbuf[j]), j, j);
        }
      }
      wave_finish(am, bv, bk, bp);
    }
    bcast(bv, bk, bp);
    if (bp < 0 || !(-bv > eps_p)) return -1;
    return bp;
  }

  // dual simplex leaving row: most negative rhs below -tol (Bland: smallest basic id)
  __device__ __forceinline__ int dual_row(bool bland, T tol) {
    __syncthreads();
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    if (wv == 0) {
      ArgMin<T> am;
      for (int i = 2 + lane; i < R; i += 64) {
        const T beta = rhsbuf[i];
        am.add(!redund[i] && beta < -tol, bland ? T(bs[i]) : beta, i, i);
      }
      wave_finish(am, bv, bk, bp);
    }
    bcast(bv, bk, bp);
    return bp;
  }

  // dual simplex entering column for leaving row r (row r published through rowbuf; the
  // objective row is read from the registers of its owners)
  __device__ __forceinline__ int dual_ratio(int r, bool bland, T eps_p, T& theta) {
    publish_row(r, rowbuf);
    __syncthreads();
    col_ready = -1;
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    if (wv == 0) {
      ArgMin<T> am;
      if (gr == 0) {
#pragma unroll
        for (int l = 0; l < CPT; ++l) {
          const int j = col_of(l);
          const T a = rowbuf[j];
          const T d = v[0][l];
          const bool ok = !((dead >> l) & 1u) && a < -eps_p;
          const T th = (d > T(0) ? d : T(0)) / (ok ? -a : T(1));
          const T key = th > T(0) ? th : a;
          am.add(ok, key, bland ? nb[j] : j, j);
        }
      }
      wave_finish(am, bv, bk, bp);
    }
    bcast(bv, bk, bp);
    theta = bv > T(0) ? bv : T(0);
    return bp;
  }

  // A dead column never enters again, and its entries influence no other entry (the
  // update of (i, j) reads only column s and row r), so killing it only sets its bit.
  __device__ __forceinline__ void kill_col(int j) {
    const int lj = j / TC, gcj = j - lj * TC;
    if (gc == gcj) dead |= 1u << lj;
  }

  __device__ __forceinline__ T get(int i, int j) {
    const int ki = i / TR, gri = i - ki * TR, lj = j / TC, gcj = j - lj * TC;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RPT; ++k)
#pragma unroll
      for (int l = 0; l < CPT; ++l)
        if (k == ki && l == lj && gr == gri && gc == gcj) *scal = v[k][l];
    __syncthreads();
    const T r = uniform(*scal);
    __syncthreads();
    return r;
  }

  __device__ __forceinline__ void set(int i, int j, T val) {
    const int ki = i / TR, gri = i - ki * TR, lj = j / TC, gcj = j - lj * TC;
#pragma unroll
    for (int k = 0; k < RPT; ++k)
#pragma unroll
      for (int l = 0; l < CPT; ++l)
        if (k == ki && l == lj) v[k][l] = (gr == gri && gc == gcj) ? val : v[k][l];
    if (j == 0 && gr == gri && gc == 0) rhsbuf[i] = val;
  }

  __device__ __forceinline__ void mark_redund(int r) {
    __syncthreads();
    if (tid == 0) redund[r] = 1;
    set(r, 0, T(0));
    __syncthreads();
  }

  __device__ __forceinline__ int min_rhs_row() {
    __syncthreads();
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    if (wv == 0) {
      ArgMin<T> am;
      for (int i = 2 + lane; i < R; i += 64) am.add(!redund[i], rhsbuf[i], i, i);
      wave_finish(am, bv, bk, bp);
    }
    bcast(bv, bk, bp);
    return bp;
  }

  __device__ __forceinline__ void setup_x0() {
    __syncthreads();
    const int x0 = C - 1, lx = x0 / TC, gcx = x0 - lx * TC;
#pragma unroll
    for (int l = 0; l < CPT; ++l) {
      if (l == lx && gc == gcx) {
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
          const int i = row_of(k);
          v[k][l] = (i >= 2 && i < R) ? (redund[i] ? T(0) : T(-1)) : (i == 1 ? T(1) : T(0));
        }
        dead &= ~(1u << l);
      }
    }
    if (gr == 1) {
#pragma unroll
      for (int l = 0; l < CPT; ++l)
        if (col_of(l) != x0) v[0][l] = T(0);
    }
    __syncthreads();
  }

  __device__ __forceinline__ int find_basic(int var) {
    __syncthreads();
    T bv = inf_of<T>();
    int bk = INT_MAX, bp = -1;
    if (wv == 0) {
      ArgMin<T> am;
      for (int i = 2 + lane; i < R; i += 64) am.add(bs[i] == var, T(0), i, i);
      wave_finish(am, bv, bk, bp);
    }
    bcast(bv, bk, bp);
    return bp;
  }

  // column of nonbasic variable var (-1 if basic), from the map the pivots keep.  No barrier:
  // a pivot writes the map before its own barrier, and only pivots write it.
  __device__ __forceinline__ int find_nonbasic(int var) { return uniform(colof[var]); }

  // pivot j of the prepared block, packed (wave-uniform)
  __device__ __forceinline__ int blk_word(int j) const { return uniform(cbW[j]); }

  // Prepares the next up-to-KB crash targets as one block.  Same targets, rows and
  // arithmetic as one absmax_row + pivot per target, so the tableau ends bit-identical:
  //   targets  the warm-list structurals from ci on that are nonbasic, stopping before a
  //            repeated column (its second turn must see the first one's pivot);
  //   A        wavefront 0 runs the pivots on the targets' columns alone (cbM): step t
  //            picks its row from column t as absmax_row does (rows pivoted by earlier
  //            steps no longer eligible) and updates the later columns; column t keeps
  //            the values step t multiplies by.  The steps' basis changes touch distinct
  //            rows, columns and variables, so lane u then makes step u's;
  //   B        one thread per column j runs the pivot rows alone (cbP): row r_t, updated
  //            by steps < t, becomes step t's scaled pivot row.
  // The caller then makes the block's pivots (cbW) through pivot() with `blk` set to the
  // step's slot, which reads the column and row from cbM / cbP and needs no barrier.  Six
  // barriers per block instead of about five per target.  Returns the number of pivots
  // (steps without an eligible row make none), or -1 when no target is left (ci = m).
  __device__ int crash_prepare(int& ci, const int8_t* sense, T eps_p) {
    __syncthreads();
    // targets: wavefront 0 reads the list 64 entries at a time and takes the nonbasic
    // structurals in order; `sel` on lane u holds target u's column
    if (wv == 0) {
      int kc = 0, c = ci, sel = -1;
      bool stop = false;
      while (!stop && c < m && kc < KB) {
        const int idx = c + lane;
        const int w = idx < m ? wl[idx] : -1;
        const int s = (w >= 0 && w < n) ? colof[w] : -1;
        unsigned long long bal = __ballot(s >= 0);
        int next = c + 64 < m ? c + 64 : m;
        while (bal != 0ull) {
          const int l = __ffsll((long long)bal) - 1;
          const int sl = __builtin_amdgcn_readlane(s, l);
          if (__ballot(lane < kc && sel == sl) != 0ull) {  // repeated: stop before it
            stop = true;
            next = c + l;
            break;
          }
          if (lane == kc) sel = sl;
          ++kc;
          bal &= bal - 1ull;
          if (kc == KB) {
            next = c + l + 1;
            break;
          }
        }
        c = next;
      }
      if (lane < kc) cbS[lane] = sel;
      if (lane == 0) {
        cbR[KB] = kc;
        cbR[KB + 1] = c;
      }
    }
    __syncthreads();
    const int kc = uniform(cbR[KB]);
    ci = uniform(cbR[KB + 1]);
    if (kc == 0) return -1;
    for (int u = 0; u < kc; ++u) publish_col(uniform(cbS[u]), cbM + u * RCAP);
    __syncthreads();

    // A: wavefront 0, the block's columns in registers (lane rows lane + 64 q)
    if (wv == 0) {
      T mv[QR][KB];
      bool el[QR];
#pragma unroll
      for (int q = 0; q < QR; ++q) {
        const int i = lane + 64 * q;
        const int x = (i >= 2 && i < R) ? bs[i] : -1;
        el[q] = x >= 0 && !redund[i] && !target[x];
#pragma unroll
        for (int u = 0; u < KB; ++u) mv[q][u] = (u < kc && i < RCAP) ? cbM[u * RCAP + i] : T(0);
      }
#pragma unroll
      for (int t = 0; t < KB; ++t) {
        if (t < kc) {
          T bv = inf_of<T>();
          int bk = INT_MAX, bp = -1;
          {
            ArgMin<T> am;
#pragma unroll
            for (int q = 0; q < QR; ++q) am.add(el[q], -tabs(mv[q][t]), lane + 64 * q, lane + 64 * q);
            wave_finish(am, bv, bk, bp);
          }
          const int r = uniform(bp);
          int rr = -1;
          if (r >= 0 && -bv > eps_p) {
            rr = r;
            const int rl = r & 63, rq = r >> 6;
            T pv = mv[0][t];
#pragma unroll
            for (int q = 1; q < QR; ++q) pv = (rq == q) ? mv[q][t] : pv;
            const T rp = T(1) / lane_read(pv, rl);
#pragma unroll
            for (int u = t + 1; u < KB; ++u) {
              if (u < kc) {
                T x = mv[0][u];
#pragma unroll
                for (int q = 1; q < QR; ++q) x = (rq == q) ? mv[q][u] : x;
                const T pr = lane_read(x, rl) * rp;
#pragma unroll
                for (int q = 0; q < QR; ++q)
                  mv[q][u] = (lane + 64 * q == r) ? pr : fma_t(-mv[q][t], pr, mv[q][u]);
              }
            }
#pragma unroll
            for (int q = 0; q < QR; ++q) el[q] = el[q] && (lane + 64 * q != r);
            if (lane == 0) cbRp[t] = rp;
          }
          if (lane == 0) cbR[t] = rr;
        }
      }
      // column t as step t multiplied by (no later step of the block changes it)
#pragma unroll
      for (int q = 0; q < QR; ++q) {
        const int i = lane + 64 * q;
        if (i < RCAP) {
#pragma unroll
          for (int u = 0; u < KB; ++u)
            if (u < kc) cbM[u * RCAP + i] = mv[q][u];
        }
      }
      // basis changes: each step has its own row, column, entering and leaving variable,
      // so lane u makes step u's
      const int r = lane < kc ? cbR[lane] : -1;
      const unsigned long long made = __ballot(r >= 0);
      if (lane == 0) cbR[KB + 2] = __popcll(made);
      if (r >= 0) {
        {
          const int s = cbS[lane];
          const int lv = bs[r], e = nb[s];
          const int kill = (lv >= n && lv < n + m && sense && sense[lv - n] == 1) ? 1 : 0;
          cbW[__popcll(made & ((1ull << lane) - 1ull))] = r | (s << 12) | (lane << 24) | (kill << 30);
          bs[r] = e;
          nb[s] = lv;
          colof[e] = -1;
          colof[lv] = s;
        }
      }
    }
    __syncthreads();
    for (int t = 0; t < kc; ++t) {
      const int r = uniform(cbR[t]);
      if (r >= 0) publish_row(r, cbP + t * CCAP);
    }
    __syncthreads();

    // B: thread per column, the block's pivot rows in registers
    for (int j = tid; j < C; j += NT) {
      T pc[KB];
#pragma unroll
      for (int u = 0; u < KB; ++u) pc[u] = u < kc ? cbP[u * CCAP + j] : T(0);
#pragma unroll
      for (int t = 0; t < KB; ++t) {
        if (t < kc) {
          const int r = cbR[t];
          if (r >= 0) {
            const int s = cbS[t];
            const T rp = cbRp[t];
            const T pr = (j == s) ? rp : pc[t] * rp;
            pc[t] = pr;
#pragma unroll
            for (int u = t + 1; u < KB; ++u) {
              if (u < kc) {
                const int ru = cbR[u];
                if (ru >= 0) pc[u] = fma_t(-cbM[t * RCAP + ru], pr, (j == s) ? T(0) : pc[u]);
              }
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < KB; ++u)
        if (u < kc) cbP[u * CCAP + j] = pc[u];
    }
    __syncthreads();
    return uniform(cbR[KB + 2]);
  }


  __device__ __forceinline__ int nb_of(int j) {
    __syncthreads();
    return uniform(nb[j]);
  }

  // column 0 is mirrored in rhsbuf (objective T[0][0], basic values T[i][0])
  __device__ __forceinline__ void write_outputs(int lp, const LpDesc& d, const Out& out, int status, int it,
                                int p1, int unb, T wfin, int last) {
    __syncthreads();
    if (tid == 0) {
      double am = s_am[0];
      for (int w = 1; w < NW; ++w) am = fmax(am, s_am[w]);
      out.amax[lp] = am;
      out.status[lp] = status;
      out.pivots[lp] = it;
      out.p1[lp] = p1;
      out.unb[lp] = unb;
      if (out.lastpiv) out.lastpiv[lp] = last;
      out.obj[lp] = (double)rhsbuf[0];
      out.wfin[lp] = (double)wfin;
    }
    double* x = out.x + d.x;
    int32_t* basis = out.basis + d.basis;
    for (int j = tid; j < n; j += NT) x[j] = 0.0;
    for (int i = tid; i < m; i += NT) basis[i] = bs[2 + i];
    __syncthreads();
    for (int i = 2 + tid; i < R; i += NT) {
      const int var = bs[i];
      if (var >= 0 && var < n) x[var] = (double)rhsbuf[i];
    }
    // final tableau and nonbasic ids for the FP64 check (B^-1 is in the slack columns)
    if (out.tab) {
      T* tp = (T*)(out.tab + 8 * d.t);
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const int i = row_of(k);
#pragma unroll
        for (int l = 0; l < CPT; ++l) {
          const int j = col_of(l);
          if (i < R && j < C) tp[(int64_t)i * C + j] = v[k][l];
        }
      }
      int32_t* np = (int32_t*)(out.tab + 8 * (d.t + (int64_t)R * C));
      for (int j = tid; j < C; j += NT) np[j] = nb[j];
    }
  }
};

// WPE: minimum waves per SIMD the register allocator must allow (1: no constraint)
template <typename T, int NW, int TC, int RPT, int CPT, int WPE, bool BC>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(WPE))) void k_tile(const LpDesc* __restrict__ desc,
                                                  const double* __restrict__ pool,
                                                  const int8_t* __restrict__ spool,
                                                  const int32_t* __restrict__ idx, int32_t count,
                                                  const int32_t* count_dev, int32_t* queue,
                                                  Out out, Opt<T> o,
                                                  const int32_t* __restrict__ warm_pool) {
  using Tab = TileTab<T, NW, TC, RPT, CPT, BC>;
  __shared__ T rhsbuf[Tab::RCAP], scal[2];
  __shared__ int bs[Tab::RCAP], nb[Tab::CCAP];
  __shared__ uint8_t redund[Tab::RCAP], target[Tab::RCAP + Tab::CCAP];
  __shared__ int wl[Tab::RCAP], colof[Tab::RCAP + Tab::CCAP];
  __shared__ Decision<T> slots[2];
  __shared__ double s_am[NW];
  __shared__ int s_q;
  __shared__ T cbM[Tab::KB * Tab::RCAP], cbP[Tab::KB * Tab::CCAP], cbRp[Tab::KB];
  __shared__ int cbR[Tab::KB + 3], cbS[Tab::KB], cbW[Tab::KB];
  Tab t;
  t.s_am = s_am;
  t.tid = threadIdx.x;
  t.lane = threadIdx.x & 63;
  t.wv = uniform((int)(threadIdx.x >> 6));
  t.gr = threadIdx.x / TC;
  t.gc = threadIdx.x - t.gr * TC;
  t.phase = 0;
  t.colbuf = cbM; t.rhsbuf = rhsbuf; t.rowbuf = cbP; t.scal = scal;  // slot 0
  t.bs = bs; t.nb = nb; t.redund = redund; t.target = target;
  t.wl = wl; t.colof = colof;
  t.slots = slots;
  t.cbM = cbM; t.cbP = cbP; t.cbRp = cbRp; t.cbR = cbR; t.cbS = cbS; t.cbW = cbW;
  t.blk = -1;
  if (count_dev) count = *count_dev;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
    __syncthreads();
    const int q = uniform(s_q);
    if (q >= count) break;
    const int lp = idx ? idx[q] : q;
    const LpDesc d = desc[lp];
    t.m = d.m; t.n = d.n; t.R = d.m + 2; t.C = d.n + 2;
    const int8_t* sense = d.s >= 0 ? spool + d.s : nullptr;
    t.init(pool + d.a, pool + d.b, pool + d.c, sense);
    const int32_t* warm = warm_pool ? warm_pool + d.basis : nullptr;
    if (!t.set_target(warm)) warm = nullptr;
    int status, it, p1, unb, last;
    T wfin;
    run_simplex<T>(t, d.m, d.n, sense, o, status, it, p1, unb, wfin, last, warm);
    t.write_outputs(lp, d, out, status, it, p1, unb, wfin, last);
  }
}

// ---------------------------------------------------------------------------------
// variant table: (waves per LP, TC, RPT, CPT); capacity (64*NW/TC*RPT) x (TC*CPT)
// (the autotuner measures every variant that holds a shape; fewer waves per LP mean fewer
// barriers per pivot, more waves mean fewer registers per thread).  w32 / w64: the
// waves-per-SIMD floors given to the register allocator for FP32 / FP64, set where a few
// bytes of spill buy one more resident workgroup per CU.  b32 / b64: block crash on (the
// variants whose register block leaves room for it without spilling).
//   id nw  tc rpt cpt w32 w64 b32 b64
#define BLPG_TILE_VARIANTS(X)              \
  X(0, 1, 8, 4, 5, 1, 1, 1, 1)             \
  X(1, 1, 8, 7, 7, 4, 1, 1, 1)             \
  X(2, 2, 16, 8, 7, 1, 1, 1, 1)            \
  X(3, 2, 16, 10, 6, 1, 1, 1, 1)           \
  X(4, 4, 16, 7, 7, 4, 1, 1, 1)            \
  X(5, 8, 16, 6, 9, 1, 1, 1, 0)            \
  X(6, 8, 16, 7, 13, 1, 1, 0, 0)           \
  X(7, 1, 8, 10, 11, 1, 1, 0, 0)           \
  X(8, 1, 16, 16, 7, 1, 1, 0, 0)           \
  X(9, 2, 16, 14, 7, 1, 1, 0, 0)           \
  X(10, 4, 32, 14, 4, 1, 1, 1, 1)          \
  X(11, 8, 16, 4, 7, 1, 1, 1, 1)

constexpr int kTileVariants = 12;
constexpr int kTileDefault = 7;  // static routing uses variants 0..6

int tile_variant_count(int) { return kTileVariants; }
int tile_default_variants() { return kTileDefault; }

TileVariant tile_variant(int, int v) {
  switch (v) {
#define X(id, nw, tc, rpt, cpt, w32, w64, b32, b64) \
  case id: return TileVariant{nw, (nw * 64) / tc, rpt, cpt};
    BLPG_TILE_VARIANTS(X)
#undef X
    default: return TileVariant{1, 8, 1, 1};
  }
}

// warm: the batch carries warm-start lists (block crash where the variant has it)
template <typename T>
static const void* tile_fn(int v, bool warm) {
  switch (v) {
#define X(id, nw, tc, rpt, cpt, w32, w64, b32, b64)                                           \
  case id:                                                                                     \
    if (warm && (sizeof(T) == 4 ? b32 : b64) != 0)                                             \
      return reinterpret_cast<const void*>(                                                    \
          &k_tile<T, nw, tc, rpt, cpt, sizeof(T) == 4 ? w32 : w64, true>);                     \
    return reinterpret_cast<const void*>(&k_tile<T, nw, tc, rpt, cpt, sizeof(T) == 4 ? w32 : w64, false>);
    BLPG_TILE_VARIANTS(X)
#undef X
    default: return nullptr;
  }
}

// the smaller of the cold and warm instantiations' residency
int tile_max_blocks_per_cu(int precision, int v) {
  int best = -1;
  for (int warm = 0; warm < 2; ++warm) {
    const void* fn = precision == 0 ? tile_fn<double>(v, warm) : tile_fn<float>(v, warm);
    if (!fn) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, tile_variant(precision, v).nw * 64,
                                                     0) != hipSuccess)
      return 0;
    best = best < 0 || nb < best ? nb : best;
  }
  return best;
}

hipError_t launch_tile(const TileLaunch& L, hipStream_t st) {
  const bool warm = L.warm != nullptr;
  const void* fn = L.precision == 0 ? tile_fn<double>(L.variant, warm) : tile_fn<float>(L.variant, warm);
  if (!fn) return hipErrorInvalidValue;
  const int nt = tile_variant(L.precision, L.variant).nw * 64;
  if (L.precision == 0) {
    Opt<double> o{L.eps_d, L.eps_p, L.tol, L.max_iter, L.bland_after};
    void* args[] = {(void*)&L.desc, (void*)&L.pool, (void*)&L.spool, (void*)&L.idx,
                    (void*)&L.count, (void*)&L.count_dev, (void*)&L.queue, (void*)&L.out,
                    (void*)&o, (void*)&L.warm};
    return hipLaunchKernel(fn, dim3(L.grid), dim3(nt), args, 0, st);
  }
  Opt<float> o{(float)L.eps_d, (float)L.eps_p, (float)L.tol, L.max_iter, L.bland_after};
  void* args[] = {(void*)&L.desc, (void*)&L.pool, (void*)&L.spool, (void*)&L.idx,
                  (void*)&L.count, (void*)&L.count_dev, (void*)&L.queue, (void*)&L.out,
                  (void*)&o, (void*)&L.warm};
  return hipLaunchKernel(fn, dim3(L.grid), dim3(nt), args, 0, st);
}

}  // namespace blpg
