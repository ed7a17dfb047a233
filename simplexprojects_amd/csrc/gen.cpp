// gen.cpp -- seeded synthetic LP generators (host code, counter-based RNG).
//
// Every random number is a pure function of (seed, lp key, field, index), so any
// subset of a batch can be regenerated bit-identically on any host.  The reference
// generates its random LPs with srand(time) + rand() (parseBenchmark.cpp:321-376) and
// never reaches the GPU with them (GPU_Tester.cu:67-79); these generators replace it.
//
//   config 1  m x n, A ~ U[0,1) with 30% of entries zeroed, b ~ U[1,10), c ~ U[0,1)
//   config 2  A ~ U[-1,1) with 40% zeroed, b ~ U[0,10) with floor(0.1 m) rows (seeded
//             choice) set to 0, c ~ N(0,1); floor(0.05*batch) LPs (seeded choice) get one
//             column j with A[:,j] = -|A[:,j]| and c_j = |c_j| + 0.1 (unbounded)
//   config 3  A ~ U[-1,1) with 50% zeroed, b ~ N(0,5), c ~ N(0,1); round(0.3 m) rows
//             (seeded choice) are equalities, the rest '<='; floor(0.03*batch) LPs
//             (seeded choice) get rows i,k with a_k = a_i, row i: a_i x <= -1 and
//             row k: a_i x >= 1 (infeasible)
//   config 4  parametric family `shape_idx`: A_f ~ U[0,1) with 30% zeroed, base b_f ~
//             U[1,10), base c_f ~ U[0,1); instance k: b = b_f (1 + 0.05 xi),
//             c = c_f (1 + 0.05 zeta), xi, zeta ~ N(0,1)
//   config 5  single LPs for the serving loop: 27x32 uses config 1, other shapes config 2
//             (no planted unbounded columns)
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/blpg/blpg.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

enum Field : uint64_t {
  F_A = 1, F_MASK = 2, F_B = 3, F_C = 4, F_BZERO = 5, F_UNB_PICK = 6, F_UNB_COL = 7,
  F_EQ = 8, F_INF_PICK = 9, F_INF_ROWS = 10, F_NOISE_B = 11, F_NOISE_C = 12, F_FAM = 13
};

struct Rng {
  uint64_t base;
  Rng(uint64_t seed, uint64_t key, uint64_t field) {
    base = mix64(mix64(mix64(seed) ^ key) ^ (field * 0x100000001B3ull));
  }
  uint64_t bits(uint64_t i) const { return mix64(base ^ (i * 0xD1B54A32D192ED03ull)); }
  double u(uint64_t i) const { return (double)(bits(i) >> 11) * 0x1.0p-53; }  // [0,1)
  double normal(uint64_t i) const {  // Box-Muller on two counters
    const double u1 = u(2 * i), u2 = u(2 * i + 1);
    return std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(6.283185307179586 * u2);
  }
};

// first `k` entries of a seeded Fisher-Yates permutation of [0, n)
std::vector<int64_t> pick(const Rng& r, int64_t n, int64_t k) {
  std::vector<int64_t> p(n);
  for (int64_t i = 0; i < n; ++i) p[i] = i;
  for (int64_t i = 0; i < k && i < n; ++i) {
    const int64_t j = i + (int64_t)(r.bits(i) % (uint64_t)(n - i));
    std::swap(p[i], p[j]);
  }
  p.resize(k < n ? k : n);
  return p;
}

inline uint64_t lp_key(int32_t shape_idx, int64_t lp) {
  return ((uint64_t)(uint32_t)shape_idx << 40) ^ (uint64_t)lp;
}

void gen_dense(uint64_t seed, uint64_t key, int m, int n, double lo, double hi, double keep,
               double* A) {
  const Rng ra(seed, key, F_A), rm(seed, key, F_MASK);
  for (int64_t e = 0; e < (int64_t)m * n; ++e) {
    const double v = lo + (hi - lo) * ra.u(e);
    A[e] = rm.u(e) < keep ? v : 0.0;
  }
}

}  // namespace

extern "C" int blpg_gen_config(int32_t config, uint64_t seed, int32_t shape_idx, int64_t first,
                               int64_t count, int64_t batch_count, int32_t m, int32_t n,
                               double* A, double* b, double* c, int8_t* sense, int32_t* flags) {
  if (m < 1 || n < 1 || count < 0 || first < 0 || !A || !b || !c) return BLPG_E_INVALID;
  if (batch_count < first + count) batch_count = first + count;
  std::vector<char> unb(batch_count, 0), inf(batch_count, 0);
  if (config == 2) {
    const Rng r(seed, (uint64_t)(uint32_t)shape_idx, F_UNB_PICK);
    for (int64_t k : pick(r, batch_count, batch_count / 20)) unb[k] = 1;
  } else if (config == 3) {
    const Rng r(seed, (uint64_t)(uint32_t)shape_idx, F_INF_PICK);
    for (int64_t k : pick(r, batch_count, (batch_count * 3) / 100)) inf[k] = 1;
  } else if (config < 1 || config > 5) {
    return BLPG_E_INVALID;
  }
  for (int64_t q = 0; q < count; ++q) {
    const int64_t lp = first + q;
    double* a = A + q * (int64_t)m * n;
    double* bb = b + q * (int64_t)m;
    double* cc = c + q * (int64_t)n;
    int8_t* ss = sense ? sense + q * (int64_t)m : nullptr;
    if (ss) std::memset(ss, 0, (size_t)m);
    int32_t fl = 0;
    const uint64_t key = lp_key(shape_idx, lp);
    const int cfg = (config == 5) ? ((m == 27 && n == 32) ? 1 : 2) : config;
    if (cfg == 1) {
      gen_dense(seed, key, m, n, 0.0, 1.0, 0.7, a);
      const Rng rb(seed, key, F_B), rc(seed, key, F_C);
      for (int i = 0; i < m; ++i) bb[i] = 1.0 + 9.0 * rb.u(i);
      for (int j = 0; j < n; ++j) cc[j] = rc.u(j);
    } else if (cfg == 2) {
      gen_dense(seed, key, m, n, -1.0, 1.0, 0.6, a);
      const Rng rb(seed, key, F_B), rc(seed, key, F_C), rz(seed, key, F_BZERO);
      for (int i = 0; i < m; ++i) bb[i] = 10.0 * rb.u(i);
      for (int64_t i : pick(rz, m, m / 10)) bb[i] = 0.0;
      for (int j = 0; j < n; ++j) cc[j] = rc.normal(j);
      if (config == 2 && unb[lp]) {
        const Rng rj(seed, key, F_UNB_COL);
        const int j = (int)(rj.bits(0) % (uint64_t)n);
        for (int i = 0; i < m; ++i) a[(int64_t)i * n + j] = -std::fabs(a[(int64_t)i * n + j]);
        cc[j] = std::fabs(cc[j]) + 0.1;
        fl |= 1;
      }
    } else if (cfg == 3) {
      gen_dense(seed, key, m, n, -1.0, 1.0, 0.5, a);
      const Rng rb(seed, key, F_B), rc(seed, key, F_C), re(seed, key, F_EQ);
      for (int i = 0; i < m; ++i) bb[i] = 5.0 * rb.normal(i);
      for (int j = 0; j < n; ++j) cc[j] = rc.normal(j);
      const int neq = (int)std::lround(0.3 * m);
      if (ss)
        for (int64_t i : pick(re, m, neq)) ss[i] = 1;
      if (inf[lp]) {
        const Rng rr(seed, key, F_INF_ROWS);
        const std::vector<int64_t> rows = pick(rr, m, 2);
        if (rows.size() == 2) {
          const int64_t i = rows[0], k = rows[1];
          for (int j = 0; j < n; ++j) a[k * n + j] = a[i * n + j];
          bb[i] = -1.0;
          bb[k] = 1.0;
          if (ss) { ss[i] = 0; ss[k] = 2; }
          fl |= 2;
        }
      }
    } else if (cfg == 4) {
      const uint64_t fkey = lp_key(shape_idx, 0x7FFFFFFFFFll);  // family-level stream
      gen_dense(seed, fkey, m, n, 0.0, 1.0, 0.7, a);
      const Rng rb(seed, fkey, F_B), rc(seed, fkey, F_C);
      const Rng nb(seed, key, F_NOISE_B), nc(seed, key, F_NOISE_C);
      for (int i = 0; i < m; ++i) bb[i] = (1.0 + 9.0 * rb.u(i)) * (1.0 + 0.05 * nb.normal(i));
      for (int j = 0; j < n; ++j) cc[j] = rc.u(j) * (1.0 + 0.05 * nc.normal(j));
    }
    if (flags) flags[q] = fl;
  }
  return BLPG_OK;
}
