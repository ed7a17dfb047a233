// tuner.hpp -- per-size-bin kernel routing.
//
// For each LP shape the tuner picks one of:
//   WAVE   the column-per-lane register kernel (smallest variant whose capacity covers
//          (m+2) x (n+2) in the precision),
//   TILE   the 2-D register-tile kernel (smallest tile variant whose capacity covers
//          (m+2) x (n+2) and whose registers fit),
//   LDS    the streamed kernel with the tableau in LDS (when it fits the CU's LDS),
//   STREAM the streamed kernel with the tableau in an HBM slab (any size).
// All three perform identical arithmetic, so the choice changes speed, never results.
// A measured preference table (set_preference, filled by blpg_autotune or imported as
// text) overrides the static order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <string>
#include <utility>

#include "../../include/blpg/blpg.h"
#include "kernels.hpp"

namespace blpg {

struct Route {
  int kernel;   // BLPG_KERNEL_TILE / LDS / STREAM
  int variant;  // tile variant
  int nt;       // stream threads per workgroup
  int rcap;
  int bin;      // shape class (distinct bins for distinct shapes)
};

class Tuner {
 public:
  void init(const hipDeviceProp_t& p) {
    lds_ = p.sharedMemPerBlock > 0 ? (size_t)p.sharedMemPerBlock : 65536;
  }
  size_t lds_limit() const { return lds_; }

  static int stream_threads(int m, int n) {
    const int work = (m + 2) * (n + 2);
    if (work <= 64 * 40) return 64;
    if (work <= 128 * 48) return 128;
    return 256;
  }

  // preference key: (precision, m, n) -> (kernel, variant or nt), with the measured time
  void set_preference(int precision, int m, int n, int kernel, int param, double ms = -1.0) {
    pref_[key(precision, m, n)] = {kernel, param};
    ms_[key(precision, m, n)] = ms;
  }
  void clear_preferences() { pref_.clear(); ms_.clear(); chk_.clear(); }
  // check plan of a shape (measured by the autotuner): bit 0 skips the first check tier, so
  // the shape's LPs go straight to the next (for workloads whose final bases never fit it)
  void set_check(int precision, int m, int n, int flags, double ms = -1.0) {
    chk_[key(precision, m, n)] = {flags, ms};
  }
  int check_flags(int precision, int m, int n) const {
    const auto it = chk_.find(key(precision, m, n));
    return it == chk_.end() ? 0 : it->second.first;
  }
  void erase_check(int precision, int m, int n) { chk_.erase(key(precision, m, n)); }
  // the stored preference of a shape (false: none)
  bool get_preference(int precision, int m, int n, int& kernel, int& param, double& ms) const {
    const auto it = pref_.find(key(precision, m, n));
    if (it == pref_.end()) return false;
    kernel = it->second.first;
    param = it->second.second;
    const auto jt = ms_.find(key(precision, m, n));
    ms = jt == ms_.end() ? -1.0 : jt->second;
    return true;
  }
  void erase_preference(int precision, int m, int n) {
    pref_.erase(key(precision, m, n));
    ms_.erase(key(precision, m, n));
  }
  // "precision m n kernel param ms" lines; kernel 0: the shape's check plan (param: flags,
  // ms: the check's measured time)
  std::string export_text() const {
    std::string out;
    char line[128];
    for (const auto& kv : pref_) {
      const long k = kv.first;
      const int p = (int)(k >> 40), m = (int)((k >> 20) & 0xFFFFF), n = (int)(k & 0xFFFFF);
      const auto it = ms_.find(k);
      snprintf(line, sizeof line, "%d %d %d %d %d %.6f\n", p, m, n, kv.second.first,
               kv.second.second, it == ms_.end() ? -1.0 : it->second);
      out += line;
    }
    for (const auto& kv : chk_) {
      const long k = kv.first;
      const int p = (int)(k >> 40), m = (int)((k >> 20) & 0xFFFFF), n = (int)(k & 0xFFFFF);
      snprintf(line, sizeof line, "%d %d %d 0 %d %.6f\n", p, m, n, kv.second.first,
               kv.second.second);
      out += line;
    }
    return out;
  }
  bool import_text(const char* text) {
    const char* p = text;
    while (p && *p) {
      int pr, m, n, kern, param, used = 0;
      double ms = -1;
      if (sscanf(p, "%d %d %d %d %d %lf%n", &pr, &m, &n, &kern, &param, &ms, &used) < 6) {
        while (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t') ++p;
        if (!*p) break;
        return false;
      }
      if (pr < 0 || pr > 1 || m < 1 || n < 1 || m > 4096 || n > 4096 || kern < 0 || kern > 4)
        return false;
      if (kern == 0) {
        if (param < 0 || param > 1) return false;
        set_check(pr, m, n, param, ms);
      } else {
        set_preference(pr, m, n, kern, param, ms);
      }
      p += used;
    }
    return true;
  }

  Route route(int m, int n, int precision, int forced) const {
    Route r{};
    r.bin = ((m + 15) / 16) * 1024 + (n + 15) / 16;
    r.rcap = m + 2;
    r.nt = stream_threads(m, n);
    const int esz = precision == BLPG_FP64 ? 8 : 4;
    const int tv = best_tile(m, n, precision);
    const int wv = best_wave(m, n, precision);
    const size_t lds_need =
        stream_lds_bytes(precision, true, r.nt, m + 2, n + 2);
    const bool lds_ok = lds_need <= lds_ && stream_max_blocks_per_cu(precision, true, r.nt, m + 2, n + 2) > 0;
    (void)esz;
    auto it = pref_.find(key(precision, m, n));
    if (forced == BLPG_KERNEL_AUTO && it != pref_.end()) {
      const int k = it->second.first;
      if (k == BLPG_KERNEL_WAVE && it->second.second >= 0 &&
          it->second.second < wave_variant_count() && fits_wave(it->second.second, m, n, precision)) {
        r.kernel = k; r.variant = it->second.second; return r;
      }
      if (k == BLPG_KERNEL_TILE && it->second.second >= 0 &&
          it->second.second < tile_variant_count(precision) &&
          fits_tile(it->second.second, m, n, precision)) {
        r.kernel = k; r.variant = it->second.second; return r;
      }
      if (k == BLPG_KERNEL_LDS && lds_ok) { r.kernel = k; if (it->second.second) r.nt = it->second.second; return r; }
      if (k == BLPG_KERNEL_STREAM) { r.kernel = k; if (it->second.second) r.nt = it->second.second; return r; }
    }
    if ((forced == BLPG_KERNEL_AUTO || forced == BLPG_KERNEL_WAVE) && wv >= 0) {
      r.kernel = BLPG_KERNEL_WAVE;
      r.variant = wv;
      return r;
    }
    if ((forced == BLPG_KERNEL_AUTO || forced == BLPG_KERNEL_TILE || forced == BLPG_KERNEL_WAVE) &&
        tv >= 0) {
      r.kernel = BLPG_KERNEL_TILE;
      r.variant = tv;
      return r;
    }
    if ((forced == BLPG_KERNEL_AUTO || forced == BLPG_KERNEL_LDS || forced == BLPG_KERNEL_TILE ||
         forced == BLPG_KERNEL_WAVE) &&
        lds_ok) {
      r.kernel = BLPG_KERNEL_LDS;
      return r;
    }
    r.kernel = BLPG_KERNEL_STREAM;
    return r;
  }

  static bool fits_wave(int v, int m, int n, int precision) {
    const WaveVariant w = wave_variant(precision, v);
    return w.enabled && w.rcap() >= m + 2 && w.ccap() >= n + 2;
  }

  // smallest-capacity wave variant that holds the LP (ties: fewer wavefronts per LP)
  static int best_wave(int m, int n, int precision) {
    int best = -1;
    long best_cap = 0;
    int best_nw = 0;
    for (int v = 0; v < wave_variant_count(); ++v) {
      if (!fits_wave(v, m, n, precision)) continue;
      const WaveVariant w = wave_variant(precision, v);
      const long cap = (long)w.rcap() * w.ccap();
      if (best < 0 || cap < best_cap || (cap == best_cap && w.nw < best_nw)) {
        best = v; best_cap = cap; best_nw = w.nw;
      }
    }
    return best;
  }

  static bool fits_tile(int v, int m, int n, int precision) {
    const TileVariant t = tile_variant(precision, v);
    return t.rcap() >= m + 2 && t.ccap() >= n + 2;
  }

  // smallest-capacity tile variant of the default set that holds the LP (the remaining
  // variants are chosen only by measurement)
  static int best_tile(int m, int n, int precision) {
    int best = -1;
    long best_cap = 0;
    for (int v = 0; v < tile_default_variants(); ++v) {
      if (!fits_tile(v, m, n, precision)) continue;
      const TileVariant t = tile_variant(precision, v);
      const long cap = (long)t.rcap() * t.ccap();
      if (best < 0 || cap < best_cap) { best = v; best_cap = cap; }
    }
    return best;
  }

 private:
  static long key(int p, int m, int n) { return ((long)p << 40) | ((long)m << 20) | n; }
  size_t lds_ = 65536;
  std::map<long, std::pair<int, int>> pref_;
  std::map<long, double> ms_;
  std::map<long, std::pair<int, double>> chk_;  // check plan flags, measured check time
};

}  // namespace blpg
