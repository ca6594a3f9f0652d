// Tile-shape autotuner for the MFMA GEMM engine.
//
// Each call site offers a few tile configurations of the same product.  The first time a
// (site, precision, shape) key is seen outside graph capture and outside an allocation-free
// call (which take the cached choice, else the first candidate), every candidate is run once and
// then timed twice with HIP events on the context's stream; the fastest is cached for the
// process and used from then on (also inside captured graphs).  All candidates accumulate over
// k in the same order, so the choice never changes the result bits.  Tuning needs an
// idempotent product (output independent of its previous contents); accumulating call sites
// use the first candidate.  AFQ_TUNE=0 in the environment disables tuning; AFQ_TUNE_VERBOSE=1
// prints every candidate's time to stderr.  A site can be pinned to one candidate (tests run
// every candidate of a site this way): Tuner::force(site, i), or AFQ_TUNE_FORCE="site:i,..." in
// the environment; a pinned choice also applies inside graph capture and is never cached.
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>

#include "gemm.hpp"
#include "zgemm_rt.hpp"
#include "ops.hpp"

namespace afq {

class Tuner {
 public:
  static Tuner& get() {
    static Tuner t;
    return t;
  }
  int lookup(const std::string& key) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = best_.find(key);
    return it == best_.end() ? -1 : it->second;
  }
  void store(const std::string& key, int idx, float ms) {
    std::lock_guard<std::mutex> g(mu_);
    best_[key] = idx;
    ms_[key] = ms;
  }
  // candidate pinned for the site of `key` ("site:precision", the text before its second ':',
  // or else "site", the text before its first), or -1
  int forced(const std::string& key, int ncand) {
    std::lock_guard<std::mutex> g(mu_);
    ncand_[key] = ncand;
    if (force_.empty()) return -1;
    const size_t c1 = key.find(':');
    const std::string site = key.substr(0, c1);
    auto it = force_.end();
    if (c1 != std::string::npos) it = force_.find(key.substr(0, key.find(':', c1 + 1)));
    if (it == force_.end()) it = force_.find(site);
    if (it == force_.end()) return -1;
    AFQ_CHECK(it->second < ncand, kErrArg,
              "site " + site + " has " + std::to_string(ncand) + " candidates, " +
                  std::to_string(it->second) + " was forced");
    return it->second;
  }
  // i < 0 removes the pin; site "" removes every pin
  void force(const std::string& site, int i) {
    std::lock_guard<std::mutex> g(mu_);
    if (site.empty()) force_.clear();
    else if (i < 0) force_.erase(site);
    else force_[site] = i;
  }
  bool enabled() const { return enabled_; }
  bool verbose() const { return verbose_; }
  std::string report() {
    std::lock_guard<std::mutex> g(mu_);
    std::string r;
    for (auto& kv : best_)
      r += kv.first + " -> " + std::to_string(kv.second) + " (" + std::to_string(ms_[kv.first]) +
           " ms) of " + std::to_string(ncand_[kv.first]) + "\n";
    for (auto& kv : ncand_)
      if (!best_.count(kv.first))
        r += kv.first + " -> untuned of " + std::to_string(kv.second) + "\n";
    return r;
  }

 private:
  Tuner() {
    const char* e = std::getenv("AFQ_TUNE");
    enabled_ = !(e && e[0] == '0');
    const char* v = std::getenv("AFQ_TUNE_VERBOSE");
    verbose_ = v && v[0] == '1';
    if (const char* f = std::getenv("AFQ_TUNE_FORCE")) {
      std::string spec(f);
      size_t p = 0;
      while (p < spec.size()) {
        size_t q = spec.find(',', p);
        if (q == std::string::npos) q = spec.size();
        const std::string item = spec.substr(p, q - p);
        const size_t c = item.rfind(':');
        if (c != std::string::npos && c > 0) force_[item.substr(0, c)] = std::atoi(item.c_str() + c + 1);
        p = q + 1;
      }
    }
  }
  std::mutex mu_;
  std::map<std::string, int> best_;
  std::map<std::string, float> ms_;
  std::map<std::string, int> ncand_;
  std::map<std::string, int> force_;
  bool enabled_ = true;
  bool verbose_ = false;
};

template <class F>
void tuned(Context& c, const std::string& key, int ncand, F&& launch, bool idempotent) {
  Tuner& t = Tuner::get();
  const int pin = t.forced(key, ncand);
  if (pin >= 0) {
    launch(pin);
    return;
  }
  int choice = t.lookup(key);
  if (choice >= 0) {
    launch(choice);
    return;
  }
  if (ncand <= 1 || !idempotent || !t.enabled() || c.no_alloc() || c.capturing()) {
    launch(0);
    return;
  }
  hipStream_t s = c.stream();
  hipEvent_t e0, e1;
  AFQ_HIP(hipEventCreate(&e0));
  AFQ_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int bi = 0;
  for (int i = 0; i < ncand; ++i) {
    launch(i);
    AFQ_HIP(hipEventRecord(e0, s));
    launch(i);
    launch(i);
    AFQ_HIP(hipEventRecord(e1, s));
    AFQ_HIP(hipEventSynchronize(e1));
    float ms = 0;
    AFQ_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (t.verbose()) std::fprintf(stderr, "tune %s candidate %d: %.4f ms\n", key.c_str(), i, ms / 2);
    if (ms < best) {
      best = ms;
      bi = i;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  t.store(key, bi, best / 2);
  launch(bi);
}

// In-place kernels (the fused propagation): candidates are timed on a scratch copy of the buffer
// they update, which they all turn into the same bits, then the chosen one runs on the buffer.
template <class F>
void tuned_on_copy(Context& c, const std::string& key, int ncand, z64* buf, size_t count, F&& run) {
  Tuner& t = Tuner::get();
  const int pin = t.forced(key, ncand);
  if (pin >= 0) {
    run(pin, buf);
    return;
  }
  int choice = t.lookup(key);
  if (choice >= 0 && choice < ncand) {
    run(choice, buf);
    return;
  }
  if (ncand <= 1 || !t.enabled() || c.no_alloc() || c.capturing()) {
    run(0, buf);
    return;
  }
  hipStream_t s = c.stream();
  z64* tmp = scratch_as<z64>(c, "tune_copy", count);
  hipEvent_t e0, e1;
  AFQ_HIP(hipEventCreate(&e0));
  AFQ_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int bi = 0;
  for (int i = 0; i < ncand; ++i) {
    AFQ_HIP(hipMemcpyAsync(tmp, buf, count * sizeof(z64), hipMemcpyDeviceToDevice, s));
    run(i, tmp);
    AFQ_HIP(hipMemcpyAsync(tmp, buf, count * sizeof(z64), hipMemcpyDeviceToDevice, s));
    AFQ_HIP(hipEventRecord(e0, s));
    run(i, tmp);
    AFQ_HIP(hipEventRecord(e1, s));
    AFQ_HIP(hipEventSynchronize(e1));
    float ms = 0;
    AFQ_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      bi = i;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  t.store(key, bi, best);
  run(bi, buf);
}

// Shapes are bucketed (each extent rounded up to a power of two) so that products of similar
// size, e.g. the walker counts of different jobs, share one tuning decision.
inline int tune_bucket(int x) {
  int b = 1;
  while (b < x) b <<= 1;
  return b;
}
inline std::string tune_key(const char* site, size_t tsize, int batch, int M, int N, int K) {
  return std::string(site) + (tsize == 8 ? ":f64:" : ":f32:") + std::to_string(tune_bucket(batch)) +
         "x" + std::to_string(tune_bucket(M)) + "x" + std::to_string(tune_bucket(N)) + "x" +
         std::to_string(tune_bucket(K));
}

// Large products (vbias, vHS, energy chunks): square-ish 2x2-wave tiles of several sizes.
template <typename T, class OA, class OB, class EPI>
void gemm_large(Context& c, const char* site, int batch, int M, int N, int K, const OA& A,
                const OB& B, const EPI& e, bool idempotent = true) {
  hipStream_t s = c.stream();
  // two LDS buffers in turn (PP) for the two smaller tiles; FP32 also offers 32-deep k slabs
  // (twice the MFMA work per staged slab).  The k order of the accumulation is the same for
  // every candidate.
  constexpr int ncand = sizeof(T) == 4 ? 11 : 8;
  tuned(c, tune_key(site, sizeof(T), batch, M, N, K), ncand, [&](int i) {
    switch (i) {
      // 128 x 128 tiles over 16 and 8 wavefronts (32 x 32 and 32 x 64 per wavefront)
      case 9: gemm::launch<T, 128, 128, 16, 4, 4, true>(s, batch, M, N, K, A, B, e); break;
      case 10: gemm::launch<T, 128, 128, 16, 4, 2, true>(s, batch, M, N, K, A, B, e); break;
      case 0: gemm::launch<T, 64, 64, 16, 2, 2>(s, batch, M, N, K, A, B, e); break;
      case 1: gemm::launch<T, 128, 128, 16, 2, 2>(s, batch, M, N, K, A, B, e); break;
      case 2: gemm::launch<T, 128, 64, 16, 2, 2>(s, batch, M, N, K, A, B, e); break;
      case 3: gemm::launch<T, 64, 128, 16, 2, 2>(s, batch, M, N, K, A, B, e); break;
      case 4: gemm::launch<T, 64, 128, 16, 2, 2, true>(s, batch, M, N, K, A, B, e); break;
      case 5: gemm::launch<T, 64, 64, 16, 2, 2, true>(s, batch, M, N, K, A, B, e); break;
      default:
        if constexpr (sizeof(T) == 4) {
          if (i == 6) gemm::launch<T, 64, 128, 32, 2, 2>(s, batch, M, N, K, A, B, e);
          else if (i == 7) gemm::launch<T, 128, 128, 32, 2, 2>(s, batch, M, N, K, A, B, e);
          else gemm::launch<T, 128, 64, 32, 2, 2>(s, batch, M, N, K, A, B, e);
        } else {
          if (i == 6) gemm::launch<T, 128, 128, 16, 4, 4, true>(s, batch, M, N, K, A, B, e);
          else gemm::launch<T, 128, 128, 16, 4, 2, true>(s, batch, M, N, K, A, B, e);
        }
        break;
    }
  }, idempotent);
}

// Large products whose output grid is small against the machine (vbias: W x M outputs over a
// long K): K is cut into chunks by a rule that depends on the shape only, so every tile
// candidate sums in the same order and the tuned choice still never changes the result.
inline void splitk_plan(int M, int N, int K, int& ksplit, int& kchunk) {
  const long tiles = (long)ceil_div(M, 64) * ceil_div(N, 64);
  int ks = (int)std::max<long>(1, std::min<long>((1024 + tiles - 1) / tiles, K / 128));
  kchunk = ((ceil_div(K, ks) + 63) / 64) * 64;
  ksplit = ceil_div(K, kchunk);
}

template <typename T, class OA, class OB, class EPI>
void gemm_large_splitk(Context& c, const char* site, int M, int N, int K, const OA& A,
                       const OB& B, const EPI& e, bool idempotent = true) {
  hipStream_t s = c.stream();
  int ks, kc;
  splitk_plan(M, N, K, ks, kc);
  cplx<T>* ws = ks > 1 ? scratch_as<cplx<T>>(c, std::string("splitk_ws") + (sizeof(T) == 8 ? "64" : "32"),
                                              (size_t)ks * M * N)
                       : nullptr;
  constexpr int ncand = sizeof(T) == 4 ? 11 : 8;
  tuned(c, tune_key(site, sizeof(T), ks, M, N, K), ncand, [&](int i) {
    switch (i) {
      case 10: gemm::launch_splitk<T, 128, 128, 16, 4, 4, true>(s, 1, M, N, K, ks, kc, ws, A, B, e); break;
      case 0: gemm::launch_splitk<T, 64, 64, 16, 2, 2>(s, 1, M, N, K, ks, kc, ws, A, B, e); break;
      case 1: gemm::launch_splitk<T, 128, 128, 16, 2, 2>(s, 1, M, N, K, ks, kc, ws, A, B, e); break;
      case 2: gemm::launch_splitk<T, 128, 64, 16, 2, 2>(s, 1, M, N, K, ks, kc, ws, A, B, e); break;
      case 3: gemm::launch_splitk<T, 64, 128, 16, 2, 2>(s, 1, M, N, K, ks, kc, ws, A, B, e); break;
      case 4: gemm::launch_splitk<T, 32, 32, 16, 1, 1>(s, 1, M, N, K, ks, kc, ws, A, B, e); break;
      case 5: gemm::launch_splitk<T, 64, 64, 16, 2, 2, true>(s, 1, M, N, K, ks, kc, ws, A, B, e); break;
      case 6: gemm::launch_splitk<T, 64, 128, 16, 2, 2, true>(s, 1, M, N, K, ks, kc, ws, A, B, e); break;
      default:
        if constexpr (sizeof(T) == 4) {
          if (i == 7) gemm::launch_splitk<T, 64, 64, 32, 2, 2>(s, 1, M, N, K, ks, kc, ws, A, B, e);
          else if (i == 8) gemm::launch_splitk<T, 128, 128, 32, 2, 2>(s, 1, M, N, K, ks, kc, ws, A, B, e);
          else gemm::launch_splitk<T, 128, 64, 32, 2, 2>(s, 1, M, N, K, ks, kc, ws, A, B, e);
        } else {
          gemm::launch_splitk<T, 128, 128, 16, 4, 4, true>(s, 1, M, N, K, ks, kc, ws, A, B, e);
        }
        break;
    }
  }, idempotent);
}

// Tall x narrow per-walker products (orbitals x electrons): BN fitted to the electron count,
// BM = 64 rows or a full-height 208-row tile (one workgroup per walker when N <= 208).
template <typename T, int BM, int WM, class OA, class OB, class EPI>
void tall_bn(hipStream_t s, int batch, int M, int N, int K, const OA& A, const OB& B,
             const EPI& e) {
  if (N <= 16)
    gemm::launch<T, BM, 16, 16, WM, 1>(s, batch, M, N, K, A, B, e);
  else if (N <= 32)
    gemm::launch<T, BM, 32, 16, WM, 1>(s, batch, M, N, K, A, B, e);
  else if (N <= 48)
    gemm::launch<T, BM, 48, 16, WM, 1>(s, batch, M, N, K, A, B, e);
  else
    gemm::launch<T, BM, 64, 16, WM, 1>(s, batch, M, N, K, A, B, e);
}

template <typename T, class OA, class OB, class EPI>
void gemm_tall(Context& c, const char* site, int batch, int M, int N, int K, const OA& A,
               const OB& B, const EPI& e, bool idempotent = true) {
  hipStream_t s = c.stream();
  tuned(c, tune_key(site, sizeof(T), batch, M, N, K), 3, [&](int i) {
    switch (i) {
      case 0: tall_bn<T, 64, 4>(s, batch, M, N, K, A, B, e); break;
      case 1: tall_bn<T, 208, 13>(s, batch, M, N, K, A, B, e); break;
      default: tall_bn<T, 128, 8>(s, batch, M, N, K, A, B, e); break;
    }
  }, idempotent);
}

// Small per-walker products (electrons x electrons, electrons x orbitals).
template <typename T, class OA, class OB, class EPI>
void gemm_small(Context& c, const char* site, int batch, int M, int N, int K, const OA& A,
                const OB& B, const EPI& e, bool idempotent = true) {
  hipStream_t s = c.stream();
  tuned(c, tune_key(site, sizeof(T), batch, M, N, K), 2, [&](int i) {
    switch (i) {
      case 0: gemm::launch<T, 32, 32, 16, 1, 1>(s, batch, M, N, K, A, B, e); break;
      default: gemm::launch<T, 48, 48, 16, 1, 1>(s, batch, M, N, K, A, B, e); break;
    }
  }, idempotent);
}

// A small batched complex FP64 product: the register-tiled kernel (its blockings tuned per
// shape) where it applies, else the LDS-staged engine.
template <class OA, class OB, class EPI>
void zgemm_small(Context& c, const char* site, int W, int M, int N, int K, const OA& A,
                 const OB& B, const EPI& e) {
  if (!zrt::supported(M, N, K, A, B)) {
    gemm_small<double>(c, site, W, M, N, K, A, B, e);
    return;
  }
  hipStream_t s = c.stream();
  tuned(c, tune_key((std::string(site) + "_rt").c_str(), sizeof(double), W, M, N, K),
        zrt::kCandidates, [&](int i) { zrt::launch(s, W, M, N, K, A, B, e, i); }, true);
}

}  // namespace afq
