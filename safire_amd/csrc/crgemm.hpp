// Register-tiled product of a complex operand and a real operand, both K-contiguous, in FP32 or
// FP64:
//
//   C(m, n) = sum_k A(m, k) * B(n, k)      A complex [M][lda], B real [N rows][K]
//
// the shape of vHS (A = the fields X [W][M], B = the Cholesky vectors' symmetric pairs) and of
// vbias (A = the walkers' G, B = the half-rotated Cholesky vectors).  Each wavefront owns a
// 16 RB x 16 CB tile of C in MFMA accumulators (v_mfma_f32_16x16x4f32 or v_mfma_f64_16x16x4f64,
// 2 MFMAs per block: re(A) B and im(A) B) and streams its operands from memory straight into
// MFMA fragments: lane (fr, fk) of k group g feeds MFMA step j with k = 16 g + 4 fk + j, four
// consecutive k values per lane, so the fragments are whole 16-byte words.  There is no LDS and
// no barrier: per k group a wavefront issues its fragment loads and 8 RB CB MFMAs, the next
// group's loads issued before the current group's MFMAs (full groups load unconditionally, from
// rows clamped into range, so the compiler waits for them only after the MFMAs).
//
// Rows past M and columns past N (or symmetric pairs outside the N x N space) read a clamped,
// valid row and are never stored.  Operands are addressed by 32-bit byte offsets from uniform
// bases (crg_supported checks the sizes).  The ragged last k group is taken first; with REVK the
// full groups then run from last to first, so an accumulation over Cholesky vectors ordered by
// decreasing norm adds the small terms first.  Split-K slices write FP32 partials that kreduce_splitk sums in slice
// order; every result depends on the shape only.
#pragma once

#include "gemm.hpp"

namespace afq {
namespace crg {

constexpr int kWaves = 4;

// B rows: plain rows of a [rows][ldb] array
struct RowsPlain {
  long ldb;  // elements between rows
  __device__ inline bool row(int n, int N, long& off) const {
    off = (long)n * ldb;
    return n < N;
  }
};

// B rows: the symmetric pairs (i, k) of a [N*N][ldb] array, in the order of gemm::sym_pair
struct RowsSymPairs {
  long ldb;
  int Nn, nb;
  __device__ inline bool row(int n, int, long& off) const {
    int i, k;
    const bool ok = gemm::sym_pair(n, nb, Nn, i, k);
    off = ok ? ((long)i * Nn + k) * ldb : 0;
    return ok;
  }
};

template <typename T, int RB, int CB>
struct Frag {
  T ar[RB][4], ai[RB][4];  // A rows: re and im of four k values
  T b[CB][4];              // B rows
};

template <int RB, int CB>
struct Ofs {
  uint32_t a[RB], b[CB];  // byte offsets of the lane's rows at its first k value
};

// the fragments of one k group; ab / bb: the operands' bases advanced to the group's first k.
// FULL: unconditional loads; else rem = K - (the lane's first k) masks k >= K to zero.
template <typename T, int RB, int CB, bool FULL>
__device__ inline void load(Frag<T, RB, CB>& f, const char* ab, const char* bb,
                            const Ofs<RB, CB>& o, int rem) {
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    if constexpr (sizeof(T) == 4) {
      const float4 z = {0.f, 0.f, 0.f, 0.f};
      float4 x = z, y = z;
      if (FULL || rem >= 4) {
        x = *reinterpret_cast<const float4*>(ab + o.a[i]);
        y = *reinterpret_cast<const float4*>(ab + o.a[i] + 16);
      } else {
        float2 t[4] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (rem > j) t[j] = *reinterpret_cast<const float2*>(ab + o.a[i] + 8 * j);
        x = float4{t[0].x, t[0].y, t[1].x, t[1].y};
        y = float4{t[2].x, t[2].y, t[3].x, t[3].y};
      }
      f.ar[i][0] = x.x; f.ai[i][0] = x.y; f.ar[i][1] = x.z; f.ai[i][1] = x.w;
      f.ar[i][2] = y.x; f.ai[i][2] = y.y; f.ar[i][3] = y.z; f.ai[i][3] = y.w;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        double2 v = {0.0, 0.0};
        if (FULL || rem > j) v = *reinterpret_cast<const double2*>(ab + o.a[i] + 16 * j);
        f.ar[i][j] = v.x;
        f.ai[i][j] = v.y;
      }
    }
  }
#pragma unroll
  for (int jb = 0; jb < CB; ++jb) {
    if constexpr (sizeof(T) == 4) {
      float4 x = {0.f, 0.f, 0.f, 0.f};
      if (FULL || rem >= 4) {
        x = *reinterpret_cast<const float4*>(bb + o.b[jb]);
      } else {
        float t[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (rem > j) t[j] = *reinterpret_cast<const float*>(bb + o.b[jb] + 4 * j);
        x = float4{t[0], t[1], t[2], t[3]};
      }
      f.b[jb][0] = x.x; f.b[jb][1] = x.y; f.b[jb][2] = x.z; f.b[jb][3] = x.w;
    } else {
      double2 x = {0.0, 0.0}, y = {0.0, 0.0};
      if (FULL || rem >= 4) {
        x = *reinterpret_cast<const double2*>(bb + o.b[jb]);
        y = *reinterpret_cast<const double2*>(bb + o.b[jb] + 16);
      } else {
        if (rem > 0) x.x = *reinterpret_cast<const double*>(bb + o.b[jb]);
        if (rem > 1) x.y = *reinterpret_cast<const double*>(bb + o.b[jb] + 8);
        if (rem > 2) y.x = *reinterpret_cast<const double*>(bb + o.b[jb] + 16);
      }
      f.b[jb][0] = x.x; f.b[jb][1] = x.y; f.b[jb][2] = y.x; f.b[jb][3] = y.y;
    }
  }
}

template <typename T, int RB, int CB>
__device__ inline void mma(const Frag<T, RB, CB>& f, typename gemm::Acc<T>::v4 (&cr)[RB][CB],
                           typename gemm::Acc<T>::v4 (&ci)[RB][CB]) {
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int jb = 0; jb < CB; ++jb) {
        cr[i][jb] = gemm::Acc<T>::mfma(f.ar[i][j], f.b[jb][j], cr[i][jb]);
        ci[i][jb] = gemm::Acc<T>::mfma(f.ai[i][j], f.b[jb][j], ci[i][jb]);
      }
}

// One wavefront per tile (tile m fastest, then n, then the K slice); the workgroups are
// renumbered so that each XCD takes a contiguous run of tiles, i.e. shares the B rows of a few
// column tiles through its L2.
template <typename T, int RB, int CB, bool REVK, class BROWS, class EPI>
__global__ void __launch_bounds__(64 * kWaves)
    k_crgemm(int M, int N, int K, const cplx<T>* __restrict__ A, long lda,
             const T* __restrict__ B, BROWS brows, int tm, int tn, int ksplit, int kchunk,
             EPI epi) {
  using V4 = typename gemm::Acc<T>::v4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int fr = lane & 15, fk = lane >> 4;
  const long total = gridDim.x, xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const long wg = xcd * (total / 8) + min(xcd, total % 8) + slot;
  const long q = wg * kWaves + wave;
  if (q >= (long)tm * tn * ksplit) return;  // uniform per wavefront; no barrier in the kernel
  const int im = (int)(q % tm), in = (int)((q / tm) % tn), ks = (int)(q / ((long)tm * tn));
  const int m0 = 16 * RB * im, n0 = 16 * CB * in;
  const int k0 = ks * kchunk, kn = min(K - k0, kchunk);  // this slice: k0 .. k0 + kn - 1
  Ofs<RB, CB> o;
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    const int r = min(m0 + 16 * i + fr, M - 1);
    o.a[i] = (uint32_t)(((long)r * lda + k0 + 4 * fk) * (long)sizeof(cplx<T>));
  }
#pragma unroll
  for (int jb = 0; jb < CB; ++jb) {
    long off;
    const int c = n0 + 16 * jb + fr;
    if (!brows.row(min(c, N - 1), N, off)) brows.row(0, N, off);  // a valid row stands in
    o.b[jb] = (uint32_t)((off + k0 + 4 * fk) * (long)sizeof(T));
  }
  V4 cr[RB][CB], ci[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int jb = 0; jb < CB; ++jb) {
      cr[i][jb] = V4{0, 0, 0, 0};
      ci[i][jb] = V4{0, 0, 0, 0};
    }
  const char* ab = reinterpret_cast<const char*>(A);
  const char* bb = reinterpret_cast<const char*>(B);
  constexpr long AS = 16 * sizeof(cplx<T>), BS = 16 * sizeof(T);  // bytes per k group
  const int NGF = kn / 16;                                      // full k groups of the slice
  const bool tail = NGF * 16 < kn;
  Frag<T, RB, CB> cur, nxt;
  if (tail) {  // the ragged (last) group first, before the main loop's registers are live
    load<T, RB, CB, false>(cur, ab + NGF * AS, bb + NGF * BS, o, kn - 16 * NGF - 4 * fk);
    mma<T, RB, CB>(cur, cr, ci);
  }
  if (NGF > 0) {
    // group order: 0 .. NGF - 1, or NGF - 1 .. 0 with REVK
    auto gid = [&](int t) { return REVK ? NGF - 1 - t : t; };
    load<T, RB, CB, true>(cur, ab + gid(0) * AS, bb + gid(0) * BS, o, 0);
    for (int t = 0; t < NGF; ++t) {
      const int gn = gid(min(t + 1, NGF - 1));  // the last group reloads itself
      load<T, RB, CB, true>(nxt, ab + gn * AS, bb + gn * BS, o, 0);
      // the prefetch stays ahead of the MFMAs and the copy behind them; the wavefront in its
      // MFMA phase issues ahead of those still loading
      __builtin_amdgcn_sched_barrier(0);
      __builtin_amdgcn_s_setprio(1);
      mma<T, RB, CB>(cur, cr, ci);
      __builtin_amdgcn_s_setprio(0);
      __builtin_amdgcn_sched_barrier(0);
      cur = nxt;
    }
  }
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int jb = 0; jb < CB; ++jb) {
      const int n = n0 + 16 * jb + fr;
      if (n >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 16 * i + gemm::Acc<T>::row(fk, r);
        if (m < M) epi(ks, m, n, cr[i][jb][r], ci[i][jb][r]);
      }
    }
}

// The kernel's requirements: 32-bit byte offsets reach every row of A and B, and the fragment
// loads are 16-byte aligned (the bases 16-byte aligned, A rows of an even number of complex
// elements in FP32, B rows of a multiple of 16 bytes).
template <typename T>
inline bool crg_supported(const void* A, long lda, long a_rows, const void* B, long ldb,
                          long b_rows) {
  return a_rows * lda * (long)sizeof(cplx<T>) < (1L << 31) &&
         b_rows * ldb * (long)sizeof(T) < (1L << 31) && (lda * sizeof(cplx<T>)) % 16 == 0 &&
         (ldb * sizeof(T)) % 16 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0 &&
         reinterpret_cast<uintptr_t>(B) % 16 == 0;
}

// The K slices of a product over M x N outputs with 16 RB x 16 CB tiles: the count whose tiles
// fill the last round of the machine's wavefront slots (256 CUs x 4 SIMDs, one wavefront each)
// best, slices of at least 256 k values and at most 8 slices, the smallest count within 2% of
// the best; the rule depends on the shape only.
inline void crg_split(int M, int N, int K, int RB, int CB, int& ksplit, int& kchunk) {
  constexpr long kSlots = 1024;
  const long tiles = (long)ceil_div(M, 16 * RB) * ceil_div(N, 16 * CB);
  // at most 8 slices: the partials cost 2 ksplit M N complex values of traffic
  const int cap = (int)std::max<long>(1, std::min<long>(K / 256, 8));
  double best = 0.0;
  int bks = 1;
  for (int ks = 1; ks <= cap; ++ks) {
    const double r = (double)(tiles * ks) / kSlots;
    const double eff = r / std::ceil(r);
    if (eff > best * 1.02) {
      best = eff;
      bks = ks;
    }
  }
  kchunk = ceil_div(ceil_div(K, bks), 16) * 16;
  ksplit = ceil_div(K, kchunk);
}

// Whether the register-tiled kernel suits a product whose M x N outputs, in ksplit K slices,
// form the 16 RB x 16 CB wavefront tiles: below half a round of the machine's 1024 wavefront
// slots its 64-row tiles leave most SIMDs idle and the LDS-staged engine's smaller tuned tiles
// are faster (vHS at N = 200, M = 1600 and 128 walkers: 360 tiles).
inline bool crg_fills(int M, int N, int RB, int CB, int ksplit = 1) {
  constexpr long kMinTiles = 512;
  return (long)ceil_div(M, 16 * RB) * ceil_div(N, 16 * CB) * ksplit >= kMinTiles;
}

// C = epi(A B^T) over M x N with K cut into ksplit slices of kchunk (a multiple of 16); more than
// one slice writes partials into ws [ksplit][M][N] and sums them in slice order (in FP64)
// before the epilogue.
template <typename T, int RB, int CB, bool REVK, class BROWS, class EPI>
void launch_crgemm(hipStream_t s, int M, int N, int K, const cplx<T>* A, long lda, const T* B,
                   const BROWS& brows, int ksplit, int kchunk, const EPI& epi, cplx<T>* ws) {
  if (M <= 0 || N <= 0) return;
  AFQ_CHECK(K > 0 && kchunk % 16 == 0 && (long)kchunk * ksplit >= K &&
                (ksplit == 1 || (long)kchunk * (ksplit - 1) < K),
            kErrArg, "bad K slicing");
  const int tm = ceil_div(M, 16 * RB), tn = ceil_div(N, 16 * CB);
  const long tiles = (long)tm * tn * ksplit;
  const long blocks = ceil_div(tiles, (long)kWaves);
  AFQ_CHECK(blocks < (1L << 31), kErrShape, "grid too large");
  if (ksplit == 1) {
    hipLaunchKernelGGL((k_crgemm<T, RB, CB, REVK, BROWS, EPI>), dim3((unsigned)blocks),
                       dim3(64 * kWaves), 0, s, M, N, K, A, lda, B, brows, tm, tn, 1, kchunk, epi);
    AFQ_LAUNCH_CHECK();
    return;
  }
  AFQ_CHECK(ws != nullptr, kErrArg, "split-K workspace");
  gemm::EpiPartial<T> ep{ws, (long)M * N, N};
  hipLaunchKernelGGL((k_crgemm<T, RB, CB, REVK, BROWS, gemm::EpiPartial<T>>),
                     dim3((unsigned)blocks), dim3(64 * kWaves), 0, s, M, N, K, A, lda, B, brows, tm,
                     tn, ksplit, kchunk, ep);
  AFQ_LAUNCH_CHECK();
  const long tot = (long)M * N;
  hipLaunchKernelGGL((gemm::kreduce_splitk<T, EPI>), dim3((unsigned)ceil_div(tot, 256L)),
                     dim3(256), 0, s, 1, M, N, ksplit, ws, epi);
  AFQ_LAUNCH_CHECK();
}

}  // namespace crg
}  // namespace afq
