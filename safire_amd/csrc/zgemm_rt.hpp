// Register-tiled batched product of two complex FP64 operands, for the small per-walker products
// of the mixed density matrix and the overlap (T = Psi^H Phi: n x n over K = N; G = T^-T Phi^T:
// n x N over K = n):
//
//   C_b(m, n) = sum_k A_b(m, k) * B_b(n, k)        (gemm::Opnd views of complex operands)
//
// One wavefront per (batch entry, block of RB 16-row tiles, block of CB 16-column tiles) holds its
// RB x CB tiles of C in MFMA accumulators, four v_mfma_f64_16x16x4f64 per complex block.  Lane
// (fr, fk) feeds k = k0 + fk of each 4-wide k step straight from memory into the MFMA operands:
// A row m0 + 16 i + fr and B row n0 + 16 j + fr, one 16-byte complex each.  Every output tile is
// accumulated over k in the same order whatever the blocking, so the blockings are tuner
// candidates that give the same bits.  The loads of the two next
// steps are in flight while a step multiplies; there is no LDS and no barrier.  Rows past M and N
// load a clamped valid row and are never stored; a ragged last k step is taken first, its
// missing k values read as zero.  The batched LDS-staged engine ran these products at a quarter
// of the FP64 MFMA rate with one wavefront per workgroup waiting on every slab.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "gemm.hpp"

namespace afq {
namespace zrt {

using gemm::f64x4;

template <int RB, int CB>
struct Frag {
  double ar[RB], ai[RB], br[CB], bi[CB];
};

template <class OP>
__device__ inline void ldz(const OP& o, long boff, int r, int k, double& re, double& im) {
  const long e = boff + (OP::kcontig ? (long)r * o.ld + k : (long)k * o.ld + r);
  const double2 v = *reinterpret_cast<const double2*>(o.p + 2 * e);
  re = v.x;
  im = OP::conj ? -v.y : v.y;
}

// the operands of k step k0 (lane k = k0 + fk); MASK: k values >= K read as zero
template <int RB, int CB, bool MASK, class OA, class OB>
__device__ inline void load_step(Frag<RB, CB>& f, const OA& A, long ab, const int (&ra)[RB],
                                 const OB& B, long bb, const int (&rb)[CB], int k0, int fk,
                                 int K) {
  const int k = MASK ? min(k0 + fk, K - 1) : k0 + fk;
#pragma unroll
  for (int i = 0; i < RB; ++i) ldz(A, ab, ra[i], k, f.ar[i], f.ai[i]);
#pragma unroll
  for (int j = 0; j < CB; ++j) ldz(B, bb, rb[j], k, f.br[j], f.bi[j]);
  if (MASK && k0 + fk >= K) {
#pragma unroll
    for (int i = 0; i < RB; ++i) f.ar[i] = f.ai[i] = 0.0;
#pragma unroll
    for (int j = 0; j < CB; ++j) f.br[j] = f.bi[j] = 0.0;
  }
}

// (the wavefront in its MFMA phase issues ahead of those still loading)
template <int RB, int CB>
__device__ inline void mma_step(const Frag<RB, CB>& f, f64x4 (&cr)[RB][CB], f64x4 (&ci)[RB][CB]) {
  using Acc = gemm::Acc<double>;
  __builtin_amdgcn_s_setprio(1);
#pragma unroll
  for (int j = 0; j < CB; ++j)
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      cr[i][j] = Acc::mfma(f.ar[i], f.br[j], cr[i][j]);
      cr[i][j] = Acc::mfma(-f.ai[i], f.bi[j], cr[i][j]);
      ci[i][j] = Acc::mfma(f.ar[i], f.bi[j], ci[i][j]);
      ci[i][j] = Acc::mfma(f.ai[i], f.br[j], ci[i][j]);
    }
  __builtin_amdgcn_s_setprio(0);
}

template <int RB, int CB, class OA, class OB, class EPI>
__global__ void __launch_bounds__(64)
    k_zgemm_rt(int M, int N, int K, OA A, OB B, int nrb, int ncb, EPI epi) {
  const int blk = blockIdx.x, b = blk / (nrb * ncb), q = blk % (nrb * ncb);
  const int m0 = (q / ncb) * 16 * RB, n0 = (q % ncb) * 16 * CB;
  const int lane = threadIdx.x, fr = lane & 15, fk = lane >> 4;
  const long ab = A.batch_offset(b), bb = B.batch_offset(b);
  int ra[RB], rb[CB];
#pragma unroll
  for (int i = 0; i < RB; ++i) ra[i] = min(m0 + 16 * i + fr, M - 1);
#pragma unroll
  for (int j = 0; j < CB; ++j) rb[j] = min(n0 + 16 * j + fr, N - 1);
  f64x4 cr[RB][CB], ci[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      cr[i][j] = f64x4{0, 0, 0, 0};
      ci[i][j] = f64x4{0, 0, 0, 0};
    }
  const int kt = K % 4, ns = K / 4;
  Frag<RB, CB> f0, f1, f2;
  if (kt) {
    load_step<RB, CB, true>(f0, A, ab, ra, B, bb, rb, 0, fk, kt);
    mma_step<RB, CB>(f0, cr, ci);
  }
  if (ns > 0) {
    // three buffers: steps s + 1 and s + 2 load while step s multiplies (prefetches past the
    // last step reload it rather than branch around the load)
    load_step<RB, CB, false>(f0, A, ab, ra, B, bb, rb, kt, fk, K);
    load_step<RB, CB, false>(f1, A, ab, ra, B, bb, rb, kt + 4 * min(1, ns - 1), fk, K);
    for (int st = 0; st < ns; st += 3) {
      load_step<RB, CB, false>(f2, A, ab, ra, B, bb, rb, kt + 4 * min(st + 2, ns - 1), fk, K);
      __builtin_amdgcn_sched_barrier(0);
      mma_step<RB, CB>(f0, cr, ci);
      __builtin_amdgcn_sched_barrier(0);
      if (st + 1 >= ns) break;
      load_step<RB, CB, false>(f0, A, ab, ra, B, bb, rb, kt + 4 * min(st + 3, ns - 1), fk, K);
      __builtin_amdgcn_sched_barrier(0);
      mma_step<RB, CB>(f1, cr, ci);
      __builtin_amdgcn_sched_barrier(0);
      if (st + 2 >= ns) break;
      load_step<RB, CB, false>(f1, A, ab, ra, B, bb, rb, kt + 4 * min(st + 4, ns - 1), fk, K);
      __builtin_amdgcn_sched_barrier(0);
      mma_step<RB, CB>(f2, cr, ci);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 16 * i + gemm::Acc<double>::row(fk, r), n = n0 + 16 * j + fr;
        if (m < M && n < N) epi(b, m, n, cr[i][j][r], ci[i][j][r]);
      }
}

// Whether the register-tiled path serves the product: 16-byte aligned complex operands and
// AFQ_ZRT unset or not "0" (the caller falls back to the LDS-staged engine otherwise).
template <class OP>
inline bool aligned16(const OP& o) {
  return reinterpret_cast<uintptr_t>(o.p) % 16 == 0;
}

inline bool enabled() {
  const char* e = std::getenv("AFQ_ZRT");
  return !(e && e[0] == '0');
}

template <class OA, class OB>
bool supported(int M, int N, int K, const OA& A, const OB& B) {
  static_assert(OA::is_cplx && OB::is_cplx && !OA::mapped && !OB::mapped && !OA::revk &&
                    !OB::revk,
                "complex operands addressed by row and k");
  static_assert(std::is_same<typename OA::in_t, double>::value &&
                    std::is_same<typename OB::in_t, double>::value,
                "FP64 operands");
  return enabled() && M >= 1 && N >= 1 && K >= 1 && aligned16(A) && aligned16(B);
}

// The blockings tried by the tuner: (RB, CB) = (rows up to 3 tiles, columns up to 3 tiles),
// (1, columns up to 3), (rows up to 3, 2), (1, 2), (1, 1).
constexpr int kCandidates = 5;
inline void candidate(int i, int M, int N, int& rb, int& cb) {
  const int rt = std::min(ceil_div(M, 16), 3), ct = std::min(ceil_div(N, 16), 3);
  switch (i) {
    case 0: rb = rt; cb = ct; break;
    case 1: rb = 1; cb = ct; break;
    case 2: rb = rt; cb = std::min(ct, 2); break;
    case 3: rb = 1; cb = std::min(ct, 2); break;
    default: rb = 1; cb = 1; break;
  }
}

template <class OA, class OB, class EPI>
void launch(hipStream_t s, int batch, int M, int N, int K, const OA& A, const OB& B, const EPI& e,
            int cand) {
  if (batch <= 0) return;
  int RB, CB;
  candidate(cand, M, N, RB, CB);
  const int nrb = ceil_div(M, 16 * RB), ncb = ceil_div(N, 16 * CB);
  AFQ_CHECK((long)batch * nrb * ncb < (1L << 31), kErrShape, "grid too large");
  const dim3 g((unsigned)(batch * nrb * ncb)), t(64);
#define AFQ_ZRT_LAUNCH(R, C)                                                                   \
  hipLaunchKernelGGL((k_zgemm_rt<R, C, OA, OB, EPI>), g, t, 0, s, M, N, K, A, B, nrb, ncb, e)
  switch (RB * 4 + CB) {
    case 5: AFQ_ZRT_LAUNCH(1, 1); break;
    case 6: AFQ_ZRT_LAUNCH(1, 2); break;
    case 7: AFQ_ZRT_LAUNCH(1, 3); break;
    case 9: AFQ_ZRT_LAUNCH(2, 1); break;
    case 10: AFQ_ZRT_LAUNCH(2, 2); break;
    case 11: AFQ_ZRT_LAUNCH(2, 3); break;
    case 13: AFQ_ZRT_LAUNCH(3, 1); break;
    case 14: AFQ_ZRT_LAUNCH(3, 2); break;
    default: AFQ_ZRT_LAUNCH(3, 3); break;
  }
#undef AFQ_ZRT_LAUNCH
  AFQ_LAUNCH_CHECK();
}

}  // namespace zrt
}  // namespace afq
