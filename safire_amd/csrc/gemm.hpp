// Batched complex GEMM on CDNA4 matrix cores (gfx950).
//
//   C[b](m, n) = sum_k A[b](m, k) * Bt[b](n, k)         (both operands addressed as [rows][K])
//
// Complex products are issued as real MFMA blocks on split re/im planes: 4 MFMAs per complex x
// complex block (Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br), 2 when one operand is real.  FP64
// uses v_mfma_f64_16x16x4f64, FP32 uses v_mfma_f32_16x16x4f32; both produce a 16x16 block from
// a 16x4 A fragment and a 4x16 B fragment with the lane layout
//   A: lane l holds A[l % 16][l / 16]      B: lane l holds B[l / 16][l % 16]
// and the accumulator register r (0..3) of lane l holding D[row][l % 16] with
//   FP32: row = 4 * (l / 16) + r           FP64: row = 4 * r + l / 16
// (measured on gfx950 with a one-hot lane probe of both builtins).
//
// A workgroup computes a BM x BN tile with WAVES_M x WAVES_N wavefronts; operands are staged
// through LDS in k-major planes ([BK][BM + pad]) so the fragment reads of a 16-lane group are
// consecutive words.  Global loads of the next k-slab are issued into registers before the
// current slab is consumed (one LDS buffer, register double-buffering).  Operand element type
// (TIn) may differ from the compute type T: FP64 walker matrices feed the FP32 path with the
// conversion done while staging.  Edges are zero-filled, so any M, N, K is valid.
#pragma once

#include "afq_common.hpp"

namespace afq {
namespace gemm {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Acc;
template <>
struct Acc<double> {
  using v4 = f64x4;
  static __device__ inline v4 mfma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of accumulator register r in lane group g = lane / 16
  static __device__ inline int row(int g, int r) { return 4 * r + g; }
};
template <>
struct Acc<float> {
  using v4 = f32x4;
  static __device__ inline v4 mfma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int g, int r) { return 4 * g + r; }
};

// Operand viewed as [rows][K]: element (r, k) of batch b lives at
//   b * bs + (KCONTIG ? r * ld + k : k * ld + r)        (units of scalar or complex elements)
// CONJ conjugates complex elements on load.
template <typename TIn, bool CPLX, bool KCONTIG, bool CONJ = false>
struct Opnd {
  using in_t = TIn;
  static constexpr bool is_cplx = CPLX;
  static constexpr bool kcontig = KCONTIG;
  static constexpr bool conj = CONJ;
  static constexpr bool mapped = false;
  static constexpr bool revk = false;
  const TIn* p;
  long ld;
  long bs;
  __device__ long row_offset(int) const { return 0; }
  __device__ long batch_offset(int b) const { return (long)b * bs; }
};

// An operand read with its k index reversed: product index k addresses element kr - k, so a
// product of two such operands accumulates from the last k to the first.  Summing the smallest
// terms first keeps the rounding of an FP32 accumulation small when the magnitudes fall with k
// (the Cholesky vectors of vHS come ordered by decreasing norm).  Both operands of a product
// must be reversed alike.
template <class OP>
struct RevK : OP {
  static constexpr bool revk = true;
  long kr;  // K - 1
  RevK(const OP& op, long K) : OP(op), kr(K - 1) {}
};

// Operand shared between batches: batch b reads block (b / div) % mod, so consecutive batches
// can share one block (div) or the blocks can repeat every mod batches.
template <typename TIn, bool CPLX, bool KCONTIG, bool CONJ = false>
struct OpndB : Opnd<TIn, CPLX, KCONTIG, CONJ> {
  int div, mod;
  __device__ long batch_offset(int b) const { return (long)((b / div) % mod) * this->bs; }
};

// Upper-triangle 8x8 blocks of an N x N pair space (i, k): block (bi, bk), bi <= bk, enumerated
// row by row, 64 pairs per block, pair t of a block = (8 bi + t / 8, 8 bk + t % 8).
__device__ inline void sym_block(int id, int nb, int& bi, int& bk) {
  float f = 0.5f * ((float)(2 * nb + 1) - sqrtf((float)((2 * nb + 1) * (2 * nb + 1) - 8 * id)));
  int b = (int)f;
  if (b < 0) b = 0;
  while (b > 0 && b * nb - b * (b - 1) / 2 > id) --b;
  while ((b + 1) * nb - (b + 1) * b / 2 <= id) ++b;
  bi = b;
  bk = b + (id - (b * nb - b * (b - 1) / 2));
}
__device__ inline bool sym_pair(int r, int nb, int N, int& i, int& k) {
  int bi, bk;
  sym_block(r >> 6, nb, bi, bk);
  i = 8 * bi + ((r & 63) >> 3);
  k = 8 * bk + (r & 7);
  return i < N && k < N;
}

// Real operand whose rows are the symmetric pairs (i, k) of a row-major [N*N][K] tensor (the
// Cholesky vectors L[i][k][m] with K = M contiguous).
template <typename TIn>
struct OpndSymPairs {
  using in_t = TIn;
  static constexpr bool is_cplx = false;
  static constexpr bool kcontig = true;
  static constexpr bool conj = false;
  static constexpr bool mapped = true;
  static constexpr bool revk = false;
  const TIn* p;
  long ld;  // K stride of one (i, k) row
  long bs;
  int N, nb;
  __device__ long row_offset(int r) const {
    int i, k;
    return sym_pair(r, nb, N, i, k) ? ((long)i * N + k) * ld : -1;
  }
  __device__ long batch_offset(int b) const { return (long)b * bs; }
};

// ---- epilogues: called once per output element with the accumulated (re, im) ----------------

// C = alpha * acc (+ C when accumulate)
template <typename T, typename TOut>
struct EpiStore {
  cplx<TOut>* c;
  long ldc, bs;
  T ar, ai;
  int accumulate;
  __device__ inline void operator()(int b, int m, int n, T re, T im) const {
    long idx = (long)b * bs + (long)m * ldc + n;
    T r = ar * re - ai * im, i = ar * im + ai * re;
    cplx<TOut> out;
    if (accumulate) {
      cplx<TOut> o = c[idx];
      out = {(TOut)(r + (T)o.re), (TOut)(i + (T)o.im)};
    } else {
      out = {(TOut)r, (TOut)i};
    }
    c[idx] = out;
  }
};

// symmetric pair output: column n is the pair (i, k) of OpndSymPairs; writes
// C[b][m][i][k] = C[b][m][k][i] = alpha * acc (C row m holds an N x N matrix)
template <typename T, typename TOut>
struct EpiStoreSym {
  cplx<TOut>* c;
  long ldc, bs;  // ldc = N * N
  T ar, ai;
  int N, nb;
  __device__ inline void operator()(int b, int m, int n, T re, T im) const {
    int i, k;
    if (!sym_pair(n, nb, N, i, k)) return;
    T r = ar * re - ai * im, q = ar * im + ai * re;
    cplx<TOut> out = {(TOut)r, (TOut)q};
    long base = (long)b * bs + (long)m * ldc;
    c[base + (long)i * N + k] = out;
    c[base + (long)k * N + i] = out;
  }
};

// real output: C = alpha * Re(acc)   (real x real products)
template <typename T>
struct EpiStoreReal {
  double* c;
  long ldc, bs;
  T alpha;
  __device__ inline void operator()(int b, int m, int n, T re, T) const {
    c[(long)b * bs + (long)m * ldc + n] = (double)(alpha * re);
  }
};

// C = alpha * acc and D = alpha * acc (two copies: the first Taylor operand and the running sum)
template <typename T>
struct EpiStore2 {
  cplx<T>* c;
  cplx<T>* d;
  long ldc, bs;
  __device__ inline void operator()(int b, int m, int n, T re, T im) const {
    long idx = (long)b * bs + (long)m * ldc + n;
    c[idx] = {re, im};
    d[idx] = {re, im};
  }
};

// Taylor term: T_k = f * acc written to C, and S += T_k  (apply_expM, propagate.hpp:60-66)
template <typename T>
struct EpiTaylor {
  cplx<T>* c;
  cplx<T>* s;
  long ldc, bs;
  T fr, fi;
  __device__ inline void operator()(int b, int m, int n, T re, T im) const {
    long idx = (long)b * bs + (long)m * ldc + n;
    T r = fr * re - fi * im, i = fr * im + fi * re;
    c[idx] = {r, i};
    cplx<T> o = s[idx];
    s[idx] = {o.re + r, o.im + i};
  }
};

// ---- operand staging -------------------------------------------------------------------------

template <typename T, class OP, int ROWS, int BK, int NT, int LDP>
struct Stager {
  static constexpr int TOT = ROWS * BK;
  static constexpr int E = (TOT + NT - 1) / NT;
  static constexpr bool EXACT = (TOT % NT) == 0;
  using TIn = typename OP::in_t;
  TIn re[E];
  TIn im[OP::is_cplx ? E : 1];
  long roff[OP::mapped ? E : 1];

  // row offsets of a mapped operand (rows are fixed per thread for the whole tile)
  __device__ inline void init_rows(const OP& op, int r0, int R, int tid) {
    if constexpr (OP::mapped) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        int gr = r0 + (tid + e * NT) / BK;
        roff[e] = (gr < R && (EXACT || tid + e * NT < TOT)) ? op.row_offset(gr) : -1;
      }
    }
  }

  __device__ inline void load(const OP& op, int b, int r0, int k0, int R, int K, int tid) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      int idx = tid + e * NT;
      int r, k;
      if (OP::kcontig) {
        r = idx / BK;
        k = idx % BK;
      } else {
        r = idx % ROWS;
        k = idx / ROWS;
      }
      int gr = r0 + r, gk = k0 + k;
      bool in = (gr < R) && (gk < K) && (EXACT || idx < TOT);
      long ka = gk;
      if constexpr (OP::revk) ka = op.kr - gk;
      long a;
      if constexpr (OP::mapped) {
        in = (roff[e] >= 0) && (gk < K) && (EXACT || idx < TOT);
        a = op.batch_offset(b) + roff[e] + ka;
      } else {
        a = op.batch_offset(b) + (OP::kcontig ? (long)gr * op.ld + ka : ka * op.ld + gr);
      }
      if (OP::is_cplx) {
        cplx<TIn> v = in ? reinterpret_cast<const cplx<TIn>*>(op.p)[a] : cplx<TIn>{0, 0};
        re[e] = v.re;
        im[e] = OP::conj ? -v.im : v.im;
      } else {
        re[e] = in ? op.p[a] : TIn(0);
      }
    }
  }
  __device__ inline void store(T* sre, T* sim, int tid) const {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      int idx = tid + e * NT;
      int r, k;
      if (OP::kcontig) {
        r = idx / BK;
        k = idx % BK;
      } else {
        r = idx % ROWS;
        k = idx / ROWS;
      }
      if (!EXACT && idx >= TOT) continue;
      sre[k * LDP + r] = (T)re[e];
      if (OP::is_cplx) sim[k * LDP + r] = (T)im[e];
    }
  }
};

// Partial products of a split-K launch: slice ks of batch b goes to
// ws[(b * ksplit + ks) * M * N + m * N + n] (the kernel hands the epilogue b * ksplit + ks).
template <typename T>
struct EpiPartial {
  cplx<T>* ws;
  long mn;   // M * N
  int ldn;   // N
  __device__ inline void operator()(int bk, int m, int n, T re, T im) const {
    ws[(long)bk * mn + (long)m * ldn + n] = {re, im};
  }
};

// PP: two LDS buffers used in turn (the next slab is stored into the other buffer after the
// current one is multiplied), one barrier per slab instead of two.  The MFMA sequence, and so
// the result, is the same either way.
template <typename T, int BM, int BN, int BK, int WAVES_M, int WAVES_N, class OA, class OB, class EPI,
          bool PP = false>
__global__ void __launch_bounds__(WAVES_M * WAVES_N * 64)
    kgemm(int M, int N, int K, int tiles_m, int tiles_n, int ksplit, int kchunk, OA A, OB B,
          EPI epi) {
  constexpr int NT = WAVES_M * WAVES_N * 64;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int RB = WM / 16, CB = WN / 16;
  static_assert(WM % 16 == 0 && WN % 16 == 0 && BK % 4 == 0, "bad tile");
  constexpr int LDA = BM + 4, LDB = BN + 4;
  constexpr bool CA = OA::is_cplx, CBX = OB::is_cplx;
  constexpr int NBUF = PP ? 2 : 1;
  using V4 = typename Acc<T>::v4;

  __shared__ T sAr[NBUF * BK * LDA];
  __shared__ T sAi[CA ? NBUF * BK * LDA : 1];
  __shared__ T sBr[NBUF * BK * LDB];
  __shared__ T sBi[CBX ? NBUF * BK * LDB : 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WAVES_N) * WM, wn0 = (wave % WAVES_N) * WN;
  // Workgroups are dispatched round-robin over the 8 XCDs; renumber them so that each XCD takes a
  // contiguous run of tiles, row tiles fastest: the workgroups resident on one XCD at a time
  // then share a few column tiles of B (the Cholesky operand of vbias/vHS) through its L2.
  const long total = gridDim.x, xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const long q = xcd * (total / 8) + min(xcd, total % 8) + slot;
  const int tm = (int)(q % tiles_m);
  const int tn = (int)((q / tiles_m) % tiles_n);
  const long rest = q / ((long)tiles_n * tiles_m);
  const int ks = (int)(rest % ksplit);
  const int b = (int)(rest / ksplit);
  const int m0 = tm * BM, n0 = tn * BN;
  // k range of this block: [kb, ke)
  const int kb = ks * kchunk;
  const int ke = min(K, kb + kchunk);

  V4 accr[RB][CB], acci[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      accr[i][j] = V4{0, 0, 0, 0};
      acci[i][j] = V4{0, 0, 0, 0};
    }

  Stager<T, OA, BM, BK, NT, LDA> ga;
  Stager<T, OB, BN, BK, NT, LDB> gb;
  ga.init_rows(A, m0, M, tid);
  gb.init_rows(B, n0, N, tid);
  ga.load(A, b, m0, kb, M, ke, tid);
  gb.load(B, b, n0, kb, N, ke, tid);

  const int fr = lane & 15, fk = lane >> 4;
  // one BK slab from LDS buffer (pa, pai, pb, pbi)
  auto slab = [&](const T* pa, const T* pai, const T* pb, const T* pbi) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kl = kk + fk;
      T ar[RB], ai[RB], br[CB], bi[CB];
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        ar[i] = pa[kl * LDA + wm0 + i * 16 + fr];
        if (CA) ai[i] = pai[kl * LDA + wm0 + i * 16 + fr];
      }
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        br[j] = pb[kl * LDB + wn0 + j * 16 + fr];
        if (CBX) bi[j] = pbi[kl * LDB + wn0 + j * 16 + fr];
      }
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
          accr[i][j] = Acc<T>::mfma(ar[i], br[j], accr[i][j]);
          if (CA && CBX) {
            accr[i][j] = Acc<T>::mfma(-ai[i], bi[j], accr[i][j]);
            acci[i][j] = Acc<T>::mfma(ar[i], bi[j], acci[i][j]);
            acci[i][j] = Acc<T>::mfma(ai[i], br[j], acci[i][j]);
          } else if (CA) {
            acci[i][j] = Acc<T>::mfma(ai[i], br[j], acci[i][j]);
          } else if (CBX) {
            acci[i][j] = Acc<T>::mfma(ar[i], bi[j], acci[i][j]);
          }
        }
    }
  };
  if constexpr (PP) {
    // buffer (it & 1) holds slab it; slab it + 1 goes into the other buffer, whose last readers
    // (slab it - 1) all passed the barrier that ended the previous iteration
    ga.store(sAr, sAi, tid);
    gb.store(sBr, sBi, tid);
    __syncthreads();
    int cur = 0;
    for (int k0 = kb; k0 < ke; k0 += BK) {
      const bool more = k0 + BK < ke;
      if (more) {
        ga.load(A, b, m0, k0 + BK, M, ke, tid);
        gb.load(B, b, n0, k0 + BK, N, ke, tid);
      }
      slab(sAr + cur * BK * LDA, sAi + (CA ? cur * BK * LDA : 0), sBr + cur * BK * LDB,
           sBi + (CBX ? cur * BK * LDB : 0));
      if (more) {
        const int nxt = cur ^ 1;
        ga.store(sAr + nxt * BK * LDA, sAi + (CA ? nxt * BK * LDA : 0), tid);
        gb.store(sBr + nxt * BK * LDB, sBi + (CBX ? nxt * BK * LDB : 0), tid);
      }
      __syncthreads();
      cur ^= 1;
    }
  } else {
    for (int k0 = kb; k0 < ke; k0 += BK) {
      __syncthreads();
      ga.store(sAr, sAi, tid);
      gb.store(sBr, sBi, tid);
      __syncthreads();
      if (k0 + BK < ke) {
        ga.load(A, b, m0, k0 + BK, M, ke, tid);
        gb.load(B, b, n0, k0 + BK, N, ke, tid);
      }
      slab(sAr, sAi, sBr, sBi);
    }
  }

#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int m = m0 + wm0 + i * 16 + Acc<T>::row(fk, r);
        int n = n0 + wn0 + j * 16 + fr;
        if (m < M && n < N) epi(b * ksplit + ks, m, n, accr[i][j][r], acci[i][j][r]);
      }
}

// Sum of the split-K partial slices in slice order, then the final epilogue.
template <typename T, class EPI>
__global__ void __launch_bounds__(256)
    kreduce_splitk(int batch, int M, int N, int ksplit, const cplx<T>* __restrict__ ws, EPI epi) {
  const long mn = (long)M * N;
  long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)batch * mn) return;
  const int b = (int)(t / mn);
  const long e = t % mn;
  const cplx<T>* p = ws + (long)b * ksplit * mn + e;
  double sr = 0, si = 0;  // FP32 slices are summed in FP64
  for (int k = 0; k < ksplit; ++k) {
    cplx<T> v = p[(long)k * mn];
    sr += v.re;
    si += v.im;
  }
  epi(b, (int)(e / N), (int)(e % N), (T)sr, (T)si);
}

// Host launcher.  batch = number of independent products (grid = batch * tiles).
template <typename T, int BM, int BN, int BK, int WAVES_M, int WAVES_N, bool PP = false, class OA,
          class OB, class EPI>
inline void launch(hipStream_t s, int batch, int M, int N, int K, const OA& A, const OB& B,
                   const EPI& epi) {
  if (batch <= 0 || M <= 0 || N <= 0) return;
  int tm = ceil_div(M, BM), tn = ceil_div(N, BN);
  long blocks = (long)batch * tm * tn;
  AFQ_CHECK(blocks < (1L << 31), kErrShape, "grid too large");
  hipLaunchKernelGGL((kgemm<T, BM, BN, BK, WAVES_M, WAVES_N, OA, OB, EPI, PP>),
                     dim3((unsigned)blocks), dim3(WAVES_M * WAVES_N * 64), 0, s, M, N, K, tm, tn,
                     1, K > 0 ? K : 1, A, B, epi);
  AFQ_LAUNCH_CHECK();
}

// Split-K launch: K cut into ksplit chunks of kchunk (a multiple of BK) computed by separate
// workgroups into ws (ksplit * batch * M * N complex), then summed in chunk order and passed to
// the epilogue. The summation order depends only on kchunk, not on the tile.
template <typename T, int BM, int BN, int BK, int WAVES_M, int WAVES_N, bool PP = false, class OA,
          class OB, class EPI>
inline void launch_splitk(hipStream_t s, int batch, int M, int N, int K, int ksplit, int kchunk,
                          cplx<T>* ws, const OA& A, const OB& B, const EPI& epi) {
  if (batch <= 0 || M <= 0 || N <= 0) return;
  if (ksplit <= 1) {
    launch<T, BM, BN, BK, WAVES_M, WAVES_N, PP>(s, batch, M, N, K, A, B, epi);
    return;
  }
  AFQ_CHECK(kchunk % BK == 0 && (long)kchunk * ksplit >= K, kErrArg, "bad split-K chunking");
  int tm = ceil_div(M, BM), tn = ceil_div(N, BN);
  long blocks = (long)batch * tm * tn * ksplit;
  AFQ_CHECK(blocks < (1L << 31), kErrShape, "grid too large");
  EpiPartial<T> ep{ws, (long)M * N, N};
  hipLaunchKernelGGL((kgemm<T, BM, BN, BK, WAVES_M, WAVES_N, OA, OB, EpiPartial<T>, PP>),
                     dim3((unsigned)blocks), dim3(WAVES_M * WAVES_N * 64), 0, s, M, N, K, tm, tn,
                     ksplit, kchunk, A, B, ep);
  AFQ_LAUNCH_CHECK();
  long tot = (long)batch * M * N;
  hipLaunchKernelGGL((kreduce_splitk<T, EPI>), dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, s,
                     batch, M, N, ksplit, ws, epi);
  AFQ_LAUNCH_CHECK();
}

}  // namespace gemm
}  // namespace afq
