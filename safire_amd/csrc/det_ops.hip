// Slater-determinant operations on the device and the scratch-buffer context.
//
//   log_overlap           det_ops::Log_Overlap / log_overlap_impl (density_matrix.hpp:113-156, :431-447)
//   mixed_density_matrix  det_ops::MixedDensityMatrix (density_matrix.hpp:535-601)
//   inverse_logdet        det_ops::detail::inverse_logdet (density_matrix.hpp:82-111)
//   propagate             det_ops::detail::propagate_impl + apply_expM (propagate.hpp:41-70, :241-298)
//   orthogonalize         det_ops::orthogonalize (orthogonalize.hpp:131-169)
//
// Every product is a batched MFMA GEMM (gemm.hpp) over walkers; the n x n LU/Cholesky work runs
// one wavefront per walker (dense_small.hip).  The orthogonalization uses Cholesky-QR2 instead
// of Householder: with R's diagonal positive and real (the reference rescales the Householder
// Q by sign(Re R_ii)), Q is unique, so both produce the same Q to rounding while the
// GEMM-shaped Gram/solve steps keep the work on the matrix cores.
#include <cstdlib>

#include "gemm.hpp"
#include "ops.hpp"
#include "tuner.hpp"

namespace afq {

// ---------------------------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------------------------
Context::Context(int device, hipStream_t stream) : device_(device), stream_(stream), own_stream_(false) {
  AFQ_HIP(hipSetDevice(device));
  if (stream_ == nullptr) {
    AFQ_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = true;
  }
}

Context::~Context() {
  (void)hipSetDevice(device_);
  (void)hipStreamSynchronize(stream_);
  for (auto& kv : bufs_)
    if (kv.second.p) (void)hipFree(kv.second.p);
  if (own_stream_) (void)hipStreamDestroy(stream_);
}

bool Context::capturing() const {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream_, &st) != hipSuccess) return false;
  return st == hipStreamCaptureStatusActive;
}

void* Context::scratch(const std::string& name, size_t bytes) {
  auto it = bufs_.find(name);
  if (it != bufs_.end() && it->second.bytes >= bytes && it->second.p) return it->second.p;
  AFQ_CHECK(!no_alloc_, kErrState,
            "scratch buffer '" + name + "' requested by an allocation-free call");
  Buf& b = bufs_[name];
  AFQ_CHECK(!capturing(), kErrState,
            "scratch buffer '" + name + "' would grow during graph capture");
  if (b.p) {
    AFQ_HIP(hipStreamSynchronize(stream_));
    AFQ_HIP(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
    ++gen_;
  }
  size_t sz = bytes < 256 ? 256 : bytes;
  if (hipMalloc(&b.p, sz) != hipSuccess) {
    b.p = nullptr;
    throw Error(kErrAlloc, "scratch allocation of " + std::to_string(sz) + " bytes failed");
  }
  b.bytes = sz;
  return b.p;
}

size_t Context::scratch_bytes() const {
  size_t t = 0;
  for (auto& kv : bufs_) t += kv.second.bytes;
  return t;
}

namespace det_ops {

using gemm::EpiStore;
using gemm::EpiStore2;
using gemm::EpiTaylor;
using gemm::Opnd;

void log_overlap(Context& c, const z64* PsiT, int n, int N, const z64* Phi, long phi_bs, int W,
                 z64* ovlp, double scale, z64 add, const LuKeep* keep) {
  if (n == 0 || W == 0) return;
  hipStream_t s = c.stream();
  z64* T = keep ? keep->T : scratch_as<z64>(c, "lo_T", (size_t)W * n * n);
  Opnd<double, true, true> A{&PsiT->re, N, 0};
  Opnd<double, true, false> B{&Phi->re, n, phi_bs};
  const EpiStore<double, double> e{T, n, (long)n * n, 1, 0, 0};
  zgemm_small(c, "lo_T", W, n, n, N, A, B, e);
  if (keep)
    lu_inverse_logdet(s, T, (long)n * n, W, n, keep->Tinv, (long)n * n, ovlp, scale, add, true,
                      keep, false);
  else
    lu_inverse_logdet(s, T, (long)n * n, W, n, nullptr, 0, ovlp, scale, add, false);
}

void mixed_density_matrix(Context& c, const z64* PsiT, int n, int N, const z64* Phi, long phi_bs,
                          int W, z64* G, long g_bs, z64* ovlp, bool compact, double scale,
                          c32* G32, z64* ws_T, z64* ws_Gc, const LuKeep* keep) {
  AFQ_CHECK(!G32 || compact, kErrArg, "the complex64 copy is for the compact density matrix");
  AFQ_CHECK(ws_T ? (compact || ws_Gc) : !ws_Gc, kErrArg,
            "workspace: T, and Gc for the full density matrix, or neither");
  if (W == 0) return;
  hipStream_t s = c.stream();
  if (n == 0) {
    AFQ_CHECK(compact || g_bs == (long)N * N, kErrShape, "empty density matrix in a strided batch");
    if (!compact) fill_z64(s, (long)W * g_bs, {0, 0}, G);
    return;
  }
  z64* T = ws_T ? ws_T : scratch_as<z64>(c, "mdm_T", (size_t)W * n * n);
  {
    Opnd<double, true, true> A{&PsiT->re, N, 0};
    Opnd<double, true, false> B{&Phi->re, n, phi_bs};
    const EpiStore<double, double> e{T, n, (long)n * n, 1, 0, 0};
    zgemm_small(c, "mdm_T", W, n, n, N, A, B, e);
  }
  z64* Ti = keep ? keep->Tinv : T;
  lu_inverse_logdet(s, T, (long)n * n, W, n, Ti, (long)n * n, ovlp, scale, {0, 0}, true, keep,
                    true);
  // compact G[a][k] = sum_b Tinv[b][a] Phi[k][b]    (product<'T','T'>(TNN, B, C))
  z64* Gc = compact ? G : ws_Gc ? ws_Gc : scratch_as<z64>(c, "mdm_Gc", (size_t)W * n * N);
  long gc_bs = compact ? g_bs : (long)n * N;
  {
    Opnd<double, true, false> A{&Ti->re, n, (long)n * n};
    Opnd<double, true, true> B{&Phi->re, n, phi_bs};
    if (G32) {
      const EpiStore<double, float> e{G32, N, g_bs, 1, 0, 0};
      zgemm_small(c, "mdm_G", W, n, N, n, A, B, e);
    } else {
      const EpiStore<double, double> e{Gc, N, gc_bs, 1, 0, 0};
      zgemm_small(c, "mdm_G", W, n, N, n, A, B, e);
    }
  }
  if (!compact) {
    // G[i][k] = sum_a PsiT[a][i] Gc[a][k]           (product<'T'>(A, TNM, C))
    Opnd<double, true, false> A{&PsiT->re, N, 0};
    Opnd<double, true, false> B{&Gc->re, N, (long)n * N};
    zgemm_small(c, "mdm_full", W, N, N, n, A, B, EpiStore<double, double>{G, N, g_bs, 1, 0, 0});
  }
}

void mixed_density_matrix_left(Context& c, const z64* L, long l_bs, int n, int N, const z64* Phi,
                               long phi_bs, int W, z64* G, long g_bs, z64* ovlp) {
  if (W == 0) return;
  AFQ_CHECK(n > 0 && n <= N && g_bs >= (long)N * N, kErrShape, "left density matrix shape");
  hipStream_t s = c.stream();
  z64* T = scratch_as<z64>(c, "mdl_T", (size_t)W * n * n);
  {  // T[a][b] = sum_k conj(L[k][a]) Phi[k][b]
    Opnd<double, true, false, true> A{&L->re, n, l_bs};
    Opnd<double, true, false> B{&Phi->re, n, phi_bs};
    zgemm_small(c, "mdl_T", W, n, n, N, A, B, EpiStore<double, double>{T, n, (long)n * n, 1, 0, 0});
  }
  lu_inverse_logdet(s, T, (long)n * n, W, n, T, (long)n * n, ovlp, 1.0, {0, 0}, true);
  z64* Gc = scratch_as<z64>(c, "mdl_Gc", (size_t)W * n * N);
  {  // Gc[a][k] = sum_b Tinv[b][a] Phi[k][b]
    Opnd<double, true, false> A{&T->re, n, (long)n * n};
    Opnd<double, true, true> B{&Phi->re, n, phi_bs};
    zgemm_small(c, "mdm_G", W, n, N, n, A, B, EpiStore<double, double>{Gc, N, (long)n * N, 1, 0, 0});
  }
  {  // G[i][k] = sum_a conj(L[i][a]) Gc[a][k]
    Opnd<double, true, true, true> A{&L->re, n, l_bs};
    Opnd<double, true, false> B{&Gc->re, N, (long)n * N};
    zgemm_small(c, "mdl_full", W, N, N, n, A, B, EpiStore<double, double>{G, N, g_bs, 1, 0, 0});
  }
}

void mixed_density_matrix_batched(Context& c, const z64* L, long l_bs, int ldiv, int n, int N,
                                  const z64* Phi, long phi_bs, int W, int B, z64* G, long g_bs,
                                  z64* ovlp) {
  using gemm::OpndB;
  if (B == 0) return;
  AFQ_CHECK(n > 0 && n <= N && W > 0 && ldiv > 0 && (!G || g_bs >= (long)N * N), kErrShape,
            "batched density matrix shape");
  hipStream_t s = c.stream();
  const int big = 1 << 30;
  z64* T = scratch_as<z64>(c, "mdb_T", (size_t)B * n * n);
  {  // T[a][b] = sum_k conj(L[k][a]) Phi[k][b]
    OpndB<double, true, false, true> A{{&L->re, n, l_bs}, ldiv, big};
    OpndB<double, true, false> Bp{{&Phi->re, n, phi_bs}, 1, W};
    gemm_small<double>(c, "mdb_T", B, n, n, N, A, Bp, EpiStore<double, double>{T, n, (long)n * n, 1, 0, 0});
  }
  lu_inverse_logdet(s, T, (long)n * n, B, n, G ? T : nullptr, (long)n * n, ovlp, 1.0, {0, 0},
                    G != nullptr);
  if (!G) return;
  z64* Gc = scratch_as<z64>(c, "mdb_Gc", (size_t)B * n * N);
  {  // Gc[a][k] = sum_b Tinv[b][a] Phi[k][b]
    Opnd<double, true, false> A{&T->re, n, (long)n * n};
    OpndB<double, true, true> Bp{{&Phi->re, n, phi_bs}, 1, W};
    gemm_small<double>(c, "mdb_G", B, n, N, n, A, Bp, EpiStore<double, double>{Gc, N, (long)n * N, 1, 0, 0});
  }
  {  // G[i][k] = sum_a conj(L[i][a]) Gc[a][k]
    OpndB<double, true, true, true> A{{&L->re, n, l_bs}, ldiv, big};
    Opnd<double, true, false> Bp{&Gc->re, N, (long)n * N};
    gemm_small<double>(c, "mdb_full", B, N, N, n, A, Bp, EpiStore<double, double>{G, N, g_bs, 1, 0, 0});
  }
}

void log_overlap_left(Context& c, const z64* L, long l_bs, int n, int N, const z64* Phi,
                      long phi_bs, int W, z64* ovlp) {
  if (n == 0 || W == 0) return;
  AFQ_CHECK(n <= N, kErrShape, "left overlap shape");
  z64* T = scratch_as<z64>(c, "lol_T", (size_t)W * n * n);
  Opnd<double, true, false, true> A{&L->re, n, l_bs};
  Opnd<double, true, false> B{&Phi->re, n, phi_bs};
  zgemm_small(c, "mdl_T", W, n, n, N, A, B, EpiStore<double, double>{T, n, (long)n * n, 1, 0, 0});
  lu_inverse_logdet(c.stream(), T, (long)n * n, W, n, nullptr, 0, ovlp, 1.0, {0, 0}, false);
}

void inverse_logdet(Context& c, const z64* A, int nb, int n, z64* ovlp, z64* Ainv) {
  lu_inverse_logdet(c.stream(), A, (long)n * n, nb, n, Ainv, (long)n * n, ovlp, 1.0, {0, 0}, true);
}

namespace {

// ---- fused propagation ------------------------------------------------------------------------
//
// Phi <- P1 expM(V) P1 Phi in one launch (propagate_impl + apply_expM for the 'N' operator).
// The Taylor recurrence T_k = (i/k) V T_{k-1} and both P1 products act on each column of the
// walker matrix separately, so a workgroup owns one walker and one 16-column tile of it:
//   * the current term lives in LDS as the B operand, stored [column][row] so the four k values
//     a lane feeds to four consecutive MFMAs are one 16-byte word;
//   * wavefront w owns the 16-row stripes w, w + NW, ... of the output: it streams the matching
//     rows of V (or P1) from memory straight into MFMA A fragments (no LDS staging, no barrier
//     inside a product), keeps its accumulators and its part of the running sum S in registers;
//   * MFMA step j of the 16-wide k group g multiplies k = 16 g + 4 (lane / 16) + j, the same
//     grouping in A and B.
// Workgroups are numbered so that each XCD (8 of them, dispatched round-robin) processes a
// contiguous run of (walker, tile) pairs: the tiles of one walker share its V in that XCD's L2.

template <typename T, bool CPLX>
struct AFrag {
  T re[4];
  T im[CPLX ? 4 : 1];
};

// A[row][k .. k + 3] (complex elements interleaved).  FULL (k + 3 < K): unconditional loads,
// rows >= rows read row rows - 1 (their products land in output rows that write_b zeroes and
// the final store skips), so the compiler counts the loads in flight and the next groups' loads
// overlap this group's MFMAs.  Otherwise rows >= rows or k >= K read as zero.  VEC: the rows of
// A are 16-byte aligned, so fragments use vector loads.
template <typename T, bool CPLX, bool VEC, bool FULL>
__device__ inline void load_afrag(AFrag<T, CPLX>& f, const T* __restrict__ A, long lda, int row,
                                  int rows, int k, int K) {
  if (FULL) row = min(row, rows - 1);
  if (VEC && (FULL || (row < rows && k + 3 < K))) {
    if constexpr (CPLX && sizeof(T) == 4) {
      const float4* p = reinterpret_cast<const float4*>(A + 2 * ((long)row * lda + k));
      const float4 x = p[0], y = p[1];
      f.re[0] = x.x; f.im[0] = x.y; f.re[1] = x.z; f.im[1] = x.w;
      f.re[2] = y.x; f.im[2] = y.y; f.re[3] = y.z; f.im[3] = y.w;
    } else if constexpr (CPLX) {
      const double2* p = reinterpret_cast<const double2*>(A + 2 * ((long)row * lda + k));
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double2 v = p[j];
        f.re[j] = v.x;
        f.im[j] = v.y;
      }
    } else if constexpr (sizeof(T) == 4) {
      const float4 x = *reinterpret_cast<const float4*>(A + (long)row * lda + k);
      f.re[0] = x.x; f.re[1] = x.y; f.re[2] = x.z; f.re[3] = x.w;
    } else {
      const double2* p = reinterpret_cast<const double2*>(A + (long)row * lda + k);
      const double2 x = p[0], y = p[1];
      f.re[0] = x.x; f.re[1] = x.y; f.re[2] = y.x; f.re[3] = y.y;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool in = FULL || (row < rows && k + j < K);
      const long e = (long)row * lda + k + j;
      if constexpr (CPLX) {
        f.re[j] = in ? A[2 * e] : (T)0;
        f.im[j] = in ? A[2 * e + 1] : (T)0;
      } else {
        f.re[j] = in ? A[e] : (T)0;
      }
    }
  }
}

template <typename T>
__device__ inline void read_bfrag(T (&br)[4], T (&bi)[4], const T* Br, const T* Bi, int off) {
  if constexpr (sizeof(T) == 4) {
    const float4 x = *reinterpret_cast<const float4*>(Br + off);
    const float4 y = *reinterpret_cast<const float4*>(Bi + off);
    br[0] = x.x; br[1] = x.y; br[2] = x.z; br[3] = x.w;
    bi[0] = y.x; bi[1] = y.y; bi[2] = y.z; bi[3] = y.w;
  } else {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const double2 x = *reinterpret_cast<const double2*>(Br + off + 2 * h);
      const double2 y = *reinterpret_cast<const double2*>(Bi + off + 2 * h);
      br[2 * h] = x.x; br[2 * h + 1] = x.y;
      bi[2 * h] = y.x; bi[2 * h + 1] = y.y;
    }
  }
}

// One 16-wide k group of acc[s][t] += op(A)[stripe s rows][k group] * B[k group][column tile t],
// op(A) = conj(A) when CONJ.  M3 (complex A): the three real products of Gauss's method,
// p1 = Ar Br, p2 = Ai Bi, p3 = (Ar +- Ai)(Br + Bi), in (ar, ai, a3); stripes_times_b combines
// them.  Otherwise four products straight into the real and imaginary parts (ar, ai).
template <typename T, bool CPLX, int SPW, int CB, bool CONJ = false, bool M3 = false>
__device__ inline void group_mfma(const AFrag<T, CPLX> (&f)[SPW], int g, int NS, int NW, int wave,
                                  int fr, int fk, const T* Br, const T* Bi, int LDK,
                                  typename gemm::Acc<T>::v4 (&ar)[SPW][CB],
                                  typename gemm::Acc<T>::v4 (&ai)[SPW][CB],
                                  typename gemm::Acc<T>::v4 (&a3)[SPW][CB]) {
  using gemm::Acc;
  // the wavefront in its MFMA phase issues ahead of those still loading
  __builtin_amdgcn_s_setprio(1);
#pragma unroll
  for (int t = 0; t < CB; ++t) {
    T br[4], bi[4];
    read_bfrag(br, bi, Br, Bi, (16 * t + fr) * LDK + 16 * g + 4 * fk);
#pragma unroll
    for (int s = 0; s < SPW; ++s) {
      if (wave + s * NW >= NS) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if constexpr (CPLX && M3) {
          const T as = CONJ ? f[s].re[j] - f[s].im[j] : f[s].re[j] + f[s].im[j];
          ar[s][t] = Acc<T>::mfma(f[s].re[j], br[j], ar[s][t]);
          ai[s][t] = Acc<T>::mfma(f[s].im[j], bi[j], ai[s][t]);
          a3[s][t] = Acc<T>::mfma(as, br[j] + bi[j], a3[s][t]);
          continue;
        }
        ar[s][t] = Acc<T>::mfma(f[s].re[j], br[j], ar[s][t]);
        ai[s][t] = Acc<T>::mfma(f[s].re[j], bi[j], ai[s][t]);
        if constexpr (CPLX && !CONJ) {
          ar[s][t] = Acc<T>::mfma(-f[s].im[j], bi[j], ar[s][t]);
          ai[s][t] = Acc<T>::mfma(f[s].im[j], br[j], ai[s][t]);
        } else if constexpr (CPLX) {
          ar[s][t] = Acc<T>::mfma(f[s].im[j], bi[j], ar[s][t]);
          ai[s][t] = Acc<T>::mfma(-f[s].im[j], br[j], ai[s][t]);
        }
      }
    }
  }
  __builtin_amdgcn_s_setprio(0);
}

// a stripe past the last one loads the last one (its products are never used), so every load
// of a group is issued unconditionally
template <typename T, bool CPLX, int SPW, bool VEC, bool FULL>
__device__ inline void load_group(AFrag<T, CPLX> (&f)[SPW], const T* __restrict__ A, long lda,
                                  int N, int NS, int NW, int wave, int fr, int fk, int g) {
#pragma unroll
  for (int s = 0; s < SPW; ++s)
    load_afrag<T, CPLX, VEC, FULL>(f[s], A, lda, min(wave + s * NW, NS - 1) * 16 + fr, N,
                                   16 * g + 4 * fk, N);
}

// acc = A[stripe rows][0:N] * B[0:N][tiles] for the wave's stripes (complex B in LDS). The A
// fragments of the next two k groups are in flight while a group is multiplied.  M3: complex
// products by Gauss's three-multiplication method (three MFMAs per complex block instead of four).
// The full k groups [gb, ge) (ge < 0: all of them), then the ragged last group if `ragged`: a
// part of the product for a split of k over several wavefronts.
template <typename T, bool CPLX, int SPW, int CB, bool VEC, bool CONJ = false, bool M3 = false>
__device__ inline void stripes_times_b(const T* __restrict__ A, long lda, int N, int NS,
                                       int NW, int wave, int fr, int fk, const T* Br,
                                       const T* Bi, int LDK,
                                       typename gemm::Acc<T>::v4 (&ar)[SPW][CB],
                                       typename gemm::Acc<T>::v4 (&ai)[SPW][CB], int gb = 0,
                                       int ge = -1, bool ragged = true) {
  using V4 = typename gemm::Acc<T>::v4;
  constexpr bool G3 = CPLX && M3;
  V4 a3[SPW][CB];
#pragma unroll
  for (int s = 0; s < SPW; ++s)
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      ar[s][t] = V4{0, 0, 0, 0};
      ai[s][t] = V4{0, 0, 0, 0};
      a3[s][t] = V4{0, 0, 0, 0};
    }
  AFrag<T, CPLX> f0[SPW], f1[SPW], f2[SPW];
  // the full k groups: two groups' loads in flight while one is multiplied (prefetches past the
  // last full group reload it rather than branching around the load); the scheduling barriers
  // keep each group's loads ahead of the previous group's MFMAs
  const int NGF = N / 16;
  const int g1 = ge < 0 ? NGF : ge;
  if (g1 > gb) {
    load_group<T, CPLX, SPW, VEC, true>(f0, A, lda, N, NS, NW, wave, fr, fk, gb);
    load_group<T, CPLX, SPW, VEC, true>(f1, A, lda, N, NS, NW, wave, fr, fk, min(gb + 1, g1 - 1));
    for (int g = gb; g < g1; g += 3) {
      load_group<T, CPLX, SPW, VEC, true>(f2, A, lda, N, NS, NW, wave, fr, fk, min(g + 2, g1 - 1));
      __builtin_amdgcn_sched_barrier(0);
      group_mfma<T, CPLX, SPW, CB, CONJ, G3>(f0, g, NS, NW, wave, fr, fk, Br, Bi, LDK, ar, ai, a3);
      __builtin_amdgcn_sched_barrier(0);
      if (g + 1 >= g1) break;
      load_group<T, CPLX, SPW, VEC, true>(f0, A, lda, N, NS, NW, wave, fr, fk, min(g + 3, g1 - 1));
      __builtin_amdgcn_sched_barrier(0);
      group_mfma<T, CPLX, SPW, CB, CONJ, G3>(f1, g + 1, NS, NW, wave, fr, fk, Br, Bi, LDK, ar, ai, a3);
      __builtin_amdgcn_sched_barrier(0);
      if (g + 2 >= g1) break;
      load_group<T, CPLX, SPW, VEC, true>(f1, A, lda, N, NS, NW, wave, fr, fk, min(g + 4, g1 - 1));
      __builtin_amdgcn_sched_barrier(0);
      group_mfma<T, CPLX, SPW, CB, CONJ, G3>(f2, g + 2, NS, NW, wave, fr, fk, Br, Bi, LDK, ar, ai, a3);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  if (ragged && NGF < NS) {  // the ragged last k group, masked
    load_group<T, CPLX, SPW, VEC, false>(f0, A, lda, N, NS, NW, wave, fr, fk, NGF);
    group_mfma<T, CPLX, SPW, CB, CONJ, G3>(f0, NGF, NS, NW, wave, fr, fk, Br, Bi, LDK, ar, ai, a3);
  }
  if constexpr (G3) {
    // p1 = Ar Br, p2 = Ai Bi, p3 = (Ar +- Ai)(Br + Bi):
    //   A B:        re = p1 - p2, im = p3 - p1 - p2
    //   conj(A) B:  re = p1 + p2, im = p3 - p1 + p2
#pragma unroll
    for (int s = 0; s < SPW; ++s)
#pragma unroll
      for (int t = 0; t < CB; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const T p1 = ar[s][t][r], p2 = ai[s][t][r], p3 = a3[s][t][r];
          ar[s][t][r] = CONJ ? p1 + p2 : p1 - p2;
          ai[s][t][r] = CONJ ? (p3 - p1) + p2 : (p3 - p1) - p2;
        }
  }
}

// B[column][row] <- the wave's stripes of (xr, xi); rows >= N (products of clamped A rows) are
// stored as zero, so they contribute nothing to the next product
template <typename T, int SPW, int CB>
__device__ inline void write_b(T* Br, T* Bi, int LDK, int N, int NS, int NW, int wave, int fr,
                               int fk, const typename gemm::Acc<T>::v4 (&xr)[SPW][CB],
                               const typename gemm::Acc<T>::v4 (&xi)[SPW][CB]) {
#pragma unroll
  for (int s = 0; s < SPW; ++s) {
    if (wave + s * NW >= NS) continue;
    const int r0 = (wave + s * NW) * 16;
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      const int off = (16 * t + fr) * LDK + r0;
      if constexpr (sizeof(T) == 4) {
        // FP32 accumulator register r holds row 4 (lane / 16) + r: four consecutive rows
        const int rb = r0 + 4 * fk;
        float4 vr = float4{xr[s][t][0], xr[s][t][1], xr[s][t][2], xr[s][t][3]};
        float4 vi = float4{xi[s][t][0], xi[s][t][1], xi[s][t][2], xi[s][t][3]};
        if (rb + 3 >= N) {
          if (rb >= N) vr.x = vi.x = 0.f;
          if (rb + 1 >= N) vr.y = vi.y = 0.f;
          if (rb + 2 >= N) vr.z = vi.z = 0.f;
          vr.w = vi.w = 0.f;
        }
        *reinterpret_cast<float4*>(Br + off + 4 * fk) = vr;
        *reinterpret_cast<float4*>(Bi + off + 4 * fk) = vi;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = gemm::Acc<T>::row(fk, r);
          const bool in = r0 + row < N;
          Br[off + row] = in ? xr[s][t][r] : (T)0;
          Bi[off + row] = in ? xi[s][t][r] : (T)0;
        }
      }
    }
  }
}

// One wavefront's share of the propagation of one walker: the output stripes wave + s NW
// (s < SPW) against the CB 16-column tiles of the B operand at (Br, Bi) (LDS, [column][row],
// pitch LDK), which hold walker columns c0, c0 + 1, ...  Every wavefront of the workgroup passes
// the same sequence of barriers.  CONJV: the V products use conj(V), which is V^H for the
// complex-symmetric V of vHS (the adjoint propagator, sgn = -1).
template <typename T, bool P1C, int SPW, int CB, bool CONJV, bool M3, bool VEC, bool XB = false>
__device__ __attribute__((always_inline)) inline void prop_share(
    int N, int n, int order, double sgn, const T* __restrict__ P1, const T* __restrict__ Vw,
    z64* __restrict__ Pw, int c0, int NS, int NW, int wave, int fr, int fk, T* Br, T* Bi, int LDK,
    int e) {
  using V4 = typename gemm::Acc<T>::v4;
  V4 ar[SPW][CB], ai[SPW][CB], sr[SPW][CB], si[SPW][CB];
  // T_0 = P1 Phi, S = T_0
  stripes_times_b<T, P1C, SPW, CB, VEC>(P1, N, N, NS, NW, wave, fr, fk, Br, Bi, LDK, ar, ai);
#pragma unroll
  for (int s = 0; s < SPW; ++s)
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      sr[s][t] = ar[s][t];
      si[s][t] = ai[s][t];
    }
  __syncthreads();
  write_b<T, SPW, CB>(Br, Bi, LDK, N, NS, NW, wave, fr, fk, ar, ai);
  __syncthreads();
  for (int k = 1; k <= order; ++k) {
    stripes_times_b<T, true, SPW, CB, VEC, CONJV, M3>(Vw, N, N, NS, NW, wave, fr, fk, Br, Bi,
                                                 LDK, ar, ai);
    // T_k = (i sgn / k) V T_{k-1}; S += T_k
    const T fi = (T)(sgn / (double)k);
#pragma unroll
    for (int s = 0; s < SPW; ++s)
#pragma unroll
      for (int t = 0; t < CB; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const T tr = -(fi * ai[s][t][r]), ti = fi * ar[s][t][r];
          ar[s][t][r] = tr;
          ai[s][t][r] = ti;
          sr[s][t][r] = sr[s][t][r] + tr;
          si[s][t][r] = si[s][t][r] + ti;
        }
    __syncthreads();
    write_b<T, SPW, CB>(Br, Bi, LDK, N, NS, NW, wave, fr, fk, ar, ai);
    __syncthreads();
  }
  write_b<T, SPW, CB>(Br, Bi, LDK, N, NS, NW, wave, fr, fk, sr, si);
  __syncthreads();
  // Phi = P1 S, scaled back by 2^e
  stripes_times_b<T, P1C, SPW, CB, VEC>(P1, N, N, NS, NW, wave, fr, fk, Br, Bi, LDK, ar, ai);
  if constexpr (XB) __syncthreads();  // the split wavefronts' last partials (prop_share_split)
  const double sb = e == 0 ? 1.0 : ldexp(1.0, e);
#pragma unroll
  for (int s = 0; s < SPW; ++s) {
    if (wave + s * NW >= NS) continue;
#pragma unroll
    for (int t = 0; t < CB; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = (wave + s * NW) * 16 + gemm::Acc<T>::row(fk, r), col = c0 + 16 * t + fr;
        if (row < N && col < n)
          Pw[(long)row * n + col] = z64{(double)ar[s][t][r] * sb, (double)ai[s][t][r] * sb};
      }
  }
}

// One of the P wavefronts that share stripe st (one 16-column tile): part `part` multiplies the
// full k groups [part NGF / P, (part + 1) NGF / P) (the last part also the ragged group) and
// leaves its partial product in red [tail wavefront][lane]; after the barrier the owner (part 0)
// sums the P partials in part order and carries the stripe's share of the recurrence (the running
// sum, the LDS operand and the final store).  The barrier sequence is prop_share's with XB.
template <typename T, bool P1C, int CB, bool CONJV, bool M3, bool VEC>
__device__ __attribute__((always_inline)) inline void prop_share_split(
    int N, int n, int order, double sgn, const T* __restrict__ P1, const T* __restrict__ Vw,
    z64* __restrict__ Pw, int c0, int NS, int st, int part, int P, int tw, int fr, int fk, T* Br,
    T* Bi, int LDK, int e, typename gemm::Acc<T>::v4* red) {
  using V4 = typename gemm::Acc<T>::v4;
  const int lane = fr + 16 * fk;
  const int NGF = N / 16, gb = part * NGF / P, ge = (part + 1) * NGF / P;
  const bool rag = part == P - 1;
  const bool owner = part == 0;
  V4 ar[1][CB], ai[1][CB];
  // red [tail wavefront][tile][re, im][lane], then the owner's running sum S [tile][re, im][lane]
  // (in LDS rather than registers, which the products need)
  auto slot = [&](int w, int t, int c) { return ((w * CB + t) * 2 + c) * 64 + lane; };
  const int own = tw / P;  // the owner's index among the split stripes (0 or 1)
  auto sslot = [&](int t, int c) { return ((4 * CB + own * CB + t) * 2 + c) * 64 + lane; };
  auto put = [&]() {
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      red[slot(tw, t, 0)] = ar[0][t];
      red[slot(tw, t, 1)] = ai[0][t];
    }
  };
  auto gather = [&]() {  // the owner sums the P partials in part order
    const int t0 = tw - part;
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      ar[0][t] = red[slot(t0, t, 0)];
      ai[0][t] = red[slot(t0, t, 1)];
      for (int q = 1; q < P; ++q) {
        ar[0][t] += red[slot(t0 + q, t, 0)];
        ai[0][t] += red[slot(t0 + q, t, 1)];
      }
    }
  };
  // T_0 = P1 Phi, S = T_0
  stripes_times_b<T, P1C, 1, CB, VEC>(P1, N, N, NS, NS, st, fr, fk, Br, Bi, LDK, ar, ai, gb, ge,
                                       rag);
  put();
  __syncthreads();
  if (owner) {
    gather();
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      red[sslot(t, 0)] = ar[0][t];
      red[sslot(t, 1)] = ai[0][t];
    }
    write_b<T, 1, CB>(Br, Bi, LDK, N, NS, NS, st, fr, fk, ar, ai);
  }
  __syncthreads();
  for (int k = 1; k <= order; ++k) {
    stripes_times_b<T, true, 1, CB, VEC, CONJV, M3>(Vw, N, N, NS, NS, st, fr, fk, Br, Bi, LDK, ar,
                                                    ai, gb, ge, rag);
    put();
    __syncthreads();
    if (owner) {
      gather();
      const T fi = (T)(sgn / (double)k);
#pragma unroll
      for (int t = 0; t < CB; ++t) {
        V4 sr = red[sslot(t, 0)], si = red[sslot(t, 1)];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const T tr = -(fi * ai[0][t][r]), ti = fi * ar[0][t][r];
          ar[0][t][r] = tr;
          ai[0][t][r] = ti;
          sr[r] = sr[r] + tr;
          si[r] = si[r] + ti;
        }
        red[sslot(t, 0)] = sr;
        red[sslot(t, 1)] = si;
      }
      write_b<T, 1, CB>(Br, Bi, LDK, N, NS, NS, st, fr, fk, ar, ai);
    }
    __syncthreads();
  }
  if (owner) {
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      ar[0][t] = red[sslot(t, 0)];
      ai[0][t] = red[sslot(t, 1)];
    }
    write_b<T, 1, CB>(Br, Bi, LDK, N, NS, NS, st, fr, fk, ar, ai);
  }
  __syncthreads();
  // Phi = P1 S, scaled back by 2^e
  stripes_times_b<T, P1C, 1, CB, VEC>(P1, N, N, NS, NS, st, fr, fk, Br, Bi, LDK, ar, ai, gb, ge,
                                       rag);
  put();
  __syncthreads();
  if (owner) {
    gather();
    const double sb = e == 0 ? 1.0 : ldexp(1.0, e);
#pragma unroll
    for (int t = 0; t < CB; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = st * 16 + gemm::Acc<T>::row(fk, r), col = c0 + 16 * t + fr;
        if (row < N && col < n)
          Pw[(long)row * n + col] = z64{(double)ar[0][t][r] * sb, (double)ai[0][t][r] * sb};
      }
  }
}

// One workgroup per (walker, CB 16-column tiles).  Without TAIL, NW wavefronts own the stripes
// (SPW each).  TAIL (SPW = 1): NW, a multiple of 4, wavefronts own the first NW stripes with all
// CB tiles, and one more wavefront per (stripe, tile) of the remaining NS - NW stripes owns that
// single tile, so the four SIMDs carry nearly equal shares (N = 200, CB = 3: 10, 10, 10 and 9
// stripe-tiles instead of 12, 9, 9 and 9).  Every output tile is multiplied in the same order
// either way, so the variants give the same bits.
// KT (SPW = 1): NW (a multiple of 4) wavefronts own the first NW stripes, and the
// remaining R = NS - NW (1 or 2) stripes are each split over 4 / R more wavefronts by k
// (prop_share_split), so the four SIMDs carry equal shares.  The split stripes sum their k range
// in parts, so their bits differ from the unsplit kernel's.
template <typename T, bool P1C, int SPW, int CB, bool CONJV, bool M3, bool VEC, bool TAIL,
          bool KT = false>
__global__ void __launch_bounds__(1024)
    k_prop_taylor(int W, int N, int n, int order, double sgn, const T* __restrict__ P1,
                  const T* __restrict__ V, long v_bs, z64* __restrict__ Phi, long phi_bs,
                  int NW, int LDK) {
  constexpr int BC = 16 * CB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* Br = reinterpret_cast<T*>(smem_raw);
  T* Bi = Br + BC * LDK;
  const int nct = (n + BC - 1) / BC;
  const int NS = (N + 15) / 16;
  const long total = (long)W * nct;
  const long bid = blockIdx.x, xcd = bid % 8, slot = bid / 8;
  const long q = xcd * (total / 8) + min(xcd, total % 8) + slot;
  const int w = (int)(q / nct), c0 = (int)(q % nct) * BC;
  // the wavefront index is uniform across the wavefront: held in a scalar register, the stripe
  // tests below are scalar branches rather than lane masks
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int fr = lane & 15, fk = lane >> 4;
  z64* Pw = Phi + (long)w * phi_bs;
  const T* Vw = V + 2 * (long)w * v_bs;

  int e = 0;  // binary scale of the tile (uniform)
  if constexpr (sizeof(T) == 4) {
    // FP32: the tile enters scaled by 2^-e, e the binary exponent of its largest entry, and
    // leaves scaled back by 2^e.  The products are linear in the tile and power-of-two scaling
    // is exact, so the results are the unscaled ones wherever those stay in FP32's range; walker
    // matrices that grow between orthogonalizations past FP32's range no longer overflow.
    constexpr int KMAX = 4 * CB * SPW;  // >= BC NS 16 / blockDim (checked at launch)
    __shared__ double wmx[16];
    z64 vals[KMAX];
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < KMAX; ++i) {
      const int t = tid + i * (int)blockDim.x;
      const int cc = t % BC, k = t / BC;
      const z64 v = (t < BC * NS * 16 && k < N && c0 + cc < n) ? Pw[(long)k * n + c0 + cc]
                                                              : z64{0.0, 0.0};
      vals[i] = v;
      mx = fmax(mx, fmax(fabs(v.re), fabs(v.im)));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if (lane == 0) wmx[wave] = mx;
    __syncthreads();
    mx = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) mx = fmax(mx, wmx[w]);
    if (mx > 0.0 && mx <= 1.7976931348623157e308) e = __builtin_amdgcn_readfirstlane(ilogb(mx));
    const double sc = ldexp(1.0, -e);
#pragma unroll
    for (int i = 0; i < KMAX; ++i) {
      const int t = tid + i * (int)blockDim.x;
      if (t < BC * NS * 16) {
        const int cc = t % BC, k = t / BC;
        Br[cc * LDK + k] = (T)(vals[i].re * sc);
        Bi[cc * LDK + k] = (T)(vals[i].im * sc);
      }
    }
  } else {
    for (int t = tid; t < BC * NS * 16; t += blockDim.x) {
      const int cc = t % BC, k = t / BC;
      const z64 v = (k < N && c0 + cc < n) ? Pw[(long)k * n + c0 + cc] : z64{0.0, 0.0};
      Br[cc * LDK + k] = (T)v.re;
      Bi[cc * LDK + k] = (T)v.im;
    }
  }
  __syncthreads();
  if ((!TAIL && !KT) || wave < NW) {
    prop_share<T, P1C, SPW, CB, CONJV, M3, VEC, KT>(N, n, order, sgn, P1, Vw, Pw, c0, NS, NW,
                                                    wave, fr, fk, Br, Bi, LDK, e);
  } else if constexpr (KT) {
    const int tw = wave - NW, P = 4 / (NS - NW);
    auto* red = reinterpret_cast<typename gemm::Acc<T>::v4*>(Bi + BC * LDK);
    prop_share_split<T, P1C, CB, CONJV, M3, VEC>(N, n, order, sgn, P1, Vw, Pw, c0, NS,
                                                 NW + tw / P, tw % P, P, tw, fr, fk, Br, Bi, LDK,
                                                 e, red);
  } else if constexpr (TAIL) {
    const int u = wave - NW, st = NW + u / CB, tt = u % CB;
    prop_share<T, P1C, 1, 1, CONJV, M3, VEC>(N, n, order, sgn, P1, Vw, Pw, c0 + 16 * tt, NS, NW,
                                             st, fr, fk, Br + 16 * tt * LDK, Bi + 16 * tt * LDK,
                                             LDK, e);
  }
}

struct FusedShape {
  int NW, spw, LDK;
};

inline int max_lds_per_block() {
  static int v = [] {
    int dev = 0, x = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&x, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      x = 64 * 1024;
    return x;
  }();
  return v;
}

// TAIL: the first NS rounded down to a multiple of 4 stripes take one wavefront each, the rest
// one wavefront per (stripe, tile)
// LDS for the split wavefronts: partials (4 wavefronts x CB tiles x (re, im) x 64 lanes) and the
// running sums of the (at most 2) owners (CB tiles x (re, im) x 64 lanes each)
template <typename T>
constexpr size_t split_lds(int cb) {
  return (size_t)6 * cb * 2 * 64 * sizeof(typename gemm::Acc<T>::v4);
}

// KT: the first NS rounded down to a multiple of 4 stripes take one wavefront each and four more
// wavefronts split the rest by k
template <typename T, bool P1C, int SPW, int CB, bool CONJV, bool M3, bool TAIL = false,
          bool KT = false>
void launch_fused(hipStream_t s, const PropagateArgs& a, const FusedShape& f, const T* P1, int vecP,
                  const T* V, int vecV, z64* Phi) {
  static_assert(!KT || (SPW == 1 && !TAIL), "the split tail takes one stripe per wave");
  const size_t lds = 2 * 16 * CB * (size_t)f.LDK * sizeof(T) + (KT ? split_lds<T>(CB) : 0);
  auto kern = (vecP && vecV) ? k_prop_taylor<T, P1C, SPW, CB, CONJV, M3, true, TAIL, KT>
                             : k_prop_taylor<T, P1C, SPW, CB, CONJV, M3, false, TAIL, KT>;
  if (lds > 64 * 1024)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const int NS = ceil_div(a.N, 16);
  const int nw = (TAIL || KT) ? (NS & ~3) : f.NW;
  const int waves = TAIL ? nw + (NS - nw) * CB : KT ? nw + 4 : f.NW;
  AFQ_CHECK(!KT || ((NS - nw == 1 || NS - nw == 2) && waves <= 16 && nw >= 4), kErrShape,
            "split tail shape");
  // the FP32 tile load keeps 4 CB SPW entries per thread in registers
  AFQ_CHECK(sizeof(T) == 8 || 16L * CB * NS * 16 <= 4L * CB * SPW * 64 * waves, kErrShape,
            "propagation tile larger than the FP32 load covers");
  const long blocks = (long)a.W * ceil_div(a.n, 16 * CB);
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(64 * waves), lds, s, a.W, a.N, a.n,
                     a.order, CONJV ? -1.0 : 1.0, P1, V, a.v_bs, Phi, a.phi_bs, nw, f.LDK);
  AFQ_LAUNCH_CHECK();
}

// Launches the fused kernel when the shape fits it (returns false otherwise). Candidates differ
// in the number of 16-column tiles per workgroup (CB) and in the tail arrangement; they multiply
// every output tile in the same order, so the tuner (timed on a scratch copy of the walkers)
// never changes the result.
template <typename T>
bool propagate_fused(Context& c, const PropagateArgs& a, const T* P1, bool p1c, const void* V) {
  // the adjoint needs V^H; the kernel reads V by rows, so it takes V^H = conj(V) for a
  // complex-symmetric V only (FP64, the back-propagation)
  const bool adj = a.TA == 'H';
  if (adj && !(a.v_symmetric && sizeof(T) == 8)) return false;
  const int N = a.N, n = a.n, W = a.W;
  FusedShape f;
  const int NS = ceil_div(N, 16);
  f.NW = std::min(NS, 16);
  f.spw = ceil_div(NS, f.NW);
  f.LDK = 16 * NS + 4;
  // more stripes per wave spill registers (FP32 beyond 2, FP64 beyond 1)
  if (f.spw > (sizeof(T) == 4 ? 2 : 1)) return false;
  const size_t lds1 = 2 * 16 * (size_t)f.LDK * sizeof(T);
  // wider column tiles spill in FP64 and with two stripes per wave
  const int maxcb = sizeof(T) == 4 && f.spw == 1 ? 3 : 1;
  struct Cand {
    int cb;
    bool tail, split;
  };
  Cand cands[5];
  int ncand = 0;
  auto fits = [&](int cb) {
    return lds1 * cb <= (size_t)max_lds_per_block() && (cb == 1 || 16 * (cb - 1) < n);
  };
  // FP64, one stripe per wave and 1 or 2 stripes past a multiple of 4: those split by k over four
  // more wavefronts (AFQ_PROP_SPLIT=0 keeps one wavefront per stripe).  FP32 balances the SIMDs
  // by the tail arrangement instead, which measured the same as the split there (0.995 vs
  // 0.997 ms at N = 200, n = 40) without the extra barrier.  Every split candidate sums the split
  // stripes in the same parts, so they share their bits.
  const char* kte = std::getenv("AFQ_PROP_SPLIT");
  const int rem = NS % 4;
  const bool kt0 = sizeof(T) == 8 && !adj && f.spw == 1 && (rem == 1 || rem == 2) && NS >= 5 &&
                   (NS & ~3) + 4 <= 16 && !(kte && kte[0] == '0');
  if (kt0)
    for (int cb = 1; cb <= maxcb; ++cb)
      if (fits(cb) && lds1 * cb + split_lds<T>(cb) <= (size_t)max_lds_per_block())
        cands[ncand++] = {cb, false, true};
  const bool kt = ncand > 0;
  if (!kt) {
    for (int cb = 1; cb <= maxcb; ++cb)
      if (fits(cb)) cands[ncand++] = {cb, false, false};
    // FP32 with one stripe per wave and a stripe count that is not a multiple of 4: the tail
    // arrangement, which spreads the last stripes' tiles over the SIMDs
    if (sizeof(T) == 4 && f.spw == 1 && NS > 4 && NS % 4 != 0)
      for (int cb = 2; cb <= maxcb; ++cb)
        if (fits(cb) && (NS & ~3) + (NS % 4) * cb <= 16) cands[ncand++] = {cb, true, false};
  }
  if (ncand == 0) return false;
  AFQ_CHECK((long)W * ceil_div(n, 16) < (1L << 31), kErrShape, "grid too large");
  const size_t esz = sizeof(T) * (p1c ? 2 : 1);
  const int vecP = ((size_t)N * esz) % 16 == 0 && (reinterpret_cast<uintptr_t>(P1) % 16) == 0;
  const int vecV = ((size_t)N * 2 * sizeof(T)) % 16 == 0 &&
                   ((size_t)a.v_bs * 2 * sizeof(T)) % 16 == 0 &&
                   (reinterpret_cast<uintptr_t>(V) % 16) == 0;
  hipStream_t s = c.stream();
  const T* Vt = reinterpret_cast<const T*>(V);
  // The V products by Gauss's three-multiplication method (three MFMAs per complex block
  // instead of four): always in FP64; in FP32 only when AFQ_PROP_M3=1 (the four-product form
  // keeps the FP32 rounding of the terms closer to the FP64 path's)
  const char* m3e = std::getenv("AFQ_PROP_M3");
  const bool m3 = sizeof(T) == 8 || (m3e && m3e[0] == '1');
  auto run = [&](int i, z64* Phi) {
    const int cb = cands[i].cb;
    const bool tail = cands[i].tail, split = cands[i].split;
    constexpr int SPW2 = sizeof(T) == 4 ? 2 : 1;
    if constexpr (sizeof(T) == 8) {
      if (adj) {  // FP64 takes one stripe per wave and one column tile
        if (p1c) launch_fused<T, true, 1, 1, true, true>(s, a, f, P1, vecP, Vt, vecV, Phi);
        else launch_fused<T, false, 1, 1, true, true>(s, a, f, P1, vecP, Vt, vecV, Phi);
        return;
      }
    }
#define AFQ_PT(P1C, SPW, CB, TL)                                                         \
  (m3 ? launch_fused<T, P1C, SPW, CB, false, true, TL>(s, a, f, P1, vecP, Vt, vecV, Phi) \
      : launch_fused<T, P1C, SPW, CB, false, sizeof(T) == 8, TL>(s, a, f, P1, vecP, Vt, vecV, Phi))
#define AFQ_PS(P1C, CB)                                                                          \
  (m3 ? launch_fused<T, P1C, 1, CB, false, true, false, true>(s, a, f, P1, vecP, Vt, vecV, Phi) \
      : launch_fused<T, P1C, 1, CB, false, sizeof(T) == 8, false, true>(s, a, f, P1, vecP, Vt,    \
                                                                       vecV, Phi))
    if (split) {
      if constexpr (sizeof(T) == 4) {
        if (p1c) {
          if (cb == 1) AFQ_PS(true, 1);
          else if (cb == 2) AFQ_PS(true, 2);
          else AFQ_PS(true, 3);
        } else {
          if (cb == 1) AFQ_PS(false, 1);
          else if (cb == 2) AFQ_PS(false, 2);
          else AFQ_PS(false, 3);
        }
      } else {
        if (p1c) AFQ_PS(true, 1);
        else AFQ_PS(false, 1);
      }
    } else if (f.spw == 1) {
      if constexpr (sizeof(T) == 4) {
        if (p1c) {
          if (cb == 1) AFQ_PT(true, 1, 1, false);
          else if (cb == 2) tail ? AFQ_PT(true, 1, 2, true) : AFQ_PT(true, 1, 2, false);
          else tail ? AFQ_PT(true, 1, 3, true) : AFQ_PT(true, 1, 3, false);
        } else {
          if (cb == 1) AFQ_PT(false, 1, 1, false);
          else if (cb == 2) tail ? AFQ_PT(false, 1, 2, true) : AFQ_PT(false, 1, 2, false);
          else tail ? AFQ_PT(false, 1, 3, true) : AFQ_PT(false, 1, 3, false);
        }
      } else {
        if (p1c) AFQ_PT(true, 1, 1, false);
        else AFQ_PT(false, 1, 1, false);
      }
    } else {
      if (p1c) AFQ_PT(true, SPW2, 1, false);
      else AFQ_PT(false, SPW2, 1, false);
    }
#undef AFQ_PT
#undef AFQ_PS
  };
  tuned_on_copy(c, tune_key(adj ? "prop_fused_adj" : "prop_fused", sizeof(T), W, N, n, N), ncand,
                a.Phi, (size_t)W * a.phi_bs, run);
  return true;
}

// Phi <- P1 expM(op(V)) P1 Phi in compute type T.  TIn of the walker matrix is always double.
template <typename T>
void propagate_t(Context& c, const PropagateArgs& a) {
  hipStream_t s = c.stream();
  const int W = a.W, N = a.N, n = a.n;
  const long tb = (long)N * n;
  const T* P1r = nullptr;
  const cplx<T>* P1c = nullptr;
  if constexpr (sizeof(T) == 8) {
    P1r = a.P1r;
    P1c = a.P1;
  } else {
    P1r = a.P1r32;
    P1c = a.P1c32;
  }
  if (a.fused && propagate_fused<T>(c, a, P1r ? P1r : &P1c->re, P1r == nullptr, a.V)) return;
  cplx<T>* S = scratch_as<cplx<T>>(c, std::string("prop_S") + (sizeof(T) == 8 ? "64" : "32"),
                                    (size_t)W * tb);
  cplx<T>* T1 = scratch_as<cplx<T>>(c, std::string("prop_T1") + (sizeof(T) == 8 ? "64" : "32"),
                                     (size_t)W * tb);
  cplx<T>* T2 = scratch_as<cplx<T>>(c, std::string("prop_T2") + (sizeof(T) == 8 ? "64" : "32"),
                                     (size_t)W * tb);
  // TMN = P1 * Phi  (written to S and T1)
  {
    Opnd<double, true, false> B{&a.Phi->re, n, a.phi_bs};
    EpiStore2<T> e{S, T1, n, tb};
    if (P1r) {
      Opnd<T, false, true> A{P1r, N, 0};
      gemm_tall<T>(c, "prop_P1phi_r", W, N, n, N, A, B, e);
    } else {
      Opnd<T, true, true> A{&P1c->re, N, 0};
      gemm_tall<T>(c, "prop_P1phi_c", W, N, n, N, A, B, e);
    }
  }
  // Taylor terms: T2 = (+-i/k) op(V) T1 ; S += T2.  The accumulating epilogue cannot be timed
  // in place, so the terms use the tile tuned for the P1 * Phi product of the same shape.
  const cplx<T>* V = reinterpret_cast<const cplx<T>*>(a.V);
  int tc = Tuner::get().lookup(tune_key(P1r ? "prop_P1phi_r" : "prop_P1phi_c", sizeof(T), W, N, n, N));
  auto taylor = [&](auto const& A, auto const& B, auto const& e) {
    if (tc == 1)
      tall_bn<T, 208, 13>(s, W, N, n, N, A, B, e);
    else if (tc == 2)
      tall_bn<T, 128, 8>(s, W, N, n, N, A, B, e);
    else
      tall_bn<T, 64, 4>(s, W, N, n, N, A, B, e);
  };
  for (int k = 1; k <= a.order; ++k) {
    T fi = (T)((a.TA == 'H' ? -1.0 : 1.0) / (double)k);
    Opnd<T, true, false> B{&T1->re, n, tb};
    EpiTaylor<T> e{T2, S, n, tb, (T)0, fi};
    if (a.TA == 'H') {
      Opnd<T, true, false, true> A{&V->re, N, a.v_bs};  // A(i, j) = conj(V[j][i])
      taylor(A, B, e);
    } else {
      Opnd<T, true, true> A{&V->re, N, a.v_bs};
      taylor(A, B, e);
    }
    std::swap(T1, T2);
  }
  // Phi = P1 * S
  {
    Opnd<T, true, false> B{&S->re, n, tb};
    EpiStore<T, double> e{a.Phi, n, a.phi_bs, (T)1, (T)0, 0};
    if (P1r) {
      Opnd<T, false, true> A{P1r, N, 0};
      gemm_tall<T>(c, "prop_P1S_r", W, N, n, N, A, B, e);
    } else {
      Opnd<T, true, true> A{&P1c->re, N, 0};
      gemm_tall<T>(c, "prop_P1S_c", W, N, n, N, A, B, e);
    }
  }
}

}  // namespace

void propagate(Context& c, const PropagateArgs& a) {
  if (a.W == 0 || a.n == 0) return;
  AFQ_CHECK(a.TA == 'N' || a.TA == 'H', kErrArg, "TA must be 'N' or 'H'");
  if (a.prec == kFP64) {
    AFQ_CHECK(a.P1r || a.P1, kErrArg, "P1 missing");
    propagate_t<double>(c, a);
  } else {
    AFQ_CHECK(a.P1r32 || a.P1c32, kErrArg, "FP32 P1 missing");
    propagate_t<float>(c, a);
  }
}

void orthogonalize(Context& c, z64* Phi, long phi_bs, int W, int N, int n, z64* ldet) {
  if (W == 0 || n == 0) return;
  hipStream_t s = c.stream();
  z64* S = scratch_as<z64>(c, "qr_S", (size_t)W * n * n);
  z64* Ri = scratch_as<z64>(c, "qr_Rinv", (size_t)W * n * n);
  z64* Q = scratch_as<z64>(c, "qr_Q", (size_t)W * N * n);
  // two Cholesky-QR passes: Phi -> Q -> Phi
  for (int pass = 0; pass < 2; ++pass) {
    const z64* X = pass == 0 ? Phi : Q;
    long xbs = pass == 0 ? phi_bs : (long)N * n;
    z64* Y = pass == 0 ? Q : Phi;
    long ybs = pass == 0 ? (long)N * n : phi_bs;
    {  // S = X^H X
      Opnd<double, true, false, true> A{&X->re, n, xbs};
      Opnd<double, true, false> B{&X->re, n, xbs};
      zgemm_small(c, "qr_gram", W, n, n, N, A, B, EpiStore<double, double>{S, n, (long)n * n, 1, 0, 0});
    }
    chol_upper_inverse(s, S, W, n, Ri, ldet);
    {  // Y = X R^{-1}
      Opnd<double, true, true> A{&X->re, n, xbs};
      Opnd<double, true, false> B{&Ri->re, n, (long)n * n};
      zgemm_small(c, "qr_Q", W, N, n, n, A, B, EpiStore<double, double>{Y, n, ybs, 1, 0, 0});
    }
  }
}

}  // namespace det_ops
}  // namespace afq
