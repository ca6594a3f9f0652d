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
  Buf& b = bufs_[name];
  if (b.bytes >= bytes && b.p) return b.p;
  AFQ_CHECK(!capturing(), kErrState,
            "scratch buffer '" + name + "' would grow during graph capture");
  if (b.p) {
    AFQ_HIP(hipStreamSynchronize(stream_));
    AFQ_HIP(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
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
                 z64* ovlp, double scale, z64 add) {
  if (n == 0 || W == 0) return;
  hipStream_t s = c.stream();
  z64* T = scratch_as<z64>(c, "lo_T", (size_t)W * n * n);
  Opnd<double, true, true> A{&PsiT->re, N, 0};
  Opnd<double, true, false> B{&Phi->re, n, phi_bs};
  gemm_small<double>(c, "lo_T", W, n, n, N, A, B, EpiStore<double, double>{T, n, (long)n * n, 1, 0, 0});
  lu_inverse_logdet(s, T, (long)n * n, W, n, nullptr, 0, ovlp, scale, add, false);
}

void mixed_density_matrix(Context& c, const z64* PsiT, int n, int N, const z64* Phi, long phi_bs,
                          int W, z64* G, long g_bs, z64* ovlp, bool compact, double scale) {
  if (W == 0) return;
  hipStream_t s = c.stream();
  if (n == 0) {
    if (!compact) fill_z64(s, (long)W * g_bs, {0, 0}, G);
    return;
  }
  z64* T = scratch_as<z64>(c, "mdm_T", (size_t)W * n * n);
  {
    Opnd<double, true, true> A{&PsiT->re, N, 0};
    Opnd<double, true, false> B{&Phi->re, n, phi_bs};
    gemm_small<double>(c, "mdm_T", W, n, n, N, A, B, EpiStore<double, double>{T, n, (long)n * n, 1, 0, 0});
  }
  lu_inverse_logdet(s, T, (long)n * n, W, n, T, (long)n * n, ovlp, scale, {0, 0}, true);
  // compact G[a][k] = sum_b Tinv[b][a] Phi[k][b]    (product<'T','T'>(TNN, B, C))
  z64* Gc = compact ? G : scratch_as<z64>(c, "mdm_Gc", (size_t)W * n * N);
  long gc_bs = compact ? g_bs : (long)n * N;
  {
    Opnd<double, true, false> A{&T->re, n, (long)n * n};
    Opnd<double, true, true> B{&Phi->re, n, phi_bs};
    gemm_small<double>(c, "mdm_G", W, n, N, n, A, B, EpiStore<double, double>{Gc, N, gc_bs, 1, 0, 0});
  }
  if (!compact) {
    // G[i][k] = sum_a PsiT[a][i] Gc[a][k]           (product<'T'>(A, TNM, C))
    Opnd<double, true, false> A{&PsiT->re, N, 0};
    Opnd<double, true, false> B{&Gc->re, N, (long)n * N};
    gemm_small<double>(c, "mdm_full", W, N, N, n, A, B, EpiStore<double, double>{G, N, g_bs, 1, 0, 0});
  }
}

void inverse_logdet(Context& c, const z64* A, int nb, int n, z64* ovlp, z64* Ainv) {
  lu_inverse_logdet(c.stream(), A, (long)n * n, nb, n, Ainv, (long)n * n, ovlp, 1.0, {0, 0}, true);
}

namespace {

// Phi <- P1 expM(op(V)) P1 Phi in compute type T.  TIn of the walker matrix is always double.
template <typename T>
void propagate_t(Context& c, const PropagateArgs& a) {
  hipStream_t s = c.stream();
  const int W = a.W, N = a.N, n = a.n;
  const long tb = (long)N * n;
  cplx<T>* S = scratch_as<cplx<T>>(c, std::string("prop_S") + (sizeof(T) == 8 ? "64" : "32"),
                                    (size_t)W * tb);
  cplx<T>* T1 = scratch_as<cplx<T>>(c, std::string("prop_T1") + (sizeof(T) == 8 ? "64" : "32"),
                                     (size_t)W * tb);
  cplx<T>* T2 = scratch_as<cplx<T>>(c, std::string("prop_T2") + (sizeof(T) == 8 ? "64" : "32"),
                                     (size_t)W * tb);
  const T* P1r = nullptr;
  const cplx<T>* P1c = nullptr;
  if constexpr (sizeof(T) == 8) {
    P1r = a.P1r;
    P1c = a.P1;
  } else {
    P1r = a.P1r32;
    P1c = a.P1c32;
  }
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
      gemm_small<double>(c, "qr_gram", W, n, n, N, A, B, EpiStore<double, double>{S, n, (long)n * n, 1, 0, 0});
    }
    chol_upper_inverse(s, S, W, n, Ri, ldet);
    {  // Y = X R^{-1}
      Opnd<double, true, true> A{&X->re, n, xbs};
      Opnd<double, true, false> B{&Ri->re, n, (long)n * n};
      gemm_tall<double>(c, "qr_Q", W, N, n, n, A, B, EpiStore<double, double>{Y, n, ybs, 1, 0, 0});
    }
  }
}

}  // namespace det_ops
}  // namespace afq
