// Cholesky (real 3-index factorization) hamiltonian operations on the device.
//
//   setup        RealDenseHamiltonian::getHamiltonianOperations (RealDenseHamiltonian.cpp:139-181):
//                Lnak[m][a][k] = sum_i PsiT[a][i] L[i][k][m], haj = PsiT h1,
//                vexx[i][l] = -1/2 sum_{j,m} L[i][j][m] L[l][j][m]
//   vbias        Real3IndexFactorization::vbias (Real3IndexFactorization.hpp:328-428)
//   vHS          Real3IndexFactorization::vHS (:282-326)
//   energy       Real3IndexFactorization::energy + energy_impl (:138-174, :476-590)
//   vMF          NOMSD::vMF (NOMSD.cpp:437-464) + zero_imag (AFQMCBasePropagator.cpp:105-108)
//   one-body P1  generateP1 (AFQMCBasePropagator.cpp:43-214), getOneBodyPropagatorMatrix (:99-124),
//                add_one_body_shifts (one_body.hpp), exp_hermitian (exp.hpp:28-60)
//
// The large contractions are single MFMA GEMMs over the walker batch: vbias is
// [W x nN] . [nN x M], vHS is [W x M] . [M x N^2] with a real right operand, accumulated from
// the last Cholesky vector to the first (smallest norms first).  When the trial is
// real the half-rotated tensor is real and vbias uses the real-operand (2-MFMA) form.  The
// exchange energy follows the reference's chunking over Cholesky vectors: one GEMM per chunk
// builds T[(w,b)][(m,a)] and a fixed-order reduction forms sum_ab T_ab T_ba and the Coulomb
// traces.  The one-body exponential uses scaling and squaring of a degree-20 Taylor
// polynomial on the device instead of the reference's host eigendecomposition; both agree to
// rounding for the Hermitian H1.
#include <cmath>
#include <cstdlib>
#include <vector>

#include "crgemm.hpp"
#include "gemm.hpp"
#include "ops.hpp"
#include "tuner.hpp"

namespace afq {

using gemm::EpiStore;
using gemm::EpiStoreReal;
using gemm::Opnd;

namespace {

template <typename T, class OA, class OB, class EPI>
void launch_large(hipStream_t s, int batch, int M, int N, int K, const OA& A, const OB& B,
                  const EPI& e) {
  gemm::launch<T, 64, 64, 16, 2, 2>(s, batch, M, N, K, A, B, e);
}
template <typename T, class OA, class OB, class EPI>
void launch_small(hipStream_t s, int batch, int M, int N, int K, const OA& A, const OB& B,
                  const EPI& e) {
  gemm::launch<T, 32, 32, 16, 1, 1>(s, batch, M, N, K, A, B, e);
}

template <typename T>
T* dalloc(size_t count) {
  void* p = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc(&p, count * sizeof(T)) != hipSuccess)
    throw Error(kErrAlloc, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
  return reinterpret_cast<T*>(p);
}

}  // namespace

CholeskyOps::CholeskyOps(Context& c, int walker_type, int N, int nup, int ndown, int M,
                         const double* Likn, const double* h1, const z64* PsiT, double E0,
                         int precision, int nspin_L, int nspin_h1)
    : ctx_(c), walker_type_(walker_type), N_(N), nup_(nup), ndown_(ndown), M_(M), prec_(precision),
      nl_(nspin_L), nh_(nspin_h1), E0_(E0) {
  AFQ_CHECK(walker_type == kClosed || walker_type == kCollinear, kErrArg, "walker_type");
  AFQ_CHECK(N > 0 && M > 0 && nup >= 0 && ndown >= 0 && nup <= N && ndown <= N, kErrShape,
            "invalid sizes");
  AFQ_CHECK(nup <= 64 && ndown <= 64, kErrShape, "electrons per spin must be <= 64");
  AFQ_CHECK(precision == kFP64 || precision == kMixed, kErrArg, "precision");
  AFQ_CHECK((nspin_L == 1 || nspin_L == 2) && (nspin_h1 == 1 || nspin_h1 == 2), kErrArg,
            "the Cholesky tensor and h1 have 1 or 2 spin blocks");
  // CLOSED walkers take spin-independent integrals (walkerDimsAreConvertible)
  AFQ_CHECK(walker_type == kCollinear || (nspin_L == 1 && nspin_h1 == 1), kErrArg,
            "spin-dependent integrals need COLLINEAR walkers");
  if (walker_type == kClosed) ndown_ = 0;
  hipStream_t s = c.stream();
  const int nel_ = nel();
  try {
    const size_t NNM = (size_t)N * N * M;
    Likn_ = dalloc<double>(nl_ * NNM);
    h1_ = dalloc<double>((size_t)nh_ * N * N);
    vexx_ = dalloc<double>((size_t)nl_ * N * N);
    AFQ_HIP(hipMemcpyAsync(Likn_, Likn, sizeof(double) * nl_ * NNM, hipMemcpyDeviceToDevice, s));
    AFQ_HIP(hipMemcpyAsync(h1_, h1, sizeof(double) * nh_ * N * N, hipMemcpyDeviceToDevice, s));
    sym_L_ = true;
    for (int l = 0; l < nl_; ++l) sym_L_ = symmetric_cholesky(s, N, M, Likn_ + l * NNM) && sym_L_;
    // vexx_l[i][k] = -1/2 sum_{(j,m)} Likn_l[i][(j,m)] Likn_l[k][(j,m)]
    // (RealDenseHamiltonian.cpp:165-181, one block per spin block of L)
    for (int l = 0; l < nl_; ++l) {
      Opnd<double, false, true> A{Likn_ + l * NNM, (long)N * M, 0};
      Opnd<double, false, true> B{Likn_ + l * NNM, (long)N * M, 0};
      launch_large<double>(s, 1, N, N, N * M, A, B,
                           EpiStoreReal<double>{vexx_ + (size_t)l * N * N, N, 0, -0.5});
    }
    if (prec_ == kMixed) {
      Likn32_ = dalloc<float>(nl_ * NNM);
      cast_f64_f32(s, (long)(nl_ * NNM), Likn_, Likn32_);
    }
    Rotated own;
    half_rotate_into(PsiT, own);
    PsiT_ = own.PsiT;
    Lnak_ = own.Lnak;
    haj_ = own.haj;
    Lnak_r_ = own.Lnak_r;
    Lnak_r32_ = own.Lnak_r32;
    Lnak32_ = own.Lnak32;
    real_trial_ = own.real;
    AFQ_HIP(hipStreamSynchronize(s));
  } catch (...) {
    free_all();
    throw;
  }
}

// The half-rotated tensors of one determinant (RealDenseHamiltonian.cpp:141-163): per spin
// block C[a][(k,m)] = sum_i PsiT[a][i] Likn[i][(k,m)], permuted to Lnak[m][a][k]; haj = PsiT h1
// (half_rotate_hamiltonian, one_body.hpp); the real and FP32 copies the contractions read.
void CholeskyOps::half_rotate_into(const z64* PsiT, Rotated& r) const {
  hipStream_t s = ctx_.stream();
  const int N = N_, M = M_, nel_ = nel();
  try {
    r.PsiT = dalloc<z64>((size_t)nel_ * N);
    r.Lnak = dalloc<z64>((size_t)M * nel_ * N);
    r.haj = dalloc<z64>((size_t)nel_ * N);
    AFQ_HIP(hipMemcpyAsync(r.PsiT, PsiT, sizeof(z64) * nel_ * N, hipMemcpyDeviceToDevice, s));
    // real determinant?  (NOMSD.hpp:70-86 makes the same test on the host)
    std::vector<z64> hp((size_t)nel_ * N);
    AFQ_HIP(hipMemcpyAsync(hp.data(), r.PsiT, sizeof(z64) * hp.size(), hipMemcpyDeviceToHost, s));
    AFQ_HIP(hipStreamSynchronize(s));
    r.real = true;
    for (auto& v : hp)
      if (v.im != 0.0) { r.real = false; break; }
    z64* tmp = nullptr;
    const int nmax = std::max(nup_, ndown_);
    if (nmax > 0) tmp = scratch_as<z64>(ctx_, "setup_tmp", (size_t)nmax * N * M);
    for (int sp = 0; sp < (walker_type_ == kCollinear ? 2 : 1); ++sp) {
      const int ns = sp == 0 ? nup_ : ndown_;
      if (ns == 0) continue;
      const z64* Ps = r.PsiT + (sp == 0 ? 0 : (size_t)nup_ * N);
      z64* Ls = r.Lnak + (sp == 0 ? 0 : (size_t)M * nup_ * N);
      Opnd<double, true, true> A{&Ps->re, N, 0};
      Opnd<double, false, false> B{Likn_ + (size_t)(sp % nl_) * N * N * M, (long)N * M, 0};
      launch_large<double>(s, 1, ns, N * M, N, A, B,
                           EpiStore<double, double>{tmp, (long)N * M, 0, 1, 0, 0});
      permute_akm_mak(s, ns, N, M, tmp, Ls);
    }
    if (nh_ == 1) {
      Opnd<double, true, true> A{&r.PsiT->re, N, 0};
      Opnd<double, false, false> B{h1_, N, 0};
      launch_small<double>(s, 1, nel_, N, N, A, B, EpiStore<double, double>{r.haj, N, 0, 1, 0, 0});
    } else {
      // haj of each spin block from its own h1 (half_rotate_hamiltonian, one_body.hpp:123-160)
      for (int sp = 0; sp < 2; ++sp) {
        const int ns = sp == 0 ? nup_ : ndown_;
        if (ns == 0) continue;
        const size_t off = sp == 0 ? 0 : (size_t)nup_ * N;
        Opnd<double, true, true> A{&(r.PsiT + off)->re, N, 0};
        Opnd<double, false, false> B{h1_ + (size_t)sp * N * N, N, 0};
        launch_small<double>(s, 1, ns, N, N, A, B,
                             EpiStore<double, double>{r.haj + off, N, 0, 1, 0, 0});
      }
    }
    if (r.real) {
      r.Lnak_r = dalloc<double>((size_t)M * nel_ * N);
      real_part(s, (long)M * nel_ * N, r.Lnak, r.Lnak_r);
    }
    if (prec_ == kMixed) {
      if (r.real) {
        r.Lnak_r32 = dalloc<float>((size_t)M * nel_ * N);
        real_part_f32(s, (long)M * nel_ * N, r.Lnak, r.Lnak_r32);
      } else {
        r.Lnak32 = dalloc<c32>((size_t)M * nel_ * N);
        cast_z64_c32(s, (long)M * nel_ * N, r.Lnak, r.Lnak32);
      }
    }
  } catch (...) {
    release(r);
    throw;
  }
}

Rotated CholeskyOps::rotate(const z64* PsiT) const {
  Rotated r;
  half_rotate_into(PsiT, r);
  AFQ_HIP(hipStreamSynchronize(ctx_.stream()));
  return r;
}

void CholeskyOps::release(Rotated& r) const {
  (void)hipStreamSynchronize(ctx_.stream());
  for (void* p : {(void*)r.PsiT, (void*)r.Lnak, (void*)r.haj, (void*)r.Lnak_r, (void*)r.Lnak_r32,
                  (void*)r.Lnak32})
    if (p) (void)hipFree(p);
  r = Rotated{};
}

void CholeskyOps::free_all() {
  for (void* p : {(void*)PsiT_, (void*)Likn_, (void*)Likn32_, (void*)h1_, (void*)vexx_,
                  (void*)Lnak_, (void*)Lnak_r_, (void*)Lnak_r32_, (void*)Lnak32_, (void*)haj_})
    if (p) (void)hipFree(p);
  PsiT_ = nullptr;
  Likn_ = nullptr;
  Likn32_ = nullptr;
  h1_ = vexx_ = nullptr;
  Lnak_ = nullptr;
  Lnak_r_ = nullptr;
  Lnak_r32_ = nullptr;
  Lnak32_ = nullptr;
  haj_ = nullptr;
}

CholeskyOps::~CholeskyOps() {
  (void)hipStreamSynchronize(ctx_.stream());
  free_all();
}

void CholeskyOps::vbias(const z64* G, int W, double dt, z64* v, Precision prec) const {
  if (W == 0) return;
  hipStream_t s = ctx_.stream();
  const double a = std::sqrt(dt) * (walker_type_ == kClosed ? 2.0 : 1.0);
  const int nel_ = nel();
  // spin blocks: G[w][(off+a)][k] with Lnak block s [M][ns][N]
  for (int sp = 0; sp < (walker_type_ == kCollinear ? 2 : 1); ++sp) {
    int ns = sp == 0 ? nup_ : ndown_;
    int acc = (sp == 1 && nup_ > 0) ? 1 : 0;
    if (ns == 0) {
      if (sp == 0) fill_z64(s, (long)W * M_, {0, 0}, v);
      continue;
    }
    size_t goff = sp == 0 ? 0 : (size_t)nup_ * N_;
    size_t loff = sp == 0 ? 0 : (size_t)M_ * nup_ * N_;
    const long K = (long)ns * N_;
    if (prec == kFP64) {
      // A(w, (a,k)) = G[w][off + a][k]: rows of length nel*N, K = ns*N contiguous
      Opnd<double, true, true> A{&(G + goff)->re, (long)nel_ * N_, 0};
      EpiStore<double, double> e{v, M_, 0, a, 0, acc};
      int ks = 0, kc = 0;
      if (real_trial_ && crg_path() &&
          crg::crg_supported<double>(G + goff, (long)nel_ * N_, W, Lnak_r_ + loff, K, M_))
        crg::crg_split(W, M_, (int)K, 4, 3, ks, kc);
      if (ks > 0 && crg::crg_fills(W, M_, 4, 3, ks)) {
        // register-tiled, K cut into slices so that the 64 x 48 tiles fill the machine
        z64* ws = ks > 1 ? scratch_as<z64>(ctx_, "crg_ws64", (size_t)ks * W * M_) : nullptr;
        crg::launch_crgemm<double, 4, 3, false>(s, W, M_, (int)K, G + goff, (long)nel_ * N_,
                                                Lnak_r_ + loff, crg::RowsPlain{K}, ks, kc, e, ws);
      } else if (real_trial_) {
        Opnd<double, false, true> B{Lnak_r_ + loff, K, 0};
        gemm_large_splitk<double>(ctx_, "vbias_r", W, M_, (int)K, A, B, e, acc == 0);
      } else {
        Opnd<double, true, true> B{&(Lnak_ + loff)->re, K, 0};
        gemm_large_splitk<double>(ctx_, "vbias_c", W, M_, (int)K, A, B, e, acc == 0);
      }
    } else {
      AFQ_CHECK(prec_ == kMixed, kErrState, "FP32 tensors were not built (precision FP64)");
      Opnd<double, true, true> A{&(G + goff)->re, (long)nel_ * N_, 0};
      EpiStore<float, double> e{v, M_, 0, (float)a, 0.f, acc};
      if (real_trial_) {
        Opnd<float, false, true> B{Lnak_r32_ + loff, K, 0};
        gemm_large_splitk<float>(ctx_, "vbias_r", W, M_, (int)K, A, B, e, acc == 0);
      } else {
        Opnd<float, true, true> B{&(Lnak32_ + loff)->re, K, 0};
        gemm_large_splitk<float>(ctx_, "vbias_c", W, M_, (int)K, A, B, e, acc == 0);
      }
    }
  }
}

void CholeskyOps::vHS(const void* X, int W, double dt, void* V, Precision prec) const {
  if (W == 0) return;
  hipStream_t s = ctx_.stream();
  const double a = std::sqrt(dt);
  const long NN = (long)N_ * N_;
  const size_t esz = prec == kFP64 ? sizeof(z64) : sizeof(c32);
  for (int l = 0; l < nl_; ++l) {
    // block l of V [nl][W][N][N] from block l of L
    char* Vl = reinterpret_cast<char*>(V) + (size_t)l * W * NN * esz;
    const size_t loff = (size_t)l * NN * M_;
    if (sym_L_) {
      // V[w] is complex symmetric: compute the 8x8 blocks of the upper triangle of (i, k) pairs
      // and write each value to (i, k) and (k, i)
      const int nb = ceil_div(N_, 8);
      const int npairs = nb * (nb + 1) / 2 * 64;
      if (prec == kFP64) {
        gemm::EpiStoreSym<double, double> e{reinterpret_cast<z64*>(Vl), NN, 0, a, 0, N_, nb};
        const z64* X64 = reinterpret_cast<const z64*>(X);
        if (crg_path() && crg::crg_supported<double>(X64, M_, W, Likn_ + loff, M_, NN) &&
            crg::crg_fills(W, npairs, 4, 3)) {
          // register-tiled: 64 walkers x 48 pairs per wavefront (the tile count fills the
          // machine's last round of wavefront slots), k from the last Cholesky vector
          crg::launch_crgemm<double, 4, 3, true>(s, W, npairs, M_, X64, M_, Likn_ + loff,
                                                 crg::RowsSymPairs{M_, N_, nb}, 1,
                                                 ceil_div(M_, 16) * 16, e, nullptr);
        } else {
          gemm::RevK<Opnd<double, true, true>> A({reinterpret_cast<const double*>(X), M_, 0}, M_);
          gemm::RevK<gemm::OpndSymPairs<double>> B({Likn_ + loff, M_, 0, N_, nb}, M_);
          gemm_large<double>(ctx_, "vhs_sym", 1, W, npairs, M_, A, B, e);
        }
      } else {
        AFQ_CHECK(prec_ == kMixed, kErrState, "FP32 tensors were not built (precision FP64)");
        gemm::EpiStoreSym<float, float> e{reinterpret_cast<c32*>(Vl), NN, 0, (float)a, 0.f, N_, nb};
        const c32* X32 = reinterpret_cast<const c32*>(X);
        if (crg_path() && crg::crg_supported<float>(X32, M_, W, Likn32_ + loff, M_, NN) &&
            crg::crg_fills(W, npairs, 4, 7)) {
          // register-tiled: 64 walkers x 112 pairs per wavefront, k groups from the last
          // Cholesky vector to the first
          crg::launch_crgemm<float, 4, 7, true>(s, W, npairs, M_, X32, M_, Likn32_ + loff,
                                                crg::RowsSymPairs{M_, N_, nb}, 1,
                                                ceil_div(M_, 16) * 16, e, nullptr);
        } else {
          gemm::RevK<Opnd<float, true, true>> A({reinterpret_cast<const float*>(X), M_, 0}, M_);
          gemm::RevK<gemm::OpndSymPairs<float>> B({Likn32_ + loff, M_, 0, N_, nb}, M_);
          gemm_large<float>(ctx_, "vhs_sym", 1, W, npairs, M_, A, B, e);
        }
      }
      continue;
    }
    if (prec == kFP64) {
      gemm::RevK<Opnd<double, true, true>> A({reinterpret_cast<const double*>(X), M_, 0}, M_);
      gemm::RevK<Opnd<double, false, true>> B({Likn_ + loff, M_, 0}, M_);
      gemm_large<double>(ctx_, "vhs", 1, W, (int)NN, M_, A, B,
                         EpiStore<double, double>{reinterpret_cast<z64*>(Vl), NN, 0, a, 0, 0});
    } else {
      AFQ_CHECK(prec_ == kMixed, kErrState, "FP32 tensors were not built (precision FP64)");
      gemm::RevK<Opnd<float, true, true>> A({reinterpret_cast<const float*>(X), M_, 0}, M_);
      gemm::RevK<Opnd<float, false, true>> B({Likn32_ + loff, M_, 0}, M_);
      gemm_large<float>(ctx_, "vhs", 1, W, (int)NN, M_, A, B,
                        EpiStore<float, float>{reinterpret_cast<c32*>(Vl), NN, 0, (float)a, 0.f,
                                               0});
    }
  }
}

void CholeskyOps::vbias32(const c32* G, int W, double dt, z64* v) const {
  if (W == 0) return;
  AFQ_CHECK(prec_ == kMixed, kErrState, "FP32 tensors were not built (precision FP64)");
  hipStream_t s = ctx_.stream();
  const float a = (float)(std::sqrt(dt) * (walker_type_ == kClosed ? 2.0 : 1.0));
  const int nel_ = nel();
  for (int sp = 0; sp < (walker_type_ == kCollinear ? 2 : 1); ++sp) {
    int ns = sp == 0 ? nup_ : ndown_;
    int acc = (sp == 1 && nup_ > 0) ? 1 : 0;
    if (ns == 0) {
      if (sp == 0) fill_z64(s, (long)W * M_, {0, 0}, v);
      continue;
    }
    size_t goff = sp == 0 ? 0 : (size_t)nup_ * N_;
    size_t loff = sp == 0 ? 0 : (size_t)M_ * nup_ * N_;
    const long K = (long)ns * N_;
    Opnd<float, true, true> A{&(G + goff)->re, (long)nel_ * N_, 0};
    EpiStore<float, double> e{v, M_, 0, a, 0.f, acc};
    int ks = 0, kc = 0;
    if (real_trial_ && crg_path() &&
        crg::crg_supported<float>(G + goff, (long)nel_ * N_, W, Lnak_r32_ + loff, K, M_))
      crg::crg_split(W, M_, (int)K, 4, 7, ks, kc);
    if (ks > 0 && crg::crg_fills(W, M_, 4, 7, ks)) {
      // register-tiled, K cut into slices so that the 64 x 112 tiles fill the machine
      cplx<float>* ws =
          ks > 1 ? scratch_as<cplx<float>>(ctx_, "crg_ws32", (size_t)ks * W * M_) : nullptr;
      crg::launch_crgemm<float, 4, 7, false>(s, W, M_, (int)K, G + goff, (long)nel_ * N_,
                                             Lnak_r32_ + loff, crg::RowsPlain{K}, ks, kc, e, ws);
    } else if (real_trial_) {
      Opnd<float, false, true> B{Lnak_r32_ + loff, K, 0};
      gemm_large_splitk<float>(ctx_, "vbias32_r", W, M_, (int)K, A, B, e, acc == 0);
    } else {
      Opnd<float, true, true> B{&(Lnak32_ + loff)->re, K, 0};
      gemm_large_splitk<float>(ctx_, "vbias32_c", W, M_, (int)K, A, B, e, acc == 0);
    }
  }
}

void CholeskyOps::vbias_full(const z64* G, int W, double dt, z64* v, Precision prec) const {
  if (W == 0) return;
  const double a = std::sqrt(dt) * (walker_type_ == kClosed ? 2.0 : 1.0);
  const int nspin = walker_type_ == kCollinear ? 2 : 1;
  const long NN = (long)N_ * N_;
  // v[w][m] (+)= a sum_{(i,k)} G[w][s][(i,k)] Likn[(i,k)][m], one product per spin block
  for (int sp = 0; sp < nspin; ++sp) {
    Opnd<double, true, true> A{&(G + sp * NN)->re, nspin * NN, 0};
    const size_t loff = (size_t)(sp % nl_) * NN * M_;
    if (prec == kFP64) {
      Opnd<double, false, false> B{Likn_ + loff, M_, 0};
      gemm_large_splitk<double>(ctx_, "vbias_full", W, M_, (int)NN, A, B,
                                EpiStore<double, double>{v, M_, 0, a, 0, sp}, sp == 0);
    } else {
      AFQ_CHECK(prec_ == kMixed, kErrState, "FP32 tensors were not built (precision FP64)");
      Opnd<float, false, false> B{Likn32_ + loff, M_, 0};
      gemm_large_splitk<float>(ctx_, "vbias_full32", W, M_, (int)NN, A, B,
                               EpiStore<float, double>{v, M_, 0, (float)a, 0.f, sp}, sp == 0);
    }
  }
}

// Exchange partial sums of one spin block: part [W][ceil(M / group)] (one entry per walker and
// group of exchange_group(n) Cholesky vectors, summed in group order by energy_finish) and, when
// Kl is given, Kl[w][m] += tr T^m.  Gblk [W][ns][N] contiguous.  The Cholesky vectors are taken
// in chunks whose T [W][ns][MC][ns] fits 1 GiB (max_memory_MB of energy_impl,
// Real3IndexFactorization.hpp:511-587), sized for the larger spin block so both share the buffer.
// The register-tiled products of crgemm.hpp (FP32 and FP64) for vHS and vbias; AFQ_CRGEMM=0
// selects the LDS-staged GEMM engine instead.
bool CholeskyOps::crg_path() {
  const char* e = std::getenv("AFQ_CRGEMM");
  return !(e && e[0] == '0');
}

bool CholeskyOps::exchange_fused_path(Precision prec, bool real) const {
  const char* e = std::getenv("AFQ_EXX_FUSED");  // "0": the GEMM + reduction path
  return prec != kFP64 && real && exx_fused_supported(std::max(nup_, ndown_), N_, M_) &&
         !(e && e[0] == '0');
}

int CholeskyOps::exchange_parts(Precision prec, bool real) const {
  return exchange_fused_path(prec, real)
             ? exx_fused_parts(M_)
             : ceil_div(M_, exchange_group(std::max(std::max(nup_, ndown_), 1)));
}

void CholeskyOps::exchange_spin(int sp, const z64* Gblk, int W, z64* part, z64* Kl,
                                Precision prec, const Rotated& t) const {
  hipStream_t s = ctx_.stream();
  const int ns = sp == 0 ? nup_ : ndown_;
  const int nmax = std::max(nup_, ndown_);
  if (exchange_fused_path(prec, t.real)) {
    AFQ_CHECK(prec_ == kMixed, kErrState, "FP32 tensors were not built (precision FP64)");
    c32* G32 = scratch_as<c32>(ctx_, "en_G32", (size_t)W * ns * N_);
    cast_z64_c32(s, (long)W * ns * N_, Gblk, G32);
    const size_t loff = sp == 0 ? 0 : (size_t)M_ * nup_ * N_;
    exchange_fused(s, W, ns, N_, M_, G32, t.Lnak_r32 + loff, part, exx_fused_parts(M_), Kl);
    return;
  }
  const bool t32 = prec != kFP64;
  const size_t tsz = t32 ? sizeof(c32) : sizeof(z64);
  const int mg = exchange_group(std::max(nmax, 1));
  const long per_m = (long)W * nmax * nmax * (long)tsz;
  int MC = (int)std::max<long>(1, std::min<long>(M_, (1L << 30) / std::max<long>(per_m, 1)));
  MC = std::max(mg, MC / mg * mg);
  const int nchunk = ceil_div(M_, MC);
  const int ng_total = ceil_div(M_, mg);
  void* T = ctx_.scratch("en_T", (size_t)W * nmax * nmax * MC * tsz);
  const size_t loff = sp == 0 ? 0 : (size_t)M_ * nup_ * N_;
  c32* G32 = nullptr;
  for (int ch = 0; ch < nchunk; ++ch) {
    int m0 = ch * MC;
    int mc = std::min(MC, M_ - m0);
    const size_t lc = loff + (size_t)m0 * ns * N_;
    // T[(w,b)][(m,a)] = sum_k G[w][b][k] Lnak[m0+m][a][k]
    if (!t32) {
      EpiStore<double, double> e{static_cast<z64*>(T), (long)MC * ns, 0, 1, 0, 0};
      Opnd<double, true, true> A{&Gblk->re, N_, 0};
      if (t.real) {
        Opnd<double, false, true> B{t.Lnak_r + lc, N_, 0};
        gemm_large<double>(ctx_, "exx_chunk_r", 1, W * ns, mc * ns, N_, A, B, e);
      } else {
        Opnd<double, true, true> B{&(t.Lnak + lc)->re, N_, 0};
        gemm_large<double>(ctx_, "exx_chunk", 1, W * ns, mc * ns, N_, A, B, e);
      }
      exchange_blocks<double>(s, W, ns, mg, MC, m0, mc, M_, static_cast<const z64*>(T), part, ng_total,
                              m0 / mg, Kl);
    } else {
      AFQ_CHECK(prec_ == kMixed, kErrState, "FP32 tensors were not built (precision FP64)");
      if (!G32) {
        G32 = scratch_as<c32>(ctx_, "en_G32", (size_t)W * ns * N_);
        cast_z64_c32(s, (long)W * ns * N_, Gblk, G32);
      }
      EpiStore<float, float> ef{static_cast<c32*>(T), (long)MC * ns, 0, 1.f, 0.f, 0};
      Opnd<float, true, true> A{&G32->re, N_, 0};
      if (t.real) {
        Opnd<float, false, true> B{t.Lnak_r32 + lc, N_, 0};
        gemm_large<float>(ctx_, "exx_chunk_r", 1, W * ns, mc * ns, N_, A, B, ef);
      } else {
        Opnd<float, true, true> B{&(t.Lnak32 + lc)->re, N_, 0};
        gemm_large<float>(ctx_, "exx_chunk_c", 1, W * ns, mc * ns, N_, A, B, ef);
      }
      exchange_blocks<float>(s, W, ns, mg, MC, m0, mc, M_, static_cast<const c32*>(T), part, ng_total,
                             m0 / mg, Kl);
    }
  }
}

void CholeskyOps::energy(const z64* G, int W, z64* E, bool addH1, bool addEJ, bool addEXX,
                         Precision prec, const Rotated* det) const {
  if (W == 0) return;
  AFQ_CHECK(!addEJ || addEXX, kErrArg, "addEJ and not addEXX not yet implemented");
  const Rotated t = det ? *det : self();
  hipStream_t s = ctx_.stream();
  const int nel_ = nel();
  const double c = walker_type_ == kClosed ? 2.0 : 1.0;
  // E = 0; E[:,0] = E0 (energy :156-157)
  fill_z64(s, (long)W * 3, {0, 0}, E);
  if (addH1) {
    // E[w][0] = E0 + c sum G haj over both spin blocks (G rows and haj rows share the order)
    dot_rows(s, W, (long)nel_ * N_, G, t.haj, c, 0.0, E0_, E, 3);
  }
  if (!addEXX) return;
  z64* Kl = scratch_as<z64>(ctx_, "en_Kl", (size_t)W * M_);
  fill_z64(s, (long)W * M_, {0, 0}, Kl);
  const int nspin = walker_type_ == kCollinear ? 2 : 1;
  const int ng_total = exchange_parts(prec, t.real);
  z64* part = scratch_as<z64>(ctx_, "en_part", (size_t)W * ng_total * nspin);
  for (int sp = 0; sp < nspin; ++sp) {
    int ns = sp == 0 ? nup_ : ndown_;
    if (ns == 0) {
      fill_z64(s, (long)W * ng_total, {0, 0}, part + (size_t)sp * W * ng_total);
      continue;
    }
    // contiguous spin block of G: [W][ns][N]
    const z64* Gblk = G;
    if (nspin == 2) {
      z64* Gs = scratch_as<z64>(ctx_, "en_Gs", (size_t)W * ns * N_);
      AFQ_HIP(hipMemcpy2DAsync(Gs, sizeof(z64) * ns * N_, G + (sp == 0 ? 0 : (size_t)nup_ * N_),
                               sizeof(z64) * nel_ * N_, sizeof(z64) * ns * N_, W,
                               hipMemcpyDeviceToDevice, s));
      Gblk = Gs;
    }
    exchange_spin(sp, Gblk, W, part + (size_t)sp * W * ng_total, addEJ ? Kl : nullptr, prec, t);
  }
  // E[w][1] = -1/2 c sum exx;  E[w][2] = 1/2 c^2 sum_m Kl^2  (energy :169-173)
  for (int sp = 0; sp < nspin; ++sp)
    energy_finish(s, W, M_, ng_total, part + (size_t)sp * W * ng_total, Kl, -0.5 * c, 0.5 * c * c,
                  1, (addEJ && sp == nspin - 1) ? 1 : 0, E);
}

void CholeskyOps::energy_spin(int spin, const z64* G, int W, z64* E, z64* EJn, bool addH1,
                              bool addEJ, bool addEXX, Precision prec) const {
  AFQ_CHECK(spin == 0 || (spin == 1 && walker_type_ == kCollinear), kErrArg,
            "spin must be 0 (alpha) or, for COLLINEAR walkers, 1 (beta)");
  AFQ_CHECK(!addEJ || addEXX, kErrArg, "addEJ and not addEXX not yet implemented");
  AFQ_CHECK(!addEJ || EJn, kErrArg, "addEJ needs EJn");
  if (W == 0) return;
  hipStream_t s = ctx_.stream();
  const int ns = spin == 0 ? nup_ : ndown_;
  const double c = walker_type_ == kClosed ? 2.0 : 1.0;
  fill_z64(s, (long)W * 3, {0, 0}, E);
  if (addEJ) fill_z64(s, (long)W * M_, {0, 0}, EJn);
  // by convention E0 goes to the alpha block only (:195-196)
  if (addH1 && spin == 0) fill_blocks_z64(s, W, 1, 3, {E0_, 0.0}, E);
  if (ns == 0) return;  // an empty spin sector contributes nothing (energy_impl :499-500)
  if (addH1)
    dot_rows(s, W, (long)ns * N_, G, haj_ + (spin == 0 ? 0 : (size_t)nup_ * N_), c, 0.0, 0.0, E, 3);
  if (!addEXX) return;
  const int ng_total = exchange_parts(prec, real_trial_);
  z64* part = scratch_as<z64>(ctx_, "ens_part", (size_t)W * ng_total);
  exchange_spin(spin, G, W, part, addEJ ? EJn : nullptr, prec, self());
  // E[w][1] = -1/2 c sum exx;  E[w][2] = 1/2 sum_m EJn^2 (:206-207)
  energy_finish(s, W, M_, ng_total, part, EJn, -0.5 * c, 0.5, 1, addEJ ? 1 : 0, E);
}

void CholeskyOps::vMF(double dt, z64* v) const {
  hipStream_t s = ctx_.stream();
  const int nel_ = nel();
  z64* Psi = scratch_as<z64>(ctx_, "vmf_Psi", (size_t)N_ * std::max(nel_, 1));
  z64* G = scratch_as<z64>(ctx_, "vmf_G", (size_t)nel_ * N_ + 1);
  z64* ov = scratch_as<z64>(ctx_, "vmf_ov", 2);
  fill_z64(s, 1, {0, 0}, ov);
  for (int sp = 0; sp < (walker_type_ == kCollinear ? 2 : 1); ++sp) {
    int ns = sp == 0 ? nup_ : ndown_;
    if (ns == 0) continue;
    const z64* Ps = PsiT_ + (sp == 0 ? 0 : (size_t)nup_ * N_);
    conj_transpose(s, ns, N_, Ps, Psi);
    det_ops::mixed_density_matrix(ctx_, Ps, ns, N_, Psi, (long)N_ * ns, 1,
                                  G + (sp == 0 ? 0 : (size_t)nup_ * N_), (long)nel_ * N_, ov, true);
  }
  vbias(G, 1, dt, v, kFP64);
  zero_imag(s, M_, v);
}

void CholeskyOps::one_body_propagator(const z64* vMF, double dt, z64* P1, bool inverse) const {
  hipStream_t s = ctx_.stream();
  const int N = N_;
  const long NN = (long)N * N;
  z64* shift = scratch_as<z64>(ctx_, "p1_shift", (size_t)nl_ * NN);
  z64* H = scratch_as<z64>(ctx_, "p1_H", NN);
  z64* E = scratch_as<z64>(ctx_, "p1_E", NN);
  z64* F = scratch_as<z64>(ctx_, "p1_F", NN);
  // meanfield shift = vHS(vMF) (includes sqrt(dt)), one block per spin block of L
  vHS(vMF, 1, dt, shift, kFP64);
  for (int sp = 0; sp < p1_spins(); ++sp) {
    // H1_s = dt h_s + dt vexx_s + shift_s (add_one_body_shifts, one_body.hpp:65-85)
    add_one_body(s, N, dt, h1_ + (size_t)(sp % nh_) * NN, vexx_ + (size_t)(sp % nl_) * NN,
                 shift + (size_t)(sp % nl_) * NN, H);
    hermitize_scale(s, N, inverse ? 0.5 : -0.5, H);  // hermitize(H1); H1 *= -1/2 (+1/2: inverse)
    // scaling and squaring: ||H||_1 on the host (setup path)
    std::vector<z64> hH(NN);
    AFQ_HIP(hipMemcpyAsync(hH.data(), H, sizeof(z64) * NN, hipMemcpyDeviceToHost, s));
    AFQ_HIP(hipStreamSynchronize(s));
    double norm1 = 0.0;
    for (int j = 0; j < N; ++j) {
      double cs = 0.0;
      for (int i = 0; i < N; ++i) cs += std::hypot(hH[(size_t)i * N + j].re, hH[(size_t)i * N + j].im);
      norm1 = std::max(norm1, cs);
    }
    int sq = 0;
    while (norm1 / std::ldexp(1.0, sq) > 0.25) ++sq;
    const double scl = std::ldexp(1.0, -sq);
    // Horner: E = I; for k = 20..1: E = I + (B E) / k with B = H / 2^sq
    identity(s, N, E);
    for (int k = 20; k >= 1; --k) {
      identity(s, N, F);
      Opnd<double, true, true> A{&H->re, N, 0};
      Opnd<double, true, false> B{&E->re, N, 0};
      launch_small<double>(s, 1, N, N, N, A, B, EpiStore<double, double>{F, N, 0, scl / k, 0, 1});
      std::swap(E, F);
    }
    for (int i = 0; i < sq; ++i) {
      Opnd<double, true, true> A{&E->re, N, 0};
      Opnd<double, true, false> B{&E->re, N, 0};
      launch_small<double>(s, 1, N, N, N, A, B, EpiStore<double, double>{F, N, 0, 1, 0, 0});
      std::swap(E, F);
    }
    z64* Ps = P1 + (size_t)sp * NN;
    AFQ_HIP(hipMemcpyAsync(Ps, E, sizeof(z64) * NN, hipMemcpyDeviceToDevice, s));
    truncate_small(s, NN, 1e-8, Ps);  // to_csr(P1, 1e-8) (csr_utils.hpp:43-66)
  }
}

}  // namespace afq
