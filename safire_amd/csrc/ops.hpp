// Host-side C++ classes that mirror SAFIRE's interfaces for the walker step:
//   CholeskyOps  ~ Real3IndexFactorization (HamiltonianOperations alternative):
//                  vbias, vHS, energy, getOneBodyPropagatorMatrix, vMF
//   det_ops::*   ~ det_ops::{MixedDensityMatrix, Log_Overlap, Propagate, orthogonalize}
//   Propagator   ~ AFQMCBasePropagator::{Propagate, Orthogonalize} for a CLOSED/COLLINEAR
//                  single-determinant trial, with the whole step on the device.
// All pointers are device pointers; complex data is interleaved (re, im), row-major.
#pragma once

#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "afq_internal.hpp"

namespace afq {

enum WalkerType : int { kClosed = 0, kCollinear = 1 };
enum Precision : int { kFP64 = 0, kMixed = 1 };

// Stream-ordered context: one device, one stream, named scratch buffers that grow on demand.
class Context {
 public:
  Context(int device, hipStream_t stream);
  ~Context();
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  // Scratch buffer `name` of at least `bytes`; contents are not preserved across growth.
  void* scratch(const std::string& name, size_t bytes);
  size_t scratch_bytes() const;
  bool capturing() const;
  // Incremented whenever a scratch buffer is freed to grow it: a HIP graph captured under an
  // older generation may hold the freed address and must be re-captured.
  uint64_t generation() const { return gen_; }
  // While set, scratch() refuses to allocate and the tuner times nothing (it takes the cached
  // choice or the first candidate, which give the same bits): the allocation-free entry points.
  bool no_alloc() const { return no_alloc_; }
  void set_no_alloc(bool on) { no_alloc_ = on; }

 private:
  int device_;
  hipStream_t stream_;
  bool own_stream_;
  bool no_alloc_ = false;
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  std::map<std::string, Buf> bufs_;
  uint64_t gen_ = 0;
};

template <typename T>
T* scratch_as(Context& c, const std::string& name, size_t count) {
  return reinterpret_cast<T*>(c.scratch(name, count * sizeof(T)));
}

// ---------------------------------------------------------------------------------------------
// Slater-determinant operations (density_matrix.hpp, propagate.hpp, orthogonalize.hpp)
// ---------------------------------------------------------------------------------------------
namespace det_ops {

// T = PsiT * Phi_w; ovlp[w] = scale * wrap(ovlp[w] + log det T) + add; optionally T^{-1} in Tinv.
// keep: T, its inverse and raw log det are kept for a later mixed_density_matrix of the same
// walkers (LuKeep, dense_small.hip).
void log_overlap(Context& c, const z64* PsiT, int n, int N, const z64* Phi, long phi_bs, int W,
                 z64* ovlp, double scale = 1.0, z64 add = {0, 0}, const LuKeep* keep = nullptr);
// MixedDensityMatrix: compact G [W][n][N] or full G [W][N][N]; ovlp accumulated as in
// log_overlap.  G32 (compact only): G is computed in FP64 and stored rounded to complex64 at
// G32 (same layout) instead of G.  ws_T [W][n][n] and, for the full G, ws_Gc [W][n][N]: the
// caller's workspace in place of the context's scratch buffers (both or neither).
// keep: walkers whose T equals the kept T bit for bit take the kept inverse and log det (the
// same bits the factorization would give); the others are factorized into the kept buffers.
void mixed_density_matrix(Context& c, const z64* PsiT, int n, int N, const z64* Phi, long phi_bs,
                          int W, z64* G, long g_bs, z64* ovlp, bool compact, double scale = 1.0,
                          c32* G32 = nullptr, z64* ws_T = nullptr, z64* ws_Gc = nullptr,
                          const LuKeep* keep = nullptr);
// inverse_logdet (density_matrix.hpp:82-111)
void inverse_logdet(Context& c, const z64* A, int nb, int n, z64* ovlp, z64* Ainv);
// Full mixed density matrices with a left state per walker (the 3-d MixedDensityMatrix of
// density_matrix.hpp:604-650, compact=false): T = L^H Phi, G[w] = (Phi T^-1 L^H)^T [N][N]
// at G + w * g_bs; L [W][N][n] (l_bs apart), ovlp += log det T.
void mixed_density_matrix_left(Context& c, const z64* L, long l_bs, int n, int N, const z64* Phi,
                               long phi_bs, int W, z64* G, long g_bs, z64* ovlp);
// Reference-major batches b = d W + w (B of them): left state L_b = L + (b / ldiv) l_bs
// (ldiv = W: one left state per reference shared by the walkers; 1: one per batch), walker
// Phi_(b mod W) = Phi + (b mod W) phi_bs.  ovlp[b] += log det(L_b^H Phi); G, when given, [B]
// blocks of g_bs: the full (Phi (L^H Phi)^-1 L^H)^T.
void mixed_density_matrix_batched(Context& c, const z64* L, long l_bs, int ldiv, int n, int N,
                                  const z64* Phi, long phi_bs, int W, int B, z64* G, long g_bs,
                                  z64* ovlp);
// ovlp[w] += log det(L_w^H Phi_w) for left states L [W][N][n] (batch stride l_bs)
void log_overlap_left(Context& c, const z64* L, long l_bs, int n, int N, const z64* Phi,
                      long phi_bs, int W, z64* ovlp);
// Propagate (propagate.hpp:241-298): Phi <- P1 expM(V) P1 Phi, order-K Taylor series.
// P1 [N][N] (real when P1r != nullptr), V [W][N][N] in the compute precision.
struct PropagateArgs {
  int W, N, n, order;
  char TA;                 // 'N' or 'H'
  const z64* P1;           // complex P1 (used when P1r == nullptr)
  const double* P1r;       // real P1
  const float* P1r32;      // real P1, FP32 copy (mixed)
  const c32* P1c32;        // complex P1, FP32 copy (mixed)
  const void* V;           // z64 or c32 [W][N][N]
  long v_bs;
  z64* Phi;                // [W][N][n] (FP64 master copy in both precisions)
  long phi_bs;
  Precision prec;
  bool fused = true;       // one-launch kernel when the shape allows; false: per-product GEMMs
  bool v_symmetric = false;  // V[i][j] == V[j][i] exactly (vHS output): lets 'H' use the fused kernel
};
void propagate(Context& c, const PropagateArgs& a);
// orthogonalize (orthogonalize.hpp:131-169): Phi <- Q with Phi = Q R, R upper triangular with a
// positive real diagonal (Cholesky-QR2); ldet += log det R.
void orthogonalize(Context& c, z64* Phi, long phi_bs, int W, int N, int n, z64* ldet);

}  // namespace det_ops

// ---------------------------------------------------------------------------------------------
// Cholesky (real 3-index) hamiltonian operations
// ---------------------------------------------------------------------------------------------
// Half-rotated tensors of one trial determinant (RealDenseHamiltonian.cpp:141-163,
// half_rotate_hamiltonian): device buffers owned by whoever built them.
struct Rotated {
  z64* PsiT = nullptr;        // [nel][N]
  z64* Lnak = nullptr;        // [M][nup][N], then [M][ndown][N] for COLLINEAR
  z64* haj = nullptr;         // [nel][N]
  double* Lnak_r = nullptr;   // its real part, when the determinant is real
  float* Lnak_r32 = nullptr;  // FP32 real part (mixed precision, real determinant)
  c32* Lnak32 = nullptr;      // FP32 copy (mixed precision, complex determinant)
  bool real = false;
};

class CholeskyOps {
 public:
  // Builds the half-rotated tensors on the device from the bare h1, Likn and the trial
  // (RealDenseHamiltonian::getHamiltonianOperations, RealDenseHamiltonian.cpp:52-186).
  // Likn [nspin_L][N][N][M] real, h1 [nspin_h1][N][N] real (nspin 2: alpha then beta, COLLINEAR
  // only; spin s reads block s % nspin, RealDenseHamiltonian.cpp:86-120), PsiT [nup+ndown][N]
  // complex (alpha rows then beta).
  CholeskyOps(Context& c, int walker_type, int N, int nup, int ndown, int M, const double* Likn,
              const double* h1, const z64* PsiT, double E0, int precision, int nspin_L = 1,
              int nspin_h1 = 1);
  ~CholeskyOps();

  int N() const { return N_; }
  int M() const { return M_; }
  int nup() const { return nup_; }
  int ndown() const { return ndown_; }
  int nel() const { return walker_type_ == kCollinear ? nup_ + ndown_ : nup_; }
  int walker_type() const { return walker_type_; }
  int precision() const { return prec_; }
  double E0() const { return E0_; }
  const z64* PsiT() const { return PsiT_; }
  const z64* Lnak() const { return Lnak_; }
  const z64* haj() const { return haj_; }
  const double* vexx() const { return vexx_; }
  bool real_trial() const { return real_trial_; }
  bool symmetric_L() const { return sym_L_; }
  // spin blocks of the Cholesky tensor and of h1; vHS writes nspin_L() blocks [W][N][N] and the
  // one-body propagator has p1_spins() blocks [N][N] (2 when either depends on spin)
  int nspin_L() const { return nl_; }
  int nspin_h1() const { return nh_; }
  int p1_spins() const { return (nl_ > 1 || nh_ > 1) ? 2 : 1; }

  // v[w][m] = sqrt(dt) c sum_{a,k} G[w][a][k] Lnak[m][a][k]   (vbias, :328-428)
  // prec == kMixed runs the contraction in FP32 and writes FP64 v.
  void vbias(const z64* G, int W, double dt, z64* v, Precision prec) const;
  // vbias of a complex64 copy of G (FP32 MFMA; the same values as vbias(G, ..., kMixed), which
  // rounds G to FP32 while staging it)
  void vbias32(const c32* G, int W, double dt, z64* v) const;
  // V[s][w][i][k] = sqrt(dt) sum_m L_s[i][k][m] X[w][m] for the nspin_L() blocks s (vHS,
  // :282-326, [nstot][nwalk][NMO][NMO]); V is z64 (FP64) or c32
  void vHS(const void* X, int W, double dt, void* V, Precision prec) const;
  // E[w][0..2] = E1, EXX, EJ  (energy :138-174, energy_impl :476-590); G compact [W][nel][N].
  // det: the half-rotated tensors of another determinant of a multi-determinant trial (the
  // reference's idet); null: the hamiltonian's own.
  void energy(const z64* G, int W, z64* E, bool addH1, bool addEJ, bool addEXX,
              Precision prec, const Rotated* det = nullptr) const;
  // per-spin energy (Real3IndexFactorization.hpp:176-210): G [W][n_spin][N] of one block,
  // E [W][3] (E0 in the alpha block only), EJn [W][M] = sum_a T^m_aa (may be null without
  // addEJ), E[w][2] = 1/2 sum_m EJn^2
  void energy_spin(int spin, const z64* G, int W, z64* E, z64* EJn, bool addH1, bool addEJ,
                   bool addEXX, Precision prec) const;
  // vbias of full density matrices (the ndet > 1 branch, Real3IndexFactorization.hpp:395-423):
  // G [W][nspin][N][N], v[w][m] = sqrt(dt) c sum_s sum_ik G[w][s][i][k] L[i][k][m]
  void vbias_full(const z64* G, int W, double dt, z64* v, Precision prec) const;
  // the half-rotated tensors of determinant PsiT [nel][N] (device), in the precision of this
  // hamiltonian; free them with release()
  Rotated rotate(const z64* PsiT) const;
  void release(Rotated& r) const;
  // the hamiltonian's own determinant
  Rotated self() const {
    return Rotated{PsiT_, Lnak_, haj_, Lnak_r_, Lnak_r32_, Lnak32_, real_trial_};
  }
  // vMF (NOMSD.cpp:437-464) with the charge-decomposition zero_imag (AFQMCBasePropagator.cpp:105)
  void vMF(double dt, z64* v) const;
  // P1_s = exp(-1/2 hermitize(dt (h_s + vexx_s) + vHS_s(vMF))) with |x| <= 1e-8 dropped
  // (generateP1, AFQMCBasePropagator.cpp:43-214; add_one_body_shifts, one_body.hpp:65-85).
  // P1 [p1_spins()][N][N] complex.
  // inverse: exp(+1/2 H1) instead, the P1x_inv of generateP1(true) (:118-123)
  void one_body_propagator(const z64* vMF, double dt, z64* P1, bool inverse = false) const;

  Context& ctx() const { return ctx_; }

 private:
  Context& ctx_;
  int walker_type_, N_, nup_, ndown_, M_, prec_;
  int nl_ = 1, nh_ = 1;  // spin blocks of Likn and h1
  double E0_;
  bool real_trial_ = false;
  bool sym_L_ = false;    // every L^m symmetric: vHS computes one triangle and mirrors it
  z64* PsiT_ = nullptr;   // [nel][N]
  double* Likn_ = nullptr;  // [nl][N][N][M]
  float* Likn32_ = nullptr;
  double* h1_ = nullptr;    // [nh][N][N]
  double* vexx_ = nullptr;  // [nl][N][N]
  z64* Lnak_ = nullptr;     // [M][nel][N]: alpha block [M][nup][N] then beta [M][ndown][N]
  double* Lnak_r_ = nullptr;  // real part when the trial is real
  float* Lnak_r32_ = nullptr;
  c32* Lnak32_ = nullptr;
  z64* haj_ = nullptr;      // [nel][N]
  void free_all();
  void exchange_spin(int sp, const z64* Gblk, int W, z64* part, z64* Kl, Precision prec,
                     const Rotated& t) const;
  void half_rotate_into(const z64* PsiT, Rotated& r) const;
  static bool crg_path();
  // the fused on-chip exchange applies (FP32, real half-rotated tensor, n <= 40)
  bool exchange_fused_path(Precision prec, bool real = true) const;
  // exchange partials per walker and spin block of the chosen path
  int exchange_parts(Precision prec, bool real = true) const;
};

// ---------------------------------------------------------------------------------------------
// Multi-determinant NOMSD trial (NOMSD.cpp:39-61, 159-333, 338-520): ndet general determinants
// with coefficients ci; determinant 0 is the hamiltonian's own PsiT.
// ---------------------------------------------------------------------------------------------
class NomsdTrial {
 public:
  // PsiT [ndet][nel][N] (device; determinant 0 equal to the hamiltonian's PsiT), ci [ndet] (host)
  NomsdTrial(CholeskyOps& ham, int ndet, const z64* PsiT, const z64* ci);
  ~NomsdTrial();
  int ndet() const { return ndet_; }
  CholeskyOps& ham() const { return ham_; }
  // walker blocks Pa [W][N][nup], Pb [W][N][ndown] (COLLINEAR), bs_a / bs_b complex elements
  // apart (0: contiguous)
  // NOMSD::Log_Overlap: log sum_d conj(c_d) <D_d|Phi> about the largest term (NOMSD.cpp:295-331)
  void log_overlap(const z64* Pa, const z64* Pb, int W, z64* Ov, long bs_a = 0, long bs_b = 0);
  // NOMSD::MixedDensityMatrix, full G [W][nspin][N][N] = sum_d w_d G_d / sum_d w_d with
  // w_d = conj(c_d) exp(ov_d - log_m) (log_m [W]: the walkers' OVLP); Ov = sum_d w_d + log_m,
  // as the reference leaves it (NOMSD.cpp:240-277)
  void mixed_density_matrix(const z64* Pa, const z64* Pb, int W, const z64* log_m, z64* G,
                            z64* Ov, long bs_a = 0, long bs_b = 0);
  // NOMSD::vbias with ndet > 1 (NOMSD.cpp:39-61): the full mixed density matrix, then the
  // full-G contraction (Real3IndexFactorization.hpp:395-423); v [W][M]
  void vbias(const z64* Pa, const z64* Pb, int W, const z64* log_m, double dt, z64* v,
             Precision prec, long bs_a = 0, long bs_b = 0);
  // NOMSD::Energy with ndet > 1 (NOMSD.cpp:185-216): E [W][3], Ov [W]
  void energy(const z64* Pa, const z64* Pb, int W, z64* E, z64* Ov, Precision prec,
              long bs_a = 0, long bs_b = 0);
  // NOMSD::vMF with ndet > 1 (NOMSD.cpp:438-516) and zero_imag: v [M]
  void vMF(double dt, z64* v);
  // the determinants as references (NOMSD::getReferences): rows_s [ndet][nel_s][N] in PsiT
  // form, cc [ndet] = conj(ci); device buffers of those sizes
  void references(z64* rows0, z64* rows1, z64* cc);

 private:
  CholeskyOps& ham_;
  Context& ctx_;
  int ndet_ = 0;
  std::vector<Rotated> dets_;   // [0]: the hamiltonian's own (not released)
  std::vector<z64> ci_h_;
  z64* cc_ = nullptr;           // device [ndet] conj(ci)
  z64* lcc_ = nullptr;          // device [ndet] log conj(ci), |ci| < 1e-12 as 1e-12
  z64* gmf_ = nullptr;          // device [nspin][N][N] trial density of vMF
  // determinant d against the walkers: log overlaps o [W] (alpha doubled when CLOSED) and, when
  // G is given, the density matrices (compact or full)
  void det_block(int d, const z64* Pa, const z64* Pb, int W, long bs_a, long bs_b, z64* o,
                 z64* G, bool compact);
};

// ---------------------------------------------------------------------------------------------
// Particle-hole multi-determinant trial (PHMSD, COLLINEAR walkers)
// ---------------------------------------------------------------------------------------------

// Host description of the expansion in the reference's ph_excitations numbering
// (Excitations.hpp:203-530, built by readWfn.cpp:238-340): per spin, unique configurations with
// the reference first and the rest ordered by rank; configuration u of rank r >= 1 is 2r ints at
// exc[s][exc_start[s][u]] (hole positions, then particle active orbitals). Determinant d couples
// alpha configuration det_a[d] and beta configuration det_b[d] with coefficient ci[d].
struct PhExcitations {
  int nel[2];
  int nconf[2];
  const int* rank[2];
  const int* exc_start[2];
  const int* exc[2];
  const int* refc[2];
  int ndet;
  const int* det_a;
  const int* det_b;
  const z64* ci;
};

// PHMSD (PHMSD.hpp/.cpp) over a COLLINEAR CholeskyOps whose trial rows are the active orbitals of
// each spin (ham.nup() = nact_alpha, ham.ndown() = nact_beta), as the reference builds its
// HamiltonianOperations from PsiT_MO (Wavefunction.cpp:459-474). The woodbury algorithm
// (refc[i] == i) is the one implemented.
class PhmsdTrial {
 public:
  static constexpr int kMaxRank = 6;
  PhmsdTrial(CholeskyOps& ham, const PhExcitations& ex);
  ~PhmsdTrial();
  int nup() const { return sp_[0].nel; }
  int ndn() const { return sp_[1].nel; }
  CholeskyOps& ham() const { return ham_; }
  // Walker blocks Pa [W][N][nel_a], Pb [W][N][nel_b], consecutive walkers bs_a / bs_b complex
  // elements apart (0: contiguous).
  // compact G [W][nact_a + nact_b][N] and log overlap Ov (PHMSD::MixedDensityMatrix, compact)
  void mixed_density_matrix(const z64* Pa, const z64* Pb, int W, z64* G, z64* Ov, long bs_a = 0,
                            long bs_b = 0);
  // PHMSD::Log_Overlap
  void log_overlap(const z64* Pa, const z64* Pb, int W, z64* Ov, long bs_a = 0, long bs_b = 0);
  // PHMSD::Energy (energy_alg1): E [W][3], Ov [W]
  void energy(const z64* Pa, const z64* Pb, int W, z64* E, z64* Ov, long bs_a = 0, long bs_b = 0);
  // PHMSD::vMF followed by the charge-decomposition zero_imag
  void vMF(double dt, z64* v);
  int ndet() const { return ndet_; }
  int nact(int sp) const { return sp_[sp].nact; }
  const z64* orbitals(int sp) const { return sp_[sp].orbs; }  // device [nact][N]
  // active orbital of column k of determinant d's spin-sp block, [ndet][nel_s]
  std::vector<int> reference_orbitals(int sp) const;
  // The determinants as references (PHMSD::getReferences, PHMSD.hpp:242-279), in PsiT form:
  // rows_s [ndet][nel_s][N] are the trial rows each determinant occupies, in the order of
  // ph_excitations::get_configuration (the reference string with holes replaced by particles);
  // cc [ndet] = conj(ci).  Device buffers of those sizes.
  void references(z64* rows0, z64* rows1, z64* cc);

 private:
  struct SpinDev {
    int nel = 0, nact = 0, nvir = 0, nconf = 0, rmax = 0, P = 0, P2 = 0;
    std::vector<int> rank_off;   // host: configurations of rank r are [rank_off[r], rank_off[r+1])
    std::vector<int> rank_h, exc_h;  // host copies: rank, exc [nconf][2 kMaxRank]
    int* exc = nullptr;          // device [nconf][2 kMaxRank]: holes then particles
    int* rank = nullptr;         // device [nconf]
    int* mstart = nullptr;       // device [nconf] offset of Minv of configuration u (r^2 each)
    int* rstart = nullptr;       // device [nconf] first compact-R row of configuration u
    int* hptr = nullptr;         // device [nel + 1] CSR of excitation rows by hole position
    int* hent = nullptr;         // device int4 entries: (configuration, compact-R row, particle, 0)
    int* cptr = nullptr;         // device [nconf + 1] couplings to the other spin
    int* ccol = nullptr;
    z64* cval = nullptr;         // conj(ci)
    int* kptr = nullptr;         // device [nconf + 1] coupling-chunk range of each row
    long nchunk = 0;
    int hlen_max = 0;            // longest hole list
    z64* Lext = nullptr;         // device [(M + 1)][nact][N]: half-rotated L then haj
    const z64* orbs = nullptr;   // device [nact][N] (rows of the hamiltonian's trial)
  };
  CholeskyOps& ham_;
  Context& ctx_;
  SpinDev sp_[2];
  int ndet_ = 0;
  int* det_ab_ = nullptr;        // device [ndet][2]
  z64* det_c_ = nullptr;         // device [ndet] conj(ci)
  int4* chunks_ = nullptr;       // device coupling chunks (row, begin, end, 0), alpha then beta
  int* crow_ = nullptr;          // device alpha row of each coupling [ndet]
  std::vector<int> det_ab_h_;
  std::vector<z64> det_c_h_;
  // occupied active orbitals of every configuration of spin sp, [nconf][nel]
  std::vector<int> occupations(int sp) const;
  void woodbury(int s, const z64* Phi, long phi_bs, int W, bool want_G, z64* log_ov);
  void ratios(int s, int W, bool want_inverse);
  void weights(int W, z64* Ov);
  void free_all();
};

// One OneRDM / DiagonalTwoRDM measurement (estimators.hip): one [nsp][N][N] = sum_w c_w G_w /
// sum_w weight_w and two [nsp + 1][N][N] = the diagonal two-body blocks (aaaa), (aabb)[, (bbbb)]
// with c_w = weight_w * coeff_w (coeff may be null: 1); either output may be null; accumulate
// adds to the outputs instead.  G [W] blocks of g_bs elements, each [nsp][N][N].  wmod > 0:
// the W blocks are several references' blocks of the same wmod walkers, block b has the weight
// of walker b mod wmod and the denominator sums those wmod weights.
void rdm_measure(Context& c, int W, int nsp, int N, bool collinear, const z64* G, long g_bs,
                 const z64* weight, const z64* coeff, bool accumulate, z64* one, z64* two,
                 int wmod = 0);

// ---------------------------------------------------------------------------------------------
// Propagator: one hybrid step of the whole walker set on the device
// ---------------------------------------------------------------------------------------------
struct PropagatorParams {
  double dt = 0.005;
  int order = 6;
  double vbias_bound = 50.0;
  double lower_cut = 1.0, upper_cut = 10.0;
  double weight_floor = 100.0, weight_frac = 0.1;
  double target_pop = 0.0;  // 0 -> W
  int free_projection = 0, use_cp = 0, project_force_bias = 0;
  uint64_t seed = 0;
  uint32_t stream = 1;
  int local_energy = 0;  // 1: hybrid = false (the local energy drives the weights)
};

struct WalkerState {
  z64* Phi;      // [W][N][nel]: CLOSED [W][N][nup]; COLLINEAR alpha [W][N][nup] then beta
                 //              [W][N][ndown] in separate blocks (Phi_b)
  z64* Phi_b;    // beta block (COLLINEAR) or nullptr
  z64* weight;
  z64* phase;
  z64* eloc;     // PSEUDO_ELOC_
  z64* ovlp;     // OVLP (log overlap)
  z64* theta;
  int W;
  // Walker strides in complex elements.  phi_bs / phi_b_bs: between consecutive walkers' Slater
  // blocks (0: contiguous, N * nup / N * ndown); prop_bs: between consecutive walkers'
  // properties (1: contiguous).  SAFIRE's walker buffer keeps one row per walker
  // (WalkerSetBase.cpp:75-121), so there every stride is the row length (WalkerSetBase.h:256-276,
  // 463-481).
  long phi_bs = 0, phi_b_bs = 0, prop_bs = 1;
  // E1_, EXX_, EJ_ columns (may be null), e_bs apart: the local-energy step (hybrid = false)
  // stores the energy components there (WalkerSetUpdate.hpp:180-184), and population control
  // copies them with their walkers
  z64* e1 = nullptr;
  z64* exx = nullptr;
  z64* ej = nullptr;
  long e_bs = 1;
};

class Propagator {
 public:
  // trial == nullptr: single determinant (the hamiltonian's PsiT rows, NOMSD); otherwise the
  // PHMSD trial built over `ham` supplies G, overlaps, energies and vMF.
  Propagator(CholeskyOps& ham, int W, const PropagatorParams& p, PhmsdTrial* trial = nullptr)
      : Propagator(ham, W, p, trial, nullptr) {}
  // a multi-determinant NOMSD trial over `ham` supplies G, overlaps, energies and vMF
  Propagator(CholeskyOps& ham, int W, const PropagatorParams& p, NomsdTrial* msd)
      : Propagator(ham, W, p, nullptr, msd) {}
  Propagator(CholeskyOps& ham, int W, const PropagatorParams& p, PhmsdTrial* trial,
             NomsdTrial* msd);
  ~Propagator();
  const z64* vMF() const { return vMF_; }
  // P1 [p1_spins][N][N]
  const z64* P1() const { return P1_; }
  // One step (AFQMCBasePropagator::Propagate, AFQMCBasePropagator.cpp:289-392). Uniforms u
  // [W][M] injected when given, else drawn from Philox with the step counter held on the device.
  void step(const WalkerState& s, const double* u, const double* Eshift);
  // Orthogonalize (:735-741): QR of every walker then Log_Overlap refresh of OVLP.
  void orthogonalize(const WalkerState& s);
  // Log_Overlap of the whole set into ovlp (NOMSD.cpp:280-333)
  void log_overlap(const WalkerState& s, z64* ovlp);
  // Energy of the set (NOMSD::Energy) into E [W][3] and ovlp [W]
  void energy(const WalkerState& s, z64* E, z64* ovlp, Precision prec);
  // Device step counter used by in-kernel RNG
  uint32_t* step_counter() { return step_dev_; }
  void set_step(uint32_t step);
  // comb population control (population_control.hpp:216-257); r_dev = injected uniform (device
  // scalar) or nullptr to draw it from Philox keyed by the device step counter
  void population_control(const WalkerState& s, const double* r_dev);
  // 1 when the last population control found a collapsed total weight (synchronizes)
  int status();
  // capture one step (+ optional orthogonalization and population control) in a HIP graph
  // and replay it
  void step_graph(const WalkerState& s, const double* Eshift, bool ortho, bool popctl,
                  double relax, double* energy_out);
  Precision precision() const { return prec_; }
  // HIP events around the step's vHS launch, recorded inside captured steps as well, so that
  // the duration is the kernel's as the step runs it.  Switching re-captures the step graphs.
  void set_vhs_timing(bool on);
  // duration of the last step's vHS in ms (waits for the end event)
  float vhs_ms();
  // bound-trigger counters since creation or the last reset, in the order of
  // AFQMCBasePropagator::printBoundStatistics (AFQMCBasePropagator.h:195-235): force-bias clamp
  // total / upper / lower (per walker and field), local-energy clamp total / upper / lower and
  // weight clamp total / upper / lower (per walker).  Synchronizes the stream.
  void bound_stats(long long out[9], bool reset);

  // ---- back-propagation (AFQMCBasePropagator::BackPropagate, AFQMCBasePropagator.cpp:404-560;
  // BackPropEstimator.hpp).  enable_backprop makes every later step store its fields in a ring
  // of nback slots and its path-restoration factor in a ring of 3 nback slots per walker (both
  // indexed by the device step counter, so steps must draw their own fields); population
  // control moves these rows and the anchor Slater matrices with their walkers.
  void enable_backprop(int nback);
  int backprop_length() const { return bp_n_; }
  // anchor = the current Slater matrices (BackPropEstimator::setAnchor)
  void bp_anchor(const WalkerState& s);
  // Back-propagates the trial through the last nbp stored fields (newest first, the adjoint
  // propagator, QR every ortho steps and at the end) and returns, per walker, the full mixed
  // density matrices between it and the anchor, G [W][nspin][N][N], and the weights times the
  // path-restoration factors of the last restore * nbp steps (restore 0, 1 or 2), w [W].
  void backpropagate(const WalkerState& s, int nbp, int ortho, int restore, z64* G, z64* w);
  // device views of the rings (tests): fields [W][nback][M], factors [W][3 nback]
  const z64* bp_fields() const { return bp_fields_; }
  const z64* bp_factors() const { return bp_wfac_; }

  // ---- density-matrix observables of the estimators, over every reference determinant of
  // the trial (one for NOMSD, the expansion for PHMSD).  onerdm [nspin][N][N] and diag2rdm
  // [nspin + 1][N][N] (unfolded blocks), either may be null; see rdm_measure.
  // MixedEstimator::measure (MixedEstimator.hpp:122-140): the walkers' mixed density matrices.
  void mixed_measure(const WalkerState& s, z64* onerdm, z64* diag2rdm);
  // BackPropEstimator::backPropagate (BackPropEstimator.hpp:186-241): the references
  // back-propagated over the last nbp steps against the anchors, with restored weights.
  void bp_measure(const WalkerState& s, int nbp, int ortho, int restore, z64* onerdm,
                  z64* diag2rdm);
  int number_of_references();

  // ---- back-propagation by operators evolved forward in time (TimeEvolvedBPEstimator.hpp;
  // AFQMCBasePropagator::PropagateOperators, AFQMCBasePropagator.cpp:538-684).  Uses the field
  // and factor rings of enable_backprop.  X, Y, M [W][nspin][N][N] (nspin 1 CLOSED, 2
  // COLLINEAR).  As in the reference (the estimator holds them) they belong to walker slots
  // and population control leaves them where they are; with follow_walkers it copies them
  // with their walkers, so that a copied walker keeps the operators of its own path.
  // setAnchor: X = Y = I, M = 0
  void te_anchor(bool follow_walkers = false);
  // PropagateOperators: replays the nsteps most recent stored fields, oldest first,
  // X <- P1 expM(V) P1 X and Y <- P1inv expM(-V)^H P1inv Y, then regularizes with the current
  // walker matrices S (orthonormal after a QR): Y <- S S^H Y, M += Y^H X, X <- X - S S^H X.
  void te_propagate(const WalkerState& s, int nsteps);
  // constructTimeEvolvedMeasurementInputs + the observables: per reference d,
  // G_d = M + (Ref_d^H X)^T (Gc_d conj(Y)) with Gc_d the compact mixed density matrix of the
  // reference and the walker, coefficients conj(c_d) <Ref_d|Phi> normalized over the
  // references, weights times the factors of the last restore_steps steps.  transpose_m adds
  // M^T instead of M (the orientation under which G_d equals the back-propagated estimator).
  void te_measure(const WalkerState& s, int restore_steps, bool transpose_m, z64* onerdm,
                  z64* diag2rdm);
  // device views (tests): 0 X, 1 Y, 2 M; nullptr before the first te_anchor
  const z64* te_operator(int which) const;
  // P1x_inv = exp(+1/2 H1) [N][N], built on first use
  const z64* P1_inverse();

 private:
  CholeskyOps& ham_;
  Context& ctx_;
  PhmsdTrial* trial_;
  NomsdTrial* msd_;
  int W_;
  int wup_, wdn_;   // electrons per spin of the walkers (columns of Phi, Phi_b)
  bool closed_;
  PropagatorParams p_;
  Precision prec_;
  z64* vMF_ = nullptr;
  z64* P1_ = nullptr;       // [np][N][N], np = ham.p1_spins()
  double* P1r_ = nullptr;
  float* P1r32_ = nullptr;
  c32* P1c32_ = nullptr;
  bool real_P1_ = false;
  // single determinant: T = PsiT Phi of each spin block, its inverse and raw log det as the
  // end-of-step overlap leaves them, for the next step's density matrix (which takes them for
  // the walkers whose T comes out bit for bit the same, and recomputes the others)
  LuKeep keep_[2] = {};
  const LuKeep* keep(int sp, int W) const {
    return keep_[sp].T && W <= W_ ? &keep_[sp] : nullptr;
  }
  // the propagation of spin block sp: its P1 block (the inverse one with `inverse`) and its V
  // block (V [nspin_L][W][N][N] in precision prec)
  det_ops::PropagateArgs spin_args(int sp, int W, int n, char ta, const void* V, Precision prec,
                                   z64* Phi, long phi_bs, bool inverse = false) const;
  uint32_t* step_dev_ = nullptr;
  int* status_dev_ = nullptr;
  unsigned long long* stats_dev_ = nullptr;  // [9] bound-trigger counters
  // graph cache, one executable graph per (walker buffers, ortho, relaxation) key
  struct GraphEntry {
    std::vector<const void*> key;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs_;
  uint64_t graph_gen_ = 0;  // context scratch generation the cached graphs were captured under
  // keys whose body ran eagerly once (every scratch buffer it needs exists, its tiles are tuned)
  std::vector<std::vector<const void*>> warmed_;
  bool vhs_timing_ = false;
  hipEvent_t vhs_ev_[2] = {nullptr, nullptr};
  int bp_n_ = 0, bp_hl_ = 0;
  z64* bp_fields_ = nullptr;
  z64* bp_wfac_ = nullptr;
  z64* bp_anch_a_ = nullptr;
  z64* bp_anch_b_ = nullptr;
  // reference determinants in PsiT form: ref_rows_[s] [nref][nel_s][N], ref_cc_ [nref] conj(ci)
  int nref_ = 0;
  z64* ref_rows_[2] = {nullptr, nullptr};
  z64* ref_cc_ = nullptr;
  // PHMSD: selection matrices [nact_s][nref n_s], column d n_s + k = unit vector of the active
  // orbital in column k of reference d (the references as combinations of the active orbitals)
  z64* ref_sel_[2] = {nullptr, nullptr};
  // the references as Slater matrices, [nref][N][nel_s] (the conjugate transposes of ref_rows_)
  z64* ref_cols_[2] = {nullptr, nullptr};
  // time-evolved operators and the inverse one-body propagator
  z64* te_op_[3] = {nullptr, nullptr, nullptr};
  bool te_follow_ = false;
  z64* P1inv_ = nullptr;
  double* P1invr_ = nullptr;
  void ensure_references();
  // back-propagates every reference through the last nbp stored fields; on return Lp[s] holds
  // [W][nref][N][nel_s] and ldR [W][nref] the accumulated log det R of their QR steps
  void bp_propagate_refs(const WalkerState& s, int nbp, int ortho, z64* Lp[2], z64* ldR);
  void bp_propagate_orbitals(const WalkerState& s, uint32_t C, int nbp, int ortho, z64* Lp[2],
                             z64* ldR);
  // ov_scale: 2 where the estimator doubles the CLOSED overlaps (back-propagation), else 1
  void measure_refs(int W, const z64* const Lp[2], bool per_walker, const z64* const Phi[2],
                    const z64* weight, const z64* ldR, z64* onerdm, z64* diag2rdm,
                    double ov_scale);
  void clear_graphs();
  void step_impl(const WalkerState& s, const double* u, const double* Eshift);
  void orthogonalize_impl(const WalkerState& s);
  void population_control_impl(const WalkerState& s, const double* r_dev);
  // walker strides of a state (the contiguous defaults filled in)
  long bs_a(const WalkerState& s) const { return s.phi_bs ? s.phi_bs : (long)ham_.N() * wup_; }
  long bs_b(const WalkerState& s) const { return s.phi_b_bs ? s.phi_b_bs : (long)ham_.N() * wdn_; }
  void check_strides(const WalkerState& s) const;
  // The state with its properties in contiguous scratch (gathered from strided columns; the
  // state itself when they are contiguous) and the scatter back after a call that updated them.
  WalkerState stage_properties(const WalkerState& s);
  void unstage_properties(const WalkerState& s, const WalkerState& staged);
  // A fully contiguous copy of a strided state (Slater blocks and properties) for the read-only
  // estimator calls; the state itself when it is contiguous.
  WalkerState packed(const WalkerState& s);
};

}  // namespace afq
