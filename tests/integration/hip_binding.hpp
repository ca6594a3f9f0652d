// The SAFIRE-side binding to libafq_hip.so, as a maintainer adds it to the reference's tree:
//   HipCholeskyOps   a HamiltonianOperations alternative (the member set of
//                    Real3IndexFactorization, HamiltonianOperations.h:33-103)
//   hip_det_ops      the device branches of det_ops::{MixedDensityMatrix, Log_Overlap,
//                    Propagate, orthogonalize} (density_matrix.hpp, propagate.hpp,
//                    orthogonalize.hpp)
//   walkers_view     a WalkerSet's buffer as afq_walkers, in place (WalkerSetBase.h:256-276,
//                    463-481)
// It uses only the SAFIRE names listed in tests/integration/compile_check.cpp (the types and
// accessors it reads) and is compiled against include/afq.h by tests/test_integration_binding.py.
#pragma once

#include <array>
#include <complex>
#include <utility>

#include "afq.h"

namespace sfqmc::afqmc {

namespace hip_detail {
inline void check(int rc) { utils::check(rc == AFQ_OK, afq_last_error()); }
template <class A>
double* d(A&& a) {
  return reinterpret_cast<double*>(const_cast<ComplexType*>(
      reinterpret_cast<const ComplexType*>(a.data())));
}
template <class A>
const double* cd(A const& a) {
  return reinterpret_cast<const double*>(a.data());
}
}  // namespace hip_detail

class HipCholeskyOps {
 public:
  // RealDenseHamiltonian::getHamiltonianOperations (RealDenseHamiltonian.cpp:52-186) with the
  // tensors it reads: Likn [NMO][NMO][nCV] real, hcore [NMO][NMO] real, PsiT [nup+ndown][NMO]
  // (conjugate-transposed trial, alpha rows first), all device arrays.
  template <class L3, class H2, class P2>
  HipCholeskyOps(afq_ctx* ctx, WALKER_TYPES wt, int NMO, int nup, int ndown, L3 const& Likn,
                 H2 const& hcore, P2 const& PsiT, RealType E0, bool mixed)
      : walker_type(wt), NMO(NMO), nup(nup), ndown(ndown), nCV((int)Likn.extent(2)), E0(E0),
        prec(mixed ? AFQ_MIXED : AFQ_FP64) {
    utils::check(wt == CLOSED || wt == COLLINEAR, "HipCholeskyOps: CLOSED or COLLINEAR walkers");
    hip_detail::check(afq_ham_create(ctx, wt == COLLINEAR ? AFQ_COLLINEAR : AFQ_CLOSED, NMO, nup,
                                     ndown, nCV, Likn.data(), hcore.data(),
                                     hip_detail::cd(PsiT), E0, prec, &h));
  }
  HipCholeskyOps(HipCholeskyOps&& o) noexcept { swap(o); }
  HipCholeskyOps& operator=(HipCholeskyOps&& o) noexcept {
    swap(o);
    return *this;
  }
  HipCholeskyOps(HipCholeskyOps const&) = delete;
  ~HipCholeskyOps() {
    if (h) afq_ham_destroy(h);
  }

  // Real3IndexFactorization::vbias (Real3IndexFactorization.hpp:328-428): G [nwalk][nel*NMO]
  template <class G2, class V2>
  void vbias(G2 const& G, V2&& v, double dt) {
    utils::check(v.extent(1) == nCV, "vbias: size mismatch");
    hip_detail::check(afq_vbias(h, hip_detail::cd(G), (int)G.extent(0), dt, prec, hip_detail::d(v)));
  }
  // Real3IndexFactorization::vHS (:282-326): V [1][nwalk][NMO][NMO], complex128 (FP64) -- the
  // mixed build hands the library complex64 X / V buffers through afq_vhs directly
  template <class X2>
  auto vHS(X2&& X, double dt) {
    const long nw = X.extent(0);
    memory::buffered_array<DEVICE_MEMORY, ComplexType, 4> V(std::array<long, 4>{1, nw, NMO, NMO});
    hip_detail::check(afq_vhs(h, X.data(), (int)nw, dt, AFQ_FP64, V.data()));
    return V;
  }
  // Real3IndexFactorization::energy (:138-174), E [nwalk][3]
  template <class E2, class G2>
  void energy(E2&& E, G2 const& G, int idet, bool addH1 = true, bool addEJ = true,
              bool addEXX = true) {
    utils::check(idet == 0, "HipCholeskyOps: one determinant");
    hip_detail::check(afq_energy(h, hip_detail::cd(G), (int)G.extent(0), addH1, addEJ, addEXX, prec,
                                 hip_detail::d(E)));
  }
  // the per-spin energy (:176-210), EJn [nwalk][nCV]
  template <class E2, class G2, class K2>
  void energy(SpinTypes spin, E2&& E, G2 const& G, int idet, K2&& EJn, bool addH1 = true,
              bool addEJ = true, bool addEXX = true) {
    utils::check(idet == 0, "HipCholeskyOps: one determinant");
    hip_detail::check(afq_energy_spin(h, spin == Alpha ? 0 : 1, hip_detail::cd(G),
                                      (int)G.extent(0), addH1, addEJ, addEXX, prec,
                                      hip_detail::d(E), addEJ ? hip_detail::d(EJn) : nullptr));
  }
  // generateP1 (AFQMCBasePropagator.cpp:43-214) takes getOneBodyPropagatorMatrix (:99-124),
  // hermitizes, exponentiates and truncates on the host; the library does the whole chain:
  // P1 [NMO][NMO] = exp(-1/2 H1) with |x| <= 1e-8 dropped
  template <class V1, class P2>
  void one_body_propagator(double dt, V1 const& vMF, P2&& P1) {
    hip_detail::check(afq_one_body_propagator(h, hip_detail::cd(vMF), dt, hip_detail::d(P1)));
  }
  int number_of_cholesky_vectors() const { return nCV; }
  RealType energy_offset() const { return E0; }
  std::array<int, 2> vHS_dims() const { return {NMO * NMO, nCV}; }
  template <class G2>
  void runtime_optimization(G2 const&) {}
  template <class... Args>
  void update_potentials(Args&&...) {}
  afq_ham* handle() const { return h; }

 private:
  afq_ham* h = nullptr;
  WALKER_TYPES walker_type = CLOSED;
  int NMO = 0, nup = 0, ndown = 0, nCV = 0;
  RealType E0 = 0;
  int prec = AFQ_FP64;
  void swap(HipCholeskyOps& o) noexcept {
    std::swap(h, o.h);
    std::swap(walker_type, o.walker_type);
    std::swap(NMO, o.NMO);
    std::swap(nup, o.nup);
    std::swap(ndown, o.ndown);
    std::swap(nCV, o.nCV);
    std::swap(E0, o.E0);
    std::swap(prec, o.prec);
  }
};

// det_ops device branches.  A is PsiT [NEL][NMO]; B is SlaterMatrices(s) [nbatch][NMO][NEL], whose
// walker stride (strides()[0]) is the walker row of the buffer it views.
namespace hip_det_ops {
template <class A2, class B3, class C3, class O1>
void MixedDensityMatrix(afq_ctx* ctx, A2 const& A, B3 const& B, C3&& C, O1&& ovlp, bool compact) {
  hip_detail::check(afq_mixed_density_matrix(ctx, hip_detail::cd(A), (int)A.extent(0),
                                             (int)A.extent(1), hip_detail::cd(B), B.strides()[0],
                                             (int)B.extent(0), hip_detail::d(C), hip_detail::d(ovlp),
                                             compact));
}
template <class A2, class B3, class O1>
void Log_Overlap(afq_ctx* ctx, A2 const& A, B3 const& B, O1&& ovlp) {
  hip_detail::check(afq_log_overlap(ctx, hip_detail::cd(A), (int)A.extent(0), (int)A.extent(1),
                                    hip_detail::cd(B), B.strides()[0], (int)B.extent(0),
                                    hip_detail::d(ovlp)));
}
// propagate_impl<TA> with a dense P1 (the CSR P1 with |x| <= 1e-8 dropped, densified)
template <class P2, class V3, class S3>
void Propagate(afq_ctx* ctx, char TA, P2 const& P1, V3 const& V, S3&& SM, int order) {
  hip_detail::check(afq_propagate(ctx, hip_detail::cd(P1), V.data(), hip_detail::d(SM),
                                  SM.strides()[0], (int)SM.extent(0), (int)SM.extent(1),
                                  (int)SM.extent(2), order, TA, AFQ_FP64));
}
template <class S3, class L1>
void orthogonalize(afq_ctx* ctx, S3&& SM, L1&& log_detR) {
  hip_detail::check(afq_orthogonalize(ctx, hip_detail::d(SM), SM.strides()[0], (int)SM.extent(0),
                                      (int)SM.extent(1), (int)SM.extent(2),
                                      hip_detail::d(log_detR)));
}
}  // namespace hip_det_ops

// The walker set's own buffer as afq_walkers: Slater blocks and property columns in place.
template <class WSet>
afq_walkers walkers_view(WSet& wset, bool keep_energy_components) {
  auto A = wset.SlaterMatrices(Alpha);
  afq_walkers w{};
  w.Phi = hip_detail::d(A);
  w.Phi_b = wset.walkerType() == COLLINEAR ? hip_detail::d(wset.SlaterMatrices(Beta)) : nullptr;
  auto col = [&](walker_data id) { return hip_detail::d(wset.getProperty(id)); };
  w.weight = col(WEIGHT);
  w.phase = col(PHASE);
  w.eloc = col(PSEUDO_ELOC_);
  w.ovlp = col(OVLP);
  w.theta = col(THETA);
  w.W = (int)wset.size();
  w.phi_stride = A.strides()[0];
  w.prop_stride = wset.getProperty(WEIGHT).strides()[0];
  if (keep_energy_components) {
    w.E1 = col(E1_);
    w.EXX = col(EXX_);
    w.EJ = col(EJ_);
  }
  return w;
}

}  // namespace sfqmc::afqmc
