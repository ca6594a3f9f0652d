// The walker step on the device: AFQMCBasePropagator::Propagate (AFQMCBasePropagator.cpp:289-392)
// and Orthogonalize (:735-741) for a single-determinant trial (NOMSD, CLOSED or COLLINEAR).
//
//   G, log ovlp    MixedDensityMatrix         (NOMSD::vbias, NOMSD.cpp:39-61)
//   v              vbias                      (Real3IndexFactorization.hpp:328-428)
//   X, HW, mf      construct_X + mean field   (AFQMCBasePropagator.cpp:324-336, 686-733)
//   V              vHS                        (:349)
//   Phi            P1 expM(V) P1 Phi          (apply_propagators, :354)
//   new ovlp       Log_Overlap                (:371)
//   weights        walker_update + bound      (:380-386, WalkerSetUpdate.hpp)
//
// Nothing crosses to the host inside a step, so a step (with or without the periodic QR and
// the Eshift relaxation) is captured once in a HIP graph and replayed.  In the mixed precision
// mode vbias, vHS and the Taylor propagation run in FP32 on the matrix cores; G, overlaps, the
// LU/QR, fields and weights stay FP64, and G is recomputed in FP64 from the FP64 walker matrices
// after every propagation.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "gemm.hpp"
#include "ops.hpp"
#include "tuner.hpp"

namespace afq {

namespace {

template <typename T>
T* dalloc(size_t count) {
  void* p = nullptr;
  if (hipMalloc(&p, (count ? count : 1) * sizeof(T)) != hipSuccess)
    throw Error(kErrAlloc, "device allocation failed");
  return reinterpret_cast<T*>(p);
}

}  // namespace

Propagator::Propagator(CholeskyOps& ham, int W, const PropagatorParams& p, PhmsdTrial* trial,
                       NomsdTrial* msd)
    : ham_(ham), ctx_(ham.ctx()), trial_(trial), msd_(msd), W_(W),
      wup_(trial ? trial->nup() : ham.nup()),
      wdn_(trial ? trial->ndn() : ham.ndown()),
      closed_(ham.walker_type() == kClosed), p_(p), prec_((Precision)ham.precision()) {
  AFQ_CHECK(W > 0, kErrArg, "W must be positive");
  AFQ_CHECK(!trial || &trial->ham() == &ham, kErrArg, "the trial belongs to another hamiltonian");
  AFQ_CHECK(!msd || &msd->ham() == &ham, kErrArg, "the trial belongs to another hamiltonian");
  AFQ_CHECK(!(trial && msd), kErrArg, "one trial");
  AFQ_CHECK(p.order >= 0, kErrArg, "order");
  if (p_.target_pop <= 0) p_.target_pop = W;
  hipStream_t s = ctx_.stream();
  const int N = ham.N(), M = ham.M();
  const long NP = (long)ham.p1_spins() * N * N;  // every one-body propagator block
  try {
    vMF_ = dalloc<z64>(M);
    P1_ = dalloc<z64>(NP);
    step_dev_ = dalloc<uint32_t>(1);
    AFQ_HIP(hipMemsetAsync(step_dev_, 0, sizeof(uint32_t), s));
    status_dev_ = dalloc<int>(1);
    AFQ_HIP(hipMemsetAsync(status_dev_, 0, sizeof(int), s));
    stats_dev_ = dalloc<unsigned long long>(9);
    AFQ_HIP(hipMemsetAsync(stats_dev_, 0, 9 * sizeof(unsigned long long), s));
    // generateP1: vMF (subtractMF = true), then H1 and P1
    if (trial_)
      trial_->vMF(p_.dt, vMF_);
    else if (msd_)
      msd_->vMF(p_.dt, vMF_);
    else
      ham.vMF(p_.dt, vMF_);
    ham.one_body_propagator(vMF_, p_.dt, P1_);
    std::vector<z64> h(NP);
    AFQ_HIP(hipMemcpyAsync(h.data(), P1_, sizeof(z64) * h.size(), hipMemcpyDeviceToHost, s));
    AFQ_HIP(hipStreamSynchronize(s));
    real_P1_ = true;
    for (auto& v : h)
      if (v.im != 0.0) { real_P1_ = false; break; }
    if (real_P1_) {
      P1r_ = dalloc<double>(NP);
      real_part(s, NP, P1_, P1r_);
      if (prec_ == kMixed) {
        P1r32_ = dalloc<float>(NP);
        real_part_f32(s, NP, P1_, P1r32_);
      }
    } else if (prec_ == kMixed) {
      P1c32_ = dalloc<c32>(NP);
      cast_z64_c32(s, NP, P1_, P1c32_);
    }
    // kept overlap factors (AFQ_LU_KEEP=0 switches them off): T starts as NaN, which matches
    // nothing
    const char* ke = std::getenv("AFQ_LU_KEEP");
    if (!trial_ && !msd_ && !(ke && ke[0] == '0')) {
      for (int sp = 0; sp < (closed_ ? 1 : 2); ++sp) {
        const int n = sp == 0 ? wup_ : wdn_;
        if (n < 1 || n > 48) continue;
        const size_t nn = (size_t)W * n * n;
        keep_[sp].T = dalloc<z64>(nn);
        keep_[sp].Tinv = dalloc<z64>(nn);
        keep_[sp].ldet = dalloc<z64>(W);
        fill_z64(s, (long)nn, {std::nan(""), std::nan("")}, keep_[sp].T);
      }
    }
    AFQ_HIP(hipStreamSynchronize(s));
  } catch (...) {
    for (void* q : {(void*)vMF_, (void*)P1_, (void*)P1r_, (void*)P1r32_, (void*)P1c32_,
                    (void*)step_dev_, (void*)status_dev_, (void*)stats_dev_, (void*)keep_[0].T,
                    (void*)keep_[0].Tinv, (void*)keep_[0].ldet, (void*)keep_[1].T,
                    (void*)keep_[1].Tinv, (void*)keep_[1].ldet})
      if (q) (void)hipFree(q);
    throw;
  }
}

Propagator::~Propagator() {
  (void)hipStreamSynchronize(ctx_.stream());
  for (hipEvent_t e : vhs_ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& g : graphs_) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
  }
  for (void* q : {(void*)vMF_, (void*)P1_, (void*)P1r_, (void*)P1r32_, (void*)P1c32_,
                  (void*)step_dev_, (void*)status_dev_, (void*)stats_dev_, (void*)bp_fields_,
                  (void*)bp_wfac_,
                  (void*)bp_anch_a_, (void*)bp_anch_b_, (void*)ref_rows_[0], (void*)ref_rows_[1],
                  (void*)ref_cc_, (void*)ref_sel_[0], (void*)ref_sel_[1], (void*)ref_cols_[0],
                  (void*)ref_cols_[1], (void*)te_op_[0], (void*)te_op_[1], (void*)te_op_[2],
                  (void*)P1inv_, (void*)P1invr_, (void*)keep_[0].T, (void*)keep_[0].Tinv,
                  (void*)keep_[0].ldet, (void*)keep_[1].T, (void*)keep_[1].Tinv,
                  (void*)keep_[1].ldet})
    if (q) (void)hipFree(q);
}

void Propagator::set_step(uint32_t step) {
  AFQ_HIP(hipMemcpyAsync(step_dev_, &step, sizeof(uint32_t), hipMemcpyHostToDevice, ctx_.stream()));
  AFQ_HIP(hipStreamSynchronize(ctx_.stream()));
}

void Propagator::check_strides(const WalkerState& st) const {
  const long N = ham_.N();
  AFQ_CHECK(st.phi_bs == 0 || st.phi_bs >= N * wup_, kErrShape,
            "walker stride of the alpha block is below N * nup");
  AFQ_CHECK(st.phi_b_bs == 0 || st.phi_b_bs >= N * wdn_, kErrShape,
            "walker stride of the beta block is below N * ndown");
  AFQ_CHECK(st.prop_bs >= 1, kErrShape, "property stride must be positive");
}

WalkerState Propagator::stage_properties(const WalkerState& st) {
  if (st.prop_bs == 1) return st;
  z64* pk = scratch_as<z64>(ctx_, "wk_props", (size_t)5 * st.W);
  ColumnSet cols;
  z64* src[5] = {st.weight, st.phase, st.eloc, st.ovlp, st.theta};
  for (int k = 0; k < 5; ++k) cols.p[k] = src[k];
  copy_columns(ctx_.stream(), 5, st.W, st.prop_bs, cols, pk, true);
  WalkerState o = st;
  o.weight = pk;
  o.phase = pk + st.W;
  o.eloc = pk + 2 * st.W;
  o.ovlp = pk + 3 * st.W;
  o.theta = pk + 4 * st.W;
  o.prop_bs = 1;
  return o;
}

void Propagator::unstage_properties(const WalkerState& st, const WalkerState& staged) {
  if (st.prop_bs == 1) return;
  ColumnSet cols;
  z64* dst[5] = {st.weight, st.phase, st.eloc, st.ovlp, st.theta};
  for (int k = 0; k < 5; ++k) cols.p[k] = dst[k];
  copy_columns(ctx_.stream(), 5, st.W, st.prop_bs, cols, staged.weight, false);
}

WalkerState Propagator::packed(const WalkerState& st) {
  check_strides(st);
  const long N = ham_.N();
  const bool coll = !closed_ && wdn_ > 0;
  const bool phi_c = bs_a(st) == N * wup_ && (!coll || bs_b(st) == N * wdn_);
  WalkerState o = stage_properties(st);
  if (phi_c) return o;
  hipStream_t s = ctx_.stream();
  z64* a = scratch_as<z64>(ctx_, "wk_phi_a", (size_t)st.W * N * std::max(wup_, 1));
  AFQ_HIP(hipMemcpy2DAsync(a, sizeof(z64) * N * wup_, st.Phi, sizeof(z64) * bs_a(st),
                           sizeof(z64) * N * wup_, st.W, hipMemcpyDeviceToDevice, s));
  o.Phi = a;
  o.phi_bs = 0;
  if (coll) {
    z64* b = scratch_as<z64>(ctx_, "wk_phi_b", (size_t)st.W * N * wdn_);
    AFQ_HIP(hipMemcpy2DAsync(b, sizeof(z64) * N * wdn_, st.Phi_b, sizeof(z64) * bs_b(st),
                             sizeof(z64) * N * wdn_, st.W, hipMemcpyDeviceToDevice, s));
    o.Phi_b = b;
    o.phi_b_bs = 0;
  }
  return o;
}

void Propagator::log_overlap(const WalkerState& st, z64* ovlp) {
  if (trial_) {
    trial_->log_overlap(st.Phi, st.Phi_b, st.W, ovlp, bs_a(st), bs_b(st));
    return;
  }
  if (msd_) {
    msd_->log_overlap(st.Phi, st.Phi_b, st.W, ovlp, bs_a(st), bs_b(st));
    return;
  }
  hipStream_t s = ctx_.stream();
  const int N = ham_.N(), nup = ham_.nup(), ndown = ham_.ndown();
  const bool closed = closed_;
  fill_z64(s, st.W, {0, 0}, ovlp);
  det_ops::log_overlap(ctx_, ham_.PsiT(), nup, N, st.Phi, bs_a(st), st.W, ovlp,
                       closed ? 2.0 : 1.0, {0, 0}, keep(0, st.W));
  if (!closed && ndown > 0)
    det_ops::log_overlap(ctx_, ham_.PsiT() + (size_t)nup * N, ndown, N, st.Phi_b, bs_b(st),
                         st.W, ovlp, 1.0, {0, 0}, keep(1, st.W));
  // + log(conj(ci0)) = 0 for the single determinant with unit coefficient
}

void Propagator::step_impl(const WalkerState& st, const double* u, const double* Eshift) {
  hipStream_t s = ctx_.stream();
  const int W = st.W, N = ham_.N(), M = ham_.M(), nup = wup_, ndown = wdn_;
  const int nel = ham_.nel();  // rows of G: electrons (NOMSD) or active orbitals (PHMSD)
  const bool closed = closed_;
  AFQ_CHECK(W <= W_, kErrShape, "walker count exceeds the propagator's capacity");
  z64* G = scratch_as<z64>(ctx_, "st_G", (size_t)W * nel * N);
  z64* ov = scratch_as<z64>(ctx_, "st_ov", W);
  z64* X = scratch_as<z64>(ctx_, "st_X", (size_t)W * M);
  c32* X32 = prec_ == kMixed ? scratch_as<c32>(ctx_, "st_X32", (size_t)W * M) : nullptr;
  z64* HW = scratch_as<z64>(ctx_, "st_HW", W);
  z64* mf = scratch_as<z64>(ctx_, "st_mf", W);
  z64* nov = scratch_as<z64>(ctx_, "st_nov", W);
  void* V = ctx_.scratch("st_V", (size_t)ham_.nspin_L() * W * N * N *
                                     (prec_ == kMixed ? sizeof(c32) : sizeof(z64)));

  // 1. G and vbias (NOMSD::vbias, NOMSD.cpp:39-61; PHMSD::vbias, PHMSD.cpp:53-72).  In mixed
  // precision G is computed in FP64 and stored as complex64 for the FP32 vbias contraction (the
  // values the contraction would round it to anyway; half the bytes it re-reads).
  if (trial_) {
    trial_->mixed_density_matrix(st.Phi, st.Phi_b, W, G, ov, bs_a(st), bs_b(st));
    ham_.vbias(G, W, p_.dt, X, prec_);
  } else if (msd_) {
    // full mixed density matrix of the expansion, scaled by the walkers' OVLP (NOMSD.cpp:251)
    msd_->vbias(st.Phi, st.Phi_b, W, st.ovlp, p_.dt, X, prec_, bs_a(st), bs_b(st));
  } else {
    c32* G32 = prec_ == kMixed ? scratch_as<c32>(ctx_, "st_G32", (size_t)W * nel * N) : nullptr;
    fill_z64(s, W, {0, 0}, ov);
    det_ops::mixed_density_matrix(ctx_, ham_.PsiT(), nup, N, st.Phi, bs_a(st), W, G,
                                  (long)nel * N, ov, true, closed ? 2.0 : 1.0, G32, nullptr,
                                  nullptr, keep(0, W));
    if (!closed && ndown > 0)
      det_ops::mixed_density_matrix(ctx_, ham_.PsiT() + (size_t)nup * N, ndown, N, st.Phi_b,
                                    bs_b(st), W, G + (size_t)nup * N, (long)nel * N, ov,
                                    true, 1.0, G32 ? G32 + (size_t)nup * N : nullptr, nullptr,
                                    nullptr, keep(1, W));
    if (G32)
      ham_.vbias32(G32, W, p_.dt, X);
    else
      ham_.vbias(G, W, p_.dt, X, prec_);
  }
  // 2. fields, hybrid weight, mean-field factor
  FieldParams fp{p_.vbias_bound, (p_.use_cp || p_.project_force_bias) ? 1 : 0, p_.free_projection,
                 p_.seed, p_.stream, 0, u ? nullptr : step_dev_, stats_dev_};
  construct_x(s, u, vMF_, W, M, fp, X, HW, mf, X32);
  if (bp_n_ > 0) store_fields(s, W, M, X, bp_fields_, bp_n_, step_dev_);
  // 3. vHS
  if (vhs_timing_) AFQ_HIP(hipEventRecord(vhs_ev_[0], s));
  ham_.vHS(prec_ == kMixed ? (const void*)X32 : (const void*)X, W, p_.dt, V, prec_);
  if (vhs_timing_) AFQ_HIP(hipEventRecord(vhs_ev_[1], s));
  // 4. propagate both spin blocks, each with its P1 and V blocks (one shared V and P1 for a
  // spin-independent hamiltonian)
  det_ops::PropagateArgs pa = spin_args(0, W, nup, 'N', V, prec_, st.Phi, bs_a(st));
  det_ops::propagate(ctx_, pa);
  if (!closed && ndown > 0) {
    pa = spin_args(1, W, ndown, 'N', V, prec_, st.Phi_b, bs_b(st));
    det_ops::propagate(ctx_, pa);
  }
  // 5. new overlaps (hybrid), or the local energies and overlaps (hybrid = false)
  z64* En = nullptr;
  if (p_.local_energy) {
    En = scratch_as<z64>(ctx_, "st_E", (size_t)W * 3);
    energy(st, En, nov, prec_);
  } else {
    log_overlap(st, nov);
  }
  // 6. weights
  UpdateParams up{p_.dt, ham_.E0(), p_.lower_cut, p_.upper_cut, p_.weight_floor, p_.weight_frac,
                  p_.target_pop, p_.free_projection, p_.use_cp, p_.local_energy, stats_dev_};
  walker_update(s, W, up, Eshift, nov, mf, HW, st.weight, st.phase, st.eloc, st.ovlp, st.theta,
                bp_n_ > 0 ? bp_wfac_ : nullptr, bp_hl_, step_dev_, En);
  if (En) store_energy_components(s, W, En, st.e1, st.exx, st.ej, st.e_bs);
}

det_ops::PropagateArgs Propagator::spin_args(int sp, int W, int n, char ta, const void* V,
                                             Precision prec, z64* Phi, long phi_bs,
                                             bool inverse) const {
  const int N = ham_.N();
  const long NN = (long)N * N;
  const long p1 = (long)(sp % ham_.p1_spins()) * NN;
  const size_t esz = prec == kFP64 ? sizeof(z64) : sizeof(c32);
  const void* Vs = reinterpret_cast<const char*>(V) + (size_t)(sp % ham_.nspin_L()) * W * NN * esz;
  const z64* P = inverse ? P1inv_ : P1_;
  const double* Pr = inverse ? P1invr_ : P1r_;
  det_ops::PropagateArgs pa{W, N, n, p_.order, ta, P + p1, real_P1_ ? Pr + p1 : nullptr,
                            (prec != kFP64 && !inverse && P1r32_) ? P1r32_ + p1 : nullptr,
                            (prec != kFP64 && !inverse && P1c32_) ? P1c32_ + p1 : nullptr, Vs,
                            NN, Phi, phi_bs,
                            prec};
  return pa;
}

void Propagator::bound_stats(long long out[9], bool reset) {
  unsigned long long h[9];
  AFQ_HIP(hipMemcpyAsync(h, stats_dev_, sizeof(h), hipMemcpyDeviceToHost, ctx_.stream()));
  AFQ_HIP(hipStreamSynchronize(ctx_.stream()));
  for (int i = 0; i < 9; ++i) out[i] = (long long)h[i];
  if (reset) {
    AFQ_HIP(hipMemsetAsync(stats_dev_, 0, sizeof(h), ctx_.stream()));
    AFQ_HIP(hipStreamSynchronize(ctx_.stream()));
  }
}

void Propagator::set_vhs_timing(bool on) {
  if (on && !vhs_ev_[0]) {
    // no system-scope fence at the records: the fence's cache write-back would be timed with
    // the kernel
    AFQ_HIP(hipEventCreateWithFlags(&vhs_ev_[0], hipEventDisableSystemFence));
    AFQ_HIP(hipEventCreateWithFlags(&vhs_ev_[1], hipEventDisableSystemFence));
  }
  if (on != vhs_timing_) {
    AFQ_HIP(hipStreamSynchronize(ctx_.stream()));
    clear_graphs();
    vhs_timing_ = on;
  }
}

float Propagator::vhs_ms() {
  AFQ_CHECK(vhs_ev_[0], kErrState, "vHS timing was never switched on");
  AFQ_HIP(hipEventSynchronize(vhs_ev_[1]));
  float ms = 0.0f;
  AFQ_HIP(hipEventElapsedTime(&ms, vhs_ev_[0], vhs_ev_[1]));
  return ms;
}

void Propagator::step(const WalkerState& st0, const double* u, const double* Eshift) {
  check_strides(st0);
  const WalkerState st = stage_properties(st0);
  step_impl(st, u, Eshift);
  if (!u) increment_counter(ctx_.stream(), step_dev_);
  unstage_properties(st0, st);
}

void Propagator::orthogonalize(const WalkerState& st0) {
  check_strides(st0);
  const WalkerState st = stage_properties(st0);
  orthogonalize_impl(st);
  unstage_properties(st0, st);
}

void Propagator::orthogonalize_impl(const WalkerState& st) {
  hipStream_t s = ctx_.stream();
  const int N = ham_.N(), nup = wup_, ndown = wdn_;
  z64* ldet = scratch_as<z64>(ctx_, "or_ldet", st.W);
  fill_z64(s, st.W, {0, 0}, ldet);
  det_ops::orthogonalize(ctx_, st.Phi, bs_a(st), st.W, N, nup, ldet);
  if (!closed_ && ndown > 0)
    det_ops::orthogonalize(ctx_, st.Phi_b, bs_b(st), st.W, N, ndown, ldet);
  log_overlap(st, st.ovlp);
}

void Propagator::population_control(const WalkerState& st0, const double* r_dev) {
  check_strides(st0);
  const WalkerState st = stage_properties(st0);
  population_control_impl(st, r_dev);
  unstage_properties(st0, st);
}

void Propagator::population_control_impl(const WalkerState& st, const double* r_dev) {
  const int N = ham_.N(), nup = wup_, ndown = wdn_;
  const bool coll = !closed_ && ndown > 0;
  double* cum = scratch_as<double>(ctx_, "pc_cum", st.W);
  int* src = scratch_as<int>(ctx_, "pc_src", st.W);
  WalkerRows extra;
  if (bp_n_ > 0) {
    extra.p[0] = bp_fields_;
    extra.n[0] = (long)bp_n_ * ham_.M();
    extra.p[1] = bp_wfac_;
    extra.n[1] = bp_hl_;
    extra.p[2] = bp_anch_a_;
    extra.n[2] = (long)N * nup;
    if (coll) {
      extra.p[3] = bp_anch_b_;
      extra.n[3] = (long)N * ndown;
    }
  }
  if (te_follow_ && te_op_[0]) {  // the time-evolved operators move with their walkers
    const long rowte = (long)(closed_ ? 1 : 2) * N * N;
    for (int i = 0; i < 3; ++i) {
      extra.p[4 + i] = te_op_[i];
      extra.n[4 + i] = rowte;
    }
  }
  z64* ecol[3] = {st.e1, st.exx, st.ej};
  for (int i = 0; i < 3; ++i)
    if (ecol[i]) {
      extra.p[7 + i] = ecol[i];
      extra.n[7 + i] = 1;
      extra.bs[7 + i] = st.e_bs;
    }
  afq::population_control(ctx_.stream(), st.W, r_dev, p_.seed, p_.stream + 1000u, step_dev_,
                          (long)N * nup, coll ? (long)N * ndown : 0, bs_a(st), coll ? bs_b(st) : 0,
                          st.Phi, coll ? st.Phi_b : nullptr, st.weight, st.phase, st.eloc, st.ovlp,
                          st.theta, cum, src, status_dev_, extra);
}

void Propagator::clear_graphs() {
  for (auto& g : graphs_) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
  }
  graphs_.clear();
  warmed_.clear();
}

void Propagator::enable_backprop(int nback) {
  AFQ_CHECK(nback > 0, kErrArg, "back-propagation length must be positive");
  AFQ_CHECK(bp_n_ == 0, kErrState, "back-propagation is already enabled");
  hipStream_t s = ctx_.stream();
  const int N = ham_.N(), M = ham_.M();
  z64* buf[4] = {nullptr, nullptr, nullptr, nullptr};
  const size_t cnt[4] = {(size_t)W_ * nback * M, (size_t)W_ * 3 * nback,
                         (size_t)W_ * N * std::max(wup_, 1), (size_t)W_ * N * std::max(wdn_, 1)};
  try {
    for (int i = 0; i < 4; ++i) buf[i] = dalloc<z64>(cnt[i]);
  } catch (...) {
    for (z64* q : buf)
      if (q) (void)hipFree(q);
    throw;
  }
  bp_fields_ = buf[0];
  bp_wfac_ = buf[1];
  bp_anch_a_ = buf[2];
  bp_anch_b_ = buf[3];
  bp_n_ = nback;
  bp_hl_ = 3 * nback;
  // an unwritten history entry restores nothing (WalkerSetBase::resize_bp initializes it to 1)
  fill_z64(s, (long)cnt[1], {1.0, 0.0}, bp_wfac_);
  fill_z64(s, (long)cnt[0], {0.0, 0.0}, bp_fields_);
  AFQ_HIP(hipStreamSynchronize(s));
  clear_graphs();  // captured steps do not store fields yet
}

void Propagator::bp_anchor(const WalkerState& st) {
  AFQ_CHECK(bp_n_ > 0, kErrState, "back-propagation is not enabled");
  check_strides(st);
  hipStream_t s = ctx_.stream();
  const long N = ham_.N();
  AFQ_HIP(hipMemcpy2DAsync(bp_anch_a_, sizeof(z64) * N * wup_, st.Phi, sizeof(z64) * bs_a(st),
                           sizeof(z64) * N * wup_, st.W, hipMemcpyDeviceToDevice, s));
  if (!closed_ && wdn_ > 0)
    AFQ_HIP(hipMemcpy2DAsync(bp_anch_b_, sizeof(z64) * N * wdn_, st.Phi_b, sizeof(z64) * bs_b(st),
                             sizeof(z64) * N * wdn_, st.W, hipMemcpyDeviceToDevice, s));
}

void Propagator::ensure_references() {
  if (nref_ > 0) return;
  hipStream_t s = ctx_.stream();
  const int N = ham_.N();
  const int R = trial_ ? trial_->ndet() : msd_ ? msd_->ndet() : 1;
  const int nsl[2] = {wup_, wdn_};
  z64* buf[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  try {
    for (int sp = 0; sp < 2; ++sp) {
      buf[sp] = dalloc<z64>((size_t)R * std::max(nsl[sp], 1) * N);
      buf[5 + sp] = dalloc<z64>((size_t)R * std::max(nsl[sp], 1) * N);
    }
    buf[2] = dalloc<z64>(R);
    if (trial_) {
      trial_->references(buf[0], buf[1], buf[2]);
      for (int sp = 0; sp < 2; ++sp) {
        const int na = trial_->nact(sp), n = nsl[sp];
        const std::vector<int> occ = trial_->reference_orbitals(sp);
        std::vector<z64> sel((size_t)na * R * n, z64{0.0, 0.0});
        for (int d = 0; d < R; ++d)
          for (int k = 0; k < n; ++k) sel[(size_t)occ[(size_t)d * n + k] * R * n + (size_t)d * n + k] = {1.0, 0.0};
        buf[3 + sp] = dalloc<z64>(sel.size());
        AFQ_HIP(hipMemcpyAsync(buf[3 + sp], sel.data(), sizeof(z64) * sel.size(),
                               hipMemcpyHostToDevice, s));
        AFQ_HIP(hipStreamSynchronize(s));
      }
    } else if (msd_) {
      // NOMSD with several determinants: each one a reference with conj(ci)
      msd_->references(buf[0], buf[1], buf[2]);
      AFQ_HIP(hipStreamSynchronize(s));
    } else {
      // NOMSD: the single determinant, coefficient 1
      for (int sp = 0; sp < 2; ++sp)
        if (nsl[sp] > 0)
          AFQ_HIP(hipMemcpyAsync(buf[sp], ham_.PsiT() + (sp == 0 ? 0 : (size_t)wup_ * N),
                                 sizeof(z64) * nsl[sp] * N, hipMemcpyDeviceToDevice, s));
      fill_z64(s, 1, {1.0, 0.0}, buf[2]);
      AFQ_HIP(hipStreamSynchronize(s));
    }
  } catch (...) {
    for (z64* q : buf)
      if (q) (void)hipFree(q);
    throw;
  }
  // Slater-matrix form [R][N][n]: conj-transpose all rows ([N][R n]), then each reference contiguous
  for (int sp = 0; sp < 2; ++sp) {
    if (nsl[sp] == 0) continue;
    z64* tmp = scratch_as<z64>(ctx_, "ref_tmp", (size_t)R * nsl[sp] * N);
    conj_transpose(s, R * nsl[sp], N, buf[sp], tmp);
    swap_mid(s, 1, N, R, nsl[sp], tmp, buf[5 + sp]);
  }
  AFQ_HIP(hipStreamSynchronize(s));
  ref_rows_[0] = buf[0];
  ref_rows_[1] = buf[1];
  ref_cc_ = buf[2];
  ref_sel_[0] = buf[3];
  ref_sel_[1] = buf[4];
  ref_cols_[0] = buf[5];
  ref_cols_[1] = buf[6];
  nref_ = R;
}

int Propagator::number_of_references() {
  ensure_references();
  return nref_;
}

void Propagator::bp_propagate_refs(const WalkerState& st, int nbp, int ortho, z64* Lp[2], z64* ldR) {
  AFQ_CHECK(bp_n_ > 0, kErrState, "back-propagation is not enabled");
  AFQ_CHECK(nbp > 0 && nbp <= bp_n_, kErrArg, "back-propagation length exceeds the field ring");
  AFQ_CHECK(ortho > 0, kErrArg, "orthogonalization interval must be positive");
  AFQ_CHECK(st.W <= W_, kErrShape, "walker count exceeds the propagator's capacity");
  ensure_references();
  hipStream_t s = ctx_.stream();
  const int W = st.W, N = ham_.N(), M = ham_.M(), R = nref_;
  const int nsp = closed_ ? 1 : 2;
  const int nsl[2] = {wup_, closed_ ? 0 : wdn_};
  uint32_t C = 0;  // completed steps: the newest stored field sits at slot (C - 1) % nback
  AFQ_HIP(hipMemcpyAsync(&C, step_dev_, sizeof(C), hipMemcpyDeviceToHost, s));
  AFQ_HIP(hipStreamSynchronize(s));
  if (ref_sel_[0]) {
    bp_propagate_orbitals(st, C, nbp, ortho, Lp, ldR);
    return;
  }
  // general references (one determinant, or the determinants of a NOMSD expansion): L_s
  // [R][W][N][n_s] starts from each reference's orbitals in every walker, and each reference
  // is propagated and orthogonalized as a batch of its own with the step's V
  for (int sp = 0; sp < nsp; ++sp) {
    Lp[sp] = nullptr;
    if (nsl[sp] == 0) continue;  // a spin without electrons (a polarized COLLINEAR set)
    Lp[sp] = scratch_as<z64>(ctx_, sp == 0 ? "bp_L0" : "bp_L1", (size_t)R * W * N * nsl[sp]);
    for (int d = 0; d < R; ++d)
      broadcast_z64(s, W, (long)N * nsl[sp], ref_cols_[sp] + (size_t)d * N * nsl[sp],
                    Lp[sp] + (size_t)d * W * N * nsl[sp]);
  }
  // log det R accumulates reference-major [R][W] (each reference's batch contiguous) and is
  // handed back walker-major, ldR [W][R], the layout the reference coefficients read
  z64* ldRT = R > 1 ? scratch_as<z64>(ctx_, "bp_ldRT", (size_t)R * W) : ldR;
  fill_z64(s, (long)R * W, {0, 0}, ldRT);
  z64* X = scratch_as<z64>(ctx_, "bp_X", (size_t)W * M);
  z64* V = scratch_as<z64>(ctx_, "bp_V", (size_t)ham_.nspin_L() * W * N * N);
  for (int ni = nbp - 1; ni >= 0; --ni) {
    const long back = (long)nbp - ni;  // 1 = the newest field
    const int slot = (int)((((long)C - back) % bp_n_ + bp_n_) % bp_n_);
    gather_fields(s, W, M, bp_fields_, bp_n_, slot, X);
    ham_.vHS(X, W, p_.dt, V, kFP64);
    const bool qr = ni == 0 || ni % ortho == 0;
    for (int d = 0; d < R; ++d)
      for (int sp = 0; sp < nsp; ++sp) {
        const int n = nsl[sp];
        if (n == 0) continue;
        z64* L = Lp[sp] + (size_t)d * W * N * n;
        det_ops::PropagateArgs pa = spin_args(sp, W, n, 'H', V, kFP64, L, (long)N * n);
        pa.v_symmetric = true;  // vHS writes both triangles from the same sums
        det_ops::propagate(ctx_, pa);
        // logdetR accumulates log det R of both spins (AFQMCBasePropagator.cpp:473-478)
        if (qr) det_ops::orthogonalize(ctx_, L, (long)N * n, W, N, n, ldRT + (size_t)d * W);
      }
  }
  if (R > 1) swap_mid(s, 1, R, W, 1, ldRT, ldR);
}

void Propagator::bp_propagate_orbitals(const WalkerState& st, uint32_t C, int nbp, int ortho,
                                       z64* Lp[2], z64* ldR) {
  using gemm::EpiStore;
  using gemm::Opnd;
  hipStream_t s = ctx_.stream();
  const int W = st.W, N = ham_.N(), M = ham_.M(), R = nref_;
  const int nsl[2] = {wup_, wdn_};
  z64 *O[2], *Oold[2], *Mall[2], *Mtmp[2], *Mp[2], *Tm[2];
  const char* nm[2][6] = {{"bpo_O0", "bpo_Oo0", "bpo_M0", "bpo_Mt0", "bpo_Mp0", "bpo_T0"},
                          {"bpo_O1", "bpo_Oo1", "bpo_M1", "bpo_Mt1", "bpo_Mp1", "bpo_T1"}};
  for (int sp = 0; sp < 2; ++sp) {
    const int na = trial_->nact(sp), n = nsl[sp];
    AFQ_CHECK(n <= na && na <= N && na <= 64, kErrShape, "active space for back-propagation");
    const size_t Rn = (size_t)R * n;
    z64* cols = scratch_as<z64>(ctx_, sp == 0 ? "bpo_c0" : "bpo_c1", (size_t)N * na);
    conj_transpose(s, na, N, trial_->orbitals(sp), cols);
    O[sp] = scratch_as<z64>(ctx_, nm[sp][0], (size_t)W * N * na);
    broadcast_z64(s, W, (long)N * na, cols, O[sp]);
    Oold[sp] = scratch_as<z64>(ctx_, nm[sp][1], (size_t)W * N * na);
    Mall[sp] = scratch_as<z64>(ctx_, nm[sp][2], (size_t)W * na * Rn);
    broadcast_z64(s, W, (long)(na * Rn), ref_sel_[sp], Mall[sp]);
    Mtmp[sp] = scratch_as<z64>(ctx_, nm[sp][3], (size_t)W * na * Rn);
    Mp[sp] = scratch_as<z64>(ctx_, nm[sp][4], (size_t)W * na * Rn);
    Tm[sp] = scratch_as<z64>(ctx_, nm[sp][5], (size_t)W * na * na);
  }
  fill_z64(s, (long)W * R, {0, 0}, ldR);
  z64* ld0 = scratch_as<z64>(ctx_, "bpo_ld", W);
  z64* X = scratch_as<z64>(ctx_, "bp_X", (size_t)W * M);
  z64* V = scratch_as<z64>(ctx_, "bp_V", (size_t)ham_.nspin_L() * W * N * N);
  for (int ni = nbp - 1; ni >= 0; --ni) {
    const long back = (long)nbp - ni;
    const int slot = (int)((((long)C - back) % bp_n_ + bp_n_) % bp_n_);
    gather_fields(s, W, M, bp_fields_, bp_n_, slot, X);
    ham_.vHS(X, W, p_.dt, V, kFP64);
    const bool qr = ni == 0 || ni % ortho == 0;
    for (int sp = 0; sp < 2; ++sp) {
      const int na = trial_->nact(sp), n = nsl[sp];
      const int Rn = R * n;
      det_ops::PropagateArgs pa = spin_args(sp, W, na, 'H', V, kFP64, O[sp], (long)N * na);
      pa.v_symmetric = true;
      det_ops::propagate(ctx_, pa);
      if (!qr) continue;
      AFQ_HIP(hipMemcpyAsync(Oold[sp], O[sp], sizeof(z64) * W * N * na, hipMemcpyDeviceToDevice, s));
      fill_z64(s, W, {0, 0}, ld0);
      det_ops::orthogonalize(ctx_, O[sp], (long)N * na, W, N, na, ld0);
      {  // T = Q^H O_old
        Opnd<double, true, false, true> A{&O[sp]->re, na, (long)N * na};
        Opnd<double, true, false> B{&Oold[sp]->re, na, (long)N * na};
        gemm_small<double>(ctx_, "bpo_T", W, na, na, N, A, B,
                           EpiStore<double, double>{Tm[sp], na, (long)na * na, 1, 0, 0});
      }
      {  // T M_d for every reference
        Opnd<double, true, true> A{&Tm[sp]->re, na, (long)na * na};
        Opnd<double, true, false> B{&Mall[sp]->re, Rn, (long)na * Rn};
        gemm_small<double>(ctx_, "bpo_TM", W, na, Rn, na, A, B,
                           EpiStore<double, double>{Mtmp[sp], Rn, (long)na * Rn, 1, 0, 0});
      }
      swap_mid(s, W, na, R, n, Mtmp[sp], Mp[sp]);
      det_ops::orthogonalize(ctx_, Mp[sp], (long)na * n, W * R, na, n, ldR);
      swap_mid(s, W, R, na, n, Mp[sp], Mall[sp]);
    }
  }
  // L_d = O M_d, then each reference contiguous: Lp [W][R][N][n]
  for (int sp = 0; sp < 2; ++sp) {
    const int na = trial_->nact(sp), n = nsl[sp];
    const int Rn = R * n;
    z64* L = scratch_as<z64>(ctx_, sp == 0 ? "bp_L0" : "bp_L1", (size_t)W * N * Rn);
    Opnd<double, true, true> A{&O[sp]->re, na, (long)N * na};
    Opnd<double, true, false> B{&Mall[sp]->re, Rn, (long)na * Rn};
    gemm_small<double>(ctx_, "bpo_L", W, N, Rn, na, A, B,
                       EpiStore<double, double>{L, Rn, (long)N * Rn, 1, 0, 0});
    // [W][N][R][n] -> [R][W N][n]: reference-major, each reference's walkers contiguous
    Lp[sp] = scratch_as<z64>(ctx_, sp == 0 ? "bp_Lp0" : "bp_Lp1", (size_t)W * N * Rn);
    swap_mid(s, 1, W * N, R, n, L, Lp[sp]);
  }
}

void Propagator::backpropagate(const WalkerState& st0, int nbp, int ortho, int restore, z64* G,
                               z64* wout) {
  const WalkerState st = packed(st0);
  AFQ_CHECK(restore >= 0 && restore <= 2, kErrArg, "restore must be 0, 1 or 2");
  ensure_references();
  AFQ_CHECK(nref_ == 1, kErrArg,
            "per-walker density matrices are defined for a single reference; use bp_measure");
  hipStream_t s = ctx_.stream();
  const int W = st.W, N = ham_.N();
  const int nsp = closed_ ? 1 : 2;
  const int nsl[2] = {wup_, closed_ ? 0 : wdn_};
  z64* Lp[2] = {nullptr, nullptr};
  z64* ldR = scratch_as<z64>(ctx_, "bp_ldR", (size_t)W);
  bp_propagate_refs(st, nbp, ortho, Lp, ldR);
  z64* anch[2] = {bp_anch_a_, bp_anch_b_};
  if (nsl[0] == 0 || (nsp == 2 && nsl[1] == 0)) fill_z64(s, (long)W * nsp * N * N, {0, 0}, G);
  for (int sp = 0; sp < nsp; ++sp)
    if (nsl[sp] > 0)
      det_ops::mixed_density_matrix_left(ctx_, Lp[sp], (long)N * nsl[sp], nsl[sp], N, anch[sp],
                                       (long)N * nsl[sp], W, G + (size_t)sp * N * N,
                                       (long)nsp * N * N, nullptr);
  // the factor ring holds 3 nback steps, so extra path restoration (2 nbp) always fits
  uint32_t C = 0;
  AFQ_HIP(hipMemcpyAsync(&C, step_dev_, sizeof(C), hipMemcpyDeviceToHost, s));
  AFQ_HIP(hipStreamSynchronize(s));
  restore_weights(s, W, st.weight, bp_wfac_, bp_hl_, C, restore * nbp, wout);
}

namespace {

// references per chunk of a measurement: the density matrices, the compact intermediate and the
// overlap matrices of chunk * W (reference, walker) pairs within about 512 MB
int reference_chunk(int R, int W, int nsp, int N, int n) {
  const double per = 16.0 * W * ((double)nsp * N * N + (double)n * N + (double)n * n);
  return std::max(1, std::min(R, (int)(512.0 * (1 << 20) / per)));
}

}  // namespace

// The observables of one measurement from the left states L (Slater-matrix form) of R
// references against the walker matrices Phi: reference coefficients from the log overlaps (+
// conj(ldR) when given), then per chunk of references the density matrices of every
// (reference, walker) pair and their weighted sums.  Lp[s] holds [R][W][N][n_s] left states
// (per_walker) or [R][N][n_s] shared by the walkers.
void Propagator::measure_refs(int W, const z64* const Lp[2], bool per_walker,
                              const z64* const Phi[2], const z64* weight, const z64* ldR,
                              z64* onerdm, z64* diag2rdm, double ov_scale) {
  hipStream_t s = ctx_.stream();
  const int N = ham_.N(), R = nref_;
  const bool coll = !closed_;
  const int nsp = coll ? 2 : 1;
  const int nsl[2] = {wup_, coll ? wdn_ : 0};
  const int nmax = std::max(nsl[0], nsl[1]);
  const int RC = reference_chunk(R, W, nsp, N, nmax);
  const int ldiv = per_walker ? 1 : W;
  // elements between consecutive references of spin sp
  auto rstride = [&](int sp) { return (long)(per_walker ? W : 1) * N * nsl[sp]; };
  z64* lo = scratch_as<z64>(ctx_, "ms_lo", (size_t)R * W);
  z64* cf = scratch_as<z64>(ctx_, "ms_cf", (size_t)R * W);
  fill_z64(s, (long)R * W, {0, 0}, lo);
  for (int d0 = 0; d0 < R; d0 += RC) {
    const int B = std::min(RC, R - d0) * W;
    for (int sp = 0; sp < nsp; ++sp)
      if (nsl[sp] > 0)
        det_ops::mixed_density_matrix_batched(ctx_, Lp[sp] + (size_t)d0 * rstride(sp),
                                              (long)N * nsl[sp], ldiv, nsl[sp], N, Phi[sp],
                                              (long)N * nsl[sp], W, B, nullptr, 0,
                                              lo + (size_t)d0 * W);
  }
  reference_coefficients(s, W, R, lo, ref_cc_, ldR, cf, ov_scale);
  z64* G = scratch_as<z64>(ctx_, "ms_G", (size_t)RC * W * nsp * N * N);
  // a spin without electrons keeps a zero density matrix
  if (nsl[0] == 0 || (nsp == 2 && nsl[1] == 0))
    fill_z64(s, (long)RC * W * nsp * N * N, {0, 0}, G);
  for (int d0 = 0; d0 < R; d0 += RC) {
    const int B = std::min(RC, R - d0) * W;
    for (int sp = 0; sp < nsp; ++sp)
      if (nsl[sp] > 0)
        det_ops::mixed_density_matrix_batched(ctx_, Lp[sp] + (size_t)d0 * rstride(sp),
                                              (long)N * nsl[sp], ldiv, nsl[sp], N, Phi[sp],
                                              (long)N * nsl[sp], W, B, G + (size_t)sp * N * N,
                                              (long)nsp * N * N, nullptr);
    rdm_measure(ctx_, B, nsp, N, coll, G, (long)nsp * N * N, weight, cf + (size_t)d0 * W, d0 > 0,
                onerdm, diag2rdm, W);
  }
}

void Propagator::bp_measure(const WalkerState& st0, int nbp, int ortho, int restore, z64* onerdm,
                            z64* diag2rdm) {
  const WalkerState st = packed(st0);
  AFQ_CHECK(restore >= 0 && restore <= 2, kErrArg, "restore must be 0, 1 or 2");
  ensure_references();
  hipStream_t s = ctx_.stream();
  const int W = st.W, R = nref_;
  z64* Lp[2] = {nullptr, nullptr};
  z64* ldR = scratch_as<z64>(ctx_, "bp_ldR", (size_t)W * R);
  bp_propagate_refs(st, nbp, ortho, Lp, ldR);
  // path-restored weights (BackPropEstimator.hpp:212-228)
  uint32_t C = 0;
  AFQ_HIP(hipMemcpyAsync(&C, step_dev_, sizeof(C), hipMemcpyDeviceToHost, s));
  AFQ_HIP(hipStreamSynchronize(s));
  z64* wr = scratch_as<z64>(ctx_, "bpm_w", W);
  restore_weights(s, W, st.weight, bp_wfac_, bp_hl_, C, restore * nbp, wr);
  // reference coefficients conj(ci) <L_d|anchor> e^{conj(logdetR)} (constructBPMeasurementInputs,
  // BackPropEstimator.hpp:62-104, which doubles both for CLOSED walkers); the back-propagated
  // references are [R][W][N][n]
  const z64* L[2] = {Lp[0], Lp[1]};
  const z64* anch[2] = {bp_anch_a_, bp_anch_b_};
  measure_refs(W, L, true, anch, wr, ldR, onerdm, diag2rdm, closed_ ? 2.0 : 1.0);
}

void Propagator::mixed_measure(const WalkerState& st0, z64* onerdm, z64* diag2rdm) {
  const WalkerState st = packed(st0);
  AFQ_CHECK(st.W <= W_, kErrShape, "walker count exceeds the propagator's capacity");
  ensure_references();
  // constructMixedMeasurementInputs (MixedEstimator.hpp:40-106): conj(ci) <D_d|Phi> normalized
  // over the references of each walker (the reference's shift by OVLP cancels in the ratio); its
  // overlaps take the alpha block once, also for CLOSED walkers
  const z64* L[2] = {ref_cols_[0], ref_cols_[1]};
  const z64* Phi[2] = {st.Phi, st.Phi_b};
  measure_refs(st.W, L, false, Phi, st.weight, nullptr, onerdm, diag2rdm, 1.0);
}

// ---- time-evolved operators ------------------------------------------------------------------

const z64* Propagator::P1_inverse() {
  if (P1inv_) return P1inv_;
  hipStream_t s = ctx_.stream();
  const int N = ham_.N();
  const long NP = (long)ham_.p1_spins() * N * N;
  z64* p = dalloc<z64>(NP);
  double* pr = nullptr;
  try {
    ham_.one_body_propagator(vMF_, p_.dt, p, true);
    // H1 real <=> P1 real <=> its inverse real
    if (real_P1_) {
      pr = dalloc<double>(NP);
      real_part(s, NP, p, pr);
    }
    AFQ_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(p);
    if (pr) (void)hipFree(pr);
    throw;
  }
  P1inv_ = p;
  P1invr_ = pr;
  return P1inv_;
}

const z64* Propagator::te_operator(int which) const {
  AFQ_CHECK(which >= 0 && which < 3, kErrArg, "operator index must be 0 (X), 1 (Y) or 2 (M)");
  return te_op_[which];
}

void Propagator::te_anchor(bool follow_walkers) {
  hipStream_t s = ctx_.stream();
  const int N = ham_.N();
  const int nsp = closed_ ? 1 : 2;
  const size_t cnt = (size_t)W_ * nsp * N * N;
  if (follow_walkers != te_follow_) {
    AFQ_HIP(hipStreamSynchronize(s));
    clear_graphs();  // captured population control moves different rows
    te_follow_ = follow_walkers;
  }
  if (!te_op_[0]) {
    if (te_follow_) {
      AFQ_HIP(hipStreamSynchronize(s));
      clear_graphs();
    }
    z64* buf[3] = {nullptr, nullptr, nullptr};
    try {
      for (int i = 0; i < 3; ++i) buf[i] = dalloc<z64>(cnt);
    } catch (...) {
      for (z64* q : buf)
        if (q) (void)hipFree(q);
      throw;
    }
    for (int i = 0; i < 3; ++i) te_op_[i] = buf[i];
  }
  z64* I = scratch_as<z64>(ctx_, "te_I", (size_t)N * N);
  identity(s, N, I);
  broadcast_z64(s, (long)W_ * nsp, (long)N * N, I, te_op_[0]);
  broadcast_z64(s, (long)W_ * nsp, (long)N * N, I, te_op_[1]);
  fill_z64(s, (long)cnt, {0.0, 0.0}, te_op_[2]);
}

void Propagator::te_propagate(const WalkerState& st0, int nsteps) {
  const WalkerState st = packed(st0);
  using gemm::EpiStore;
  using gemm::Opnd;
  AFQ_CHECK(bp_n_ > 0, kErrState, "back-propagation is not enabled");
  AFQ_CHECK(te_op_[0], kErrState, "the time-evolved operators have no anchor");
  AFQ_CHECK(nsteps > 0 && nsteps <= bp_n_, kErrArg,
            "operator propagation length must be positive and fit the field ring");
  AFQ_CHECK(st.W == W_, kErrShape, "the operators belong to the propagator's walker slots");
  hipStream_t s = ctx_.stream();
  const int W = st.W, N = ham_.N(), M = ham_.M();
  const int nsp = closed_ ? 1 : 2;
  const int nsl[2] = {wup_, closed_ ? 0 : wdn_};
  const long NN = (long)N * N, obs = (long)nsp * NN;  // batch stride of X, Y, M
  P1_inverse();
  uint32_t C = 0;  // completed steps: the newest stored field sits at slot (C - 1) % nback
  AFQ_HIP(hipMemcpyAsync(&C, step_dev_, sizeof(C), hipMemcpyDeviceToHost, s));
  AFQ_HIP(hipStreamSynchronize(s));
  z64* Xf = scratch_as<z64>(ctx_, "bp_X", (size_t)W * M);
  const int nl = ham_.nspin_L();
  z64* V = scratch_as<z64>(ctx_, "bp_V", (size_t)nl * W * NN);
  z64 *Xo = te_op_[0], *Yo = te_op_[1], *Mo = te_op_[2];
  for (int i = 0; i < nsteps; ++i) {  // oldest first
    const long back = (long)nsteps - i;
    const int slot = (int)((((long)C - back) % bp_n_ + bp_n_) % bp_n_);
    gather_fields(s, W, M, bp_fields_, bp_n_, slot, Xf);
    ham_.vHS(Xf, W, p_.dt, V, kFP64);
    for (int sp = 0; sp < nsp; ++sp) {
      det_ops::PropagateArgs pa = spin_args(sp, W, N, 'N', V, kFP64, Xo + sp * NN, obs);
      det_ops::propagate(ctx_, pa);
    }
    // B^{-H} = P1inv expM(-V)^H P1inv: the adjoint propagator of -V with the inverse P1
    scale_add_cplx(s, (long)nl * W * NN, V, -1.0, 0.0, V);
    for (int sp = 0; sp < nsp; ++sp) {
      det_ops::PropagateArgs pa = spin_args(sp, W, N, 'H', V, kFP64, Yo + sp * NN, obs, true);
      pa.v_symmetric = true;
      det_ops::propagate(ctx_, pa);
    }
  }
  // regularization with the walker matrices S [W][N][n]
  const z64* Sp[2] = {st.Phi, st.Phi_b};
  for (int sp = 0; sp < nsp; ++sp) {
    const int n = nsl[sp];
    z64 *X = Xo + sp * NN, *Y = Yo + sp * NN, *Mm = Mo + sp * NN;
    if (n == 0) {  // no occupied space: Y <- 0, M and X unchanged
      fill_blocks_z64(s, W, NN, obs, {0.0, 0.0}, Y);
      continue;
    }
    const z64* S = Sp[sp];
    z64* SY = scratch_as<z64>(ctx_, "te_SY", (size_t)W * n * N);
    const Opnd<double, true, false, true> SH{&S->re, n, (long)N * n};  // S^H: (a, k) = conj S[k][a]
    const Opnd<double, true, true> Sn{&S->re, n, (long)N * n};          // S: (i, a) = S[i][a]
    const Opnd<double, true, false> SYt{&SY->re, N, (long)n * N};       // (j, a) = SY[a][j]
    {  // SY = S^H Y; Y <- S SY
      Opnd<double, true, false> Bt{&Y->re, N, obs};
      gemm_small<double>(ctx_, "te_SY", W, n, N, N, SH, Bt,
                         EpiStore<double, double>{SY, N, (long)n * N, 1, 0, 0});
      gemm_small<double>(ctx_, "te_Y", W, N, N, n, Sn, SYt,
                         EpiStore<double, double>{Y, N, obs, 1, 0, 0});
    }
    {  // M += Y^H X
      Opnd<double, true, false, true> YH{&Y->re, N, obs};
      Opnd<double, true, false> Bt{&X->re, N, obs};
      gemm_small<double>(ctx_, "te_M", W, N, N, N, YH, Bt,
                         EpiStore<double, double>{Mm, N, obs, 1, 0, 1}, false);
    }
    {  // SY = S^H X; X <- X - S SY
      Opnd<double, true, false> Bt{&X->re, N, obs};
      gemm_small<double>(ctx_, "te_SY", W, n, N, N, SH, Bt,
                         EpiStore<double, double>{SY, N, (long)n * N, 1, 0, 0});
      gemm_small<double>(ctx_, "te_X", W, N, N, n, Sn, SYt,
                         EpiStore<double, double>{X, N, obs, -1, 0, 1}, false);
    }
  }
}

void Propagator::te_measure(const WalkerState& st0, int restore_steps, bool transpose_m,
                            z64* onerdm, z64* diag2rdm) {
  const WalkerState st = packed(st0);
  using gemm::EpiStore;
  using gemm::Opnd;
  AFQ_CHECK(bp_n_ > 0, kErrState, "back-propagation is not enabled");
  AFQ_CHECK(te_op_[0], kErrState, "the time-evolved operators have no anchor");
  AFQ_CHECK(st.W == W_, kErrShape, "the operators belong to the propagator's walker slots");
  AFQ_CHECK(restore_steps >= 0 && restore_steps <= bp_hl_, kErrArg,
            "path restoration over more steps than the factor ring holds");
  ensure_references();
  hipStream_t s = ctx_.stream();
  const int W = st.W, N = ham_.N(), R = nref_;
  const bool coll = !closed_;
  const int nsp = coll ? 2 : 1;
  const int nsl[2] = {wup_, coll ? wdn_ : 0};
  const long NN = (long)N * N, obs = (long)nsp * NN;
  const z64* Phi[2] = {st.Phi, st.Phi_b};
  // weights times the path-restoration factors
  uint32_t C = 0;
  AFQ_HIP(hipMemcpyAsync(&C, step_dev_, sizeof(C), hipMemcpyDeviceToHost, s));
  AFQ_HIP(hipStreamSynchronize(s));
  z64* wr = scratch_as<z64>(ctx_, "te_w", W);
  restore_weights(s, W, st.weight, bp_wfac_, bp_hl_, C, restore_steps, wr);
  // reference coefficients from <Ref_d|Phi> (the reference's OVLP shift cancels in the ratio)
  z64* lo = scratch_as<z64>(ctx_, "ms_lo", (size_t)R * W);
  z64* cf = scratch_as<z64>(ctx_, "ms_cf", (size_t)R * W);
  const int nmax = std::max(nsl[0], nsl[1]);
  const int RC = reference_chunk(R, W, nsp, N, nmax);
  fill_z64(s, (long)R * W, {0, 0}, lo);
  for (int d0 = 0; d0 < R; d0 += RC)
    for (int sp = 0; sp < nsp; ++sp)
      if (nsl[sp] > 0)
        det_ops::mixed_density_matrix_batched(ctx_, ref_cols_[sp] + (size_t)d0 * N * nsl[sp],
                                              (long)N * nsl[sp], W, nsl[sp], N, Phi[sp],
                                              (long)N * nsl[sp], W, std::min(RC, R - d0) * W,
                                              nullptr, 0, lo + (size_t)d0 * W);
  // constructTimeEvolvedMeasurementInputs doubles the CLOSED overlaps
  // (TimeEvolvedBPEstimator.hpp:85-95)
  reference_coefficients(s, W, R, lo, ref_cc_, nullptr, cf, closed_ ? 2.0 : 1.0);
  z64* Gc = scratch_as<z64>(ctx_, "te_Gc", (size_t)W * nmax * N);
  z64* GY = scratch_as<z64>(ctx_, "te_GY", (size_t)W * nmax * N);
  z64* XR = scratch_as<z64>(ctx_, "te_XR", (size_t)W * nmax * N);
  z64* G = scratch_as<z64>(ctx_, "te_G", (size_t)W * obs);
  z64* dum = scratch_as<z64>(ctx_, "te_ov", (size_t)W);
  for (int d = 0; d < R; ++d) {
    if (transpose_m)
      transpose_blocks(s, (long)W * nsp, N, te_op_[2], G);
    else
      AFQ_HIP(hipMemcpyAsync(G, te_op_[2], sizeof(z64) * W * obs, hipMemcpyDeviceToDevice, s));
    for (int sp = 0; sp < nsp; ++sp) {
      const int n = nsl[sp];
      if (n == 0) continue;
      const z64* ref = ref_rows_[sp] + (size_t)d * n * N;  // Ref_d^H [n][N]
      const z64 *X = te_op_[0] + sp * NN, *Y = te_op_[1] + sp * NN;
      // Gc = T^{-T} Phi^T with T = Ref^H Phi (the log overlap it also returns is not used)
      fill_z64(s, W, {0.0, 0.0}, dum);
      det_ops::mixed_density_matrix(ctx_, ref, n, N, Phi[sp], (long)N * n, W, Gc, (long)n * N,
                                    dum, true);
      {  // GY = Gc conj(Y)
        Opnd<double, true, true> A{&Gc->re, N, (long)n * N};
        Opnd<double, true, false, true> Bt{&Y->re, N, obs};
        gemm_small<double>(ctx_, "te_GY", W, n, N, N, A, Bt,
                           EpiStore<double, double>{GY, N, (long)n * N, 1, 0, 0});
      }
      {  // XR = Ref^H X
        Opnd<double, true, true> A{&ref->re, N, 0};
        Opnd<double, true, false> Bt{&X->re, N, obs};
        gemm_small<double>(ctx_, "te_XR", W, n, N, N, A, Bt,
                           EpiStore<double, double>{XR, N, (long)n * N, 1, 0, 0});
      }
      {  // G += XR^T GY
        Opnd<double, true, false> A{&XR->re, N, (long)n * N};
        Opnd<double, true, false> Bt{&GY->re, N, (long)n * N};
        gemm_small<double>(ctx_, "te_G", W, N, N, n, A, Bt,
                           EpiStore<double, double>{G + sp * NN, N, obs, 1, 0, 1}, false);
      }
    }
    rdm_measure(ctx_, W, nsp, N, coll, G, obs, wr, cf + (size_t)d * W, d > 0, onerdm, diag2rdm, W);
  }
}

int Propagator::status() {
  int h = 0;
  AFQ_HIP(hipMemcpyAsync(&h, status_dev_, sizeof(int), hipMemcpyDeviceToHost, ctx_.stream()));
  AFQ_HIP(hipStreamSynchronize(ctx_.stream()));
  return h;
}

void Propagator::energy(const WalkerState& st, z64* E, z64* ovlp, Precision prec) {
  check_strides(st);
  if (trial_) {  // PHMSD::Energy (energy_alg1), FP64
    trial_->energy(st.Phi, st.Phi_b, st.W, E, ovlp, bs_a(st), bs_b(st));
    return;
  }
  if (msd_) {  // NOMSD::Energy with ndet > 1
    msd_->energy(st.Phi, st.Phi_b, st.W, E, ovlp, prec, bs_a(st), bs_b(st));
    return;
  }
  hipStream_t s = ctx_.stream();
  const int W = st.W, N = ham_.N(), nup = ham_.nup(), ndown = ham_.ndown(), nel = ham_.nel();
  const bool closed = closed_;
  z64* G = scratch_as<z64>(ctx_, "en_G", (size_t)W * nel * N);
  fill_z64(s, W, {0, 0}, ovlp);
  det_ops::mixed_density_matrix(ctx_, ham_.PsiT(), nup, N, st.Phi, bs_a(st), W, G,
                                (long)nel * N, ovlp, true, closed ? 2.0 : 1.0);
  if (!closed && ndown > 0)
    det_ops::mixed_density_matrix(ctx_, ham_.PsiT() + (size_t)nup * N, ndown, N, st.Phi_b,
                                  bs_b(st), W, G + (size_t)nup * N, (long)nel * N, ovlp, true);
  ham_.energy(G, W, E, true, true, true, prec);
}

void Propagator::step_graph(const WalkerState& st0, const double* Eshift, bool ortho, bool popctl,
                            double relax, double* energy_out) {
  hipStream_t s = ctx_.stream();
  check_strides(st0);
  const WalkerState& st = st0;
  std::vector<const void*> key = {st.Phi, st.Phi_b, st.weight, st.phase, st.eloc, st.ovlp,
                                  st.theta, Eshift, energy_out, (const void*)(long)st.W,
                                  (const void*)(long)((ortho ? 1 : 0) | (popctl ? 2 : 0)),
                                  (const void*)st.phi_bs, (const void*)st.phi_b_bs,
                                  (const void*)st.prop_bs, st.e1, st.exx, st.ej,
                                  (const void*)st.e_bs};
  long rbits;
  static_assert(sizeof(rbits) == sizeof(relax), "");
  std::memcpy(&rbits, &relax, sizeof(rbits));
  key.push_back((const void*)rbits);
  auto body = [&]() {
    const WalkerState sp = stage_properties(st0);
    step_impl(sp, nullptr, Eshift);
    if (ortho) orthogonalize_impl(sp);
    if (popctl) population_control_impl(sp, nullptr);
    average_pseudo_energy(s, sp.W, sp.weight, sp.eloc, p_.dt, relax, energy_out,
                          const_cast<double*>(Eshift));
    increment_counter(s, step_dev_);
    unstage_properties(st0, sp);
  };
  // A scratch buffer that grew since the graphs were captured (another call on this context
  // needed more room) was freed: the graphs may hold its old address.
  if (!graphs_.empty() && graph_gen_ != ctx_.generation()) clear_graphs();
  for (size_t i = 0; i < graphs_.size(); ++i)
    if (graphs_[i].key == key) {
      AFQ_HIP(hipGraphLaunch(graphs_[i].exec, s));
      return;
    }
  if (std::find(warmed_.begin(), warmed_.end(), key) == warmed_.end()) {
    // first use of this key: run eagerly so that every scratch buffer its body needs exists
    // (and its products are tuned) before capture
    body();
    warmed_.push_back(key);
    return;
  }
  if (graphs_.empty()) graph_gen_ = ctx_.generation();
  AFQ_CHECK(graph_gen_ == ctx_.generation(), kErrState, "scratch generation changed between captures");
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  AFQ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  AFQ_HIP(hipStreamEndCapture(s, &g));
  AFQ_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
  graphs_.push_back({key, g, ge});
  AFQ_HIP(hipGraphLaunch(ge, s));
}

}  // namespace afq
