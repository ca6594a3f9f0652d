// Multi-determinant NOMSD trial on the device (NOMSD.cpp:39-61, 159-333, 338-520).
//
// Every quantity of an expansion sum_d c_d |D_d> of general (non-orthogonal) determinants is
// a combination of single-determinant ones, evaluated for all walkers of one determinant at a
// time with the batched Slater-determinant kernels and combined per walker:
//   Log_Overlap          log sum_d conj(c_d) <D_d|Phi>, each term exp(l_d - l_max) with l_max
//                        the term of largest real part (one thread per walker, determinants
//                        in order)
//   MixedDensityMatrix   full G = sum_d w_d G_d / sum_d w_d, w_d = conj(c_d) e^{ov_d - log_m}
//                        with log_m the walkers' OVLP (the reference's scale, which cancels)
//   vbias                the full-G contraction with the Cholesky tensor (the reference's
//                        ndet > 1 branch): one MFMA GEMM [W x nspin N^2] . [N^2 x M]
//   Energy               each determinant's energy from its compact G and its own half-rotated
//                        tensors (built once), weighted as above about the largest overlap
//   vMF                  the trial density over the pairs p <= q of determinants with the
//                        weights Re(conj(c_q) c_p <D_q|D_p>), built once on the device.
#include <cmath>

#include "ops.hpp"

namespace afq {

namespace {

constexpr int kNT = 256;

inline dim3 grid_of(long n) { return dim3((unsigned)((n + kNT - 1) / kNT)); }

__device__ inline z64 zmul(z64 a, z64 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline z64 zexp(z64 a) {
  const double e = exp(a.re);
  return {e * cos(a.im), e * sin(a.im)};
}
__device__ inline z64 zlog(z64 a) { return {log(hypot(a.re, a.im)), atan2(a.im, a.re)}; }

// Ov[w] = l_max + log sum_d exp(l[d][w] - l_max), l_max the first term of largest real part
__global__ void k_logsumexp(int ndet, int W, const z64* __restrict__ l, z64* __restrict__ Ov) {
  const int w = blockIdx.x * kNT + threadIdx.x;
  if (w >= W) return;
  z64 m = l[w];
  for (int d = 1; d < ndet; ++d)
    if (l[(long)d * W + w].re > m.re) m = l[(long)d * W + w];
  z64 s = {0.0, 0.0};
  for (int d = 0; d < ndet; ++d) {
    const z64 t = zexp({l[(long)d * W + w].re - m.re, l[(long)d * W + w].im - m.im});
    s = {s.re + t.re, s.im + t.im};
  }
  const z64 ls = zlog(s);
  Ov[w] = {m.re + ls.re, m.im + ls.im};
}

// wts[d][w] = conj(c_d) exp(ov[d][w] - m_w), m_w = scale[w] when given, else the first ov of
// largest real part; sum[w] = sum_d wts[d][w]; mout[w] = m_w
__global__ void k_weights(int ndet, int W, const z64* __restrict__ ov, const z64* __restrict__ cc,
                          const z64* __restrict__ scale, z64* __restrict__ wts,
                          z64* __restrict__ sum, z64* __restrict__ mout) {
  const int w = blockIdx.x * kNT + threadIdx.x;
  if (w >= W) return;
  z64 m;
  if (scale) {
    m = scale[w];
  } else {
    m = ov[w];
    for (int d = 1; d < ndet; ++d)
      if (ov[(long)d * W + w].re > m.re) m = ov[(long)d * W + w];
  }
  z64 s = {0.0, 0.0};
  for (int d = 0; d < ndet; ++d) {
    const z64 t = zmul(cc[d], zexp({ov[(long)d * W + w].re - m.re, ov[(long)d * W + w].im - m.im}));
    wts[(long)d * W + w] = t;
    s = {s.re + t.re, s.im + t.im};
  }
  sum[w] = s;
  if (mout) mout[w] = m;
}

// G[w][k] (+)= wts[w] Gd[w][k] for K entries per walker (wts: one determinant's weights)
__global__ void k_accumulate(long W, long K, const z64* __restrict__ wts, const z64* __restrict__ Gd,
                             z64* __restrict__ G, int first) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= W * K) return;
  const z64 v = zmul(wts[t / K], Gd[t]);
  if (first) {
    G[t] = v;
  } else {
    const z64 o = G[t];
    G[t] = {o.re + v.re, o.im + v.im};
  }
}

// G[w][k] /= den[w]; Ov[w] = den[w] + add[w]
__global__ void k_normalize(long W, long K, const z64* __restrict__ den, const z64* __restrict__ add,
                            z64* __restrict__ G, z64* __restrict__ Ov) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= W * K) return;
  const long w = t / K;
  const z64 d = den[w];
  const double q = 1.0 / (d.re * d.re + d.im * d.im);
  const z64 g = G[t];
  G[t] = {(g.re * d.re + g.im * d.im) * q, (g.im * d.re - g.re * d.im) * q};
  if (Ov && t % K == 0) Ov[w] = {d.re + add[w].re, d.im + add[w].im};
}

// E[w] = sum_d wts[d][w] Ed[d][w] / sum[w]; Ov[w] = log sum[w] + m[w]
__global__ void k_combine_energy(int ndet, int W, const z64* __restrict__ wts,
                                 const z64* __restrict__ Ed, const z64* __restrict__ sum,
                                 const z64* __restrict__ m, z64* __restrict__ E,
                                 z64* __restrict__ Ov) {
  const int w = blockIdx.x * kNT + threadIdx.x;
  if (w >= W) return;
  const z64 s = sum[w];
  const double q = 1.0 / (s.re * s.re + s.im * s.im);
  for (int c = 0; c < 3; ++c) {
    z64 a = {0.0, 0.0};
    for (int d = 0; d < ndet; ++d) {
      const z64 t = zmul(wts[(long)d * W + w], Ed[((long)d * W + w) * 3 + c]);
      a = {a.re + t.re, a.im + t.im};
    }
    E[(long)w * 3 + c] = {(a.re * s.re + a.im * s.im) * q, (a.im * s.re - a.re * s.im) * q};
  }
  const z64 ls = zlog(s);
  Ov[w] = {ls.re + m[w].re, ls.im + m[w].im};
}

// ov[d][w] += c[d]
__global__ void k_add_row_const(int ndet, int W, const z64* __restrict__ c, z64* __restrict__ ov) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= (long)ndet * W) return;
  const z64 a = c[t / W];
  ov[t] = {ov[t].re + a.re, ov[t].im + a.im};
}

// one determinant's weight wd[w] = cc exp(ov[w] - scale[w]); sum[w] (+)= wd[w]
__global__ void k_det_weight(int W, z64 cc, const z64* __restrict__ ov,
                             const z64* __restrict__ scale, z64* __restrict__ wd,
                             z64* __restrict__ sum, int first) {
  const int w = blockIdx.x * kNT + threadIdx.x;
  if (w >= W) return;
  const z64 t = zmul(cc, zexp({ov[w].re - scale[w].re, ov[w].im - scale[w].im}));
  wd[w] = t;
  sum[w] = first ? t : z64{sum[w].re + t.re, sum[w].im + t.im};
}

// y += a x (real a), n entries
__global__ void k_axpy(long n, double a, const z64* __restrict__ x, z64* __restrict__ y) {
  const long t = (long)blockIdx.x * kNT + threadIdx.x;
  if (t >= n) return;
  y[t] = {y[t].re + a * x[t].re, y[t].im + a * x[t].im};
}

// v[m] = (Re v[m], 0)
__global__ void k_zero_imag(int n, z64* __restrict__ v) {
  const int t = blockIdx.x * kNT + threadIdx.x;
  if (t < n) v[t].im = 0.0;
}

template <typename T>
T* dnew(size_t count) {
  void* p = nullptr;
  if (hipMalloc(&p, (count ? count : 1) * sizeof(T)) != hipSuccess)
    throw Error(kErrAlloc, "device allocation failed");
  return reinterpret_cast<T*>(p);
}

}  // namespace

NomsdTrial::NomsdTrial(CholeskyOps& ham, int ndet, const z64* PsiT, const z64* ci)
    : ham_(ham), ctx_(ham.ctx()), ndet_(ndet) {
  AFQ_CHECK(ndet >= 1, kErrArg, "a NOMSD trial needs at least one determinant");
  AFQ_CHECK(PsiT && ci, kErrArg, "NULL argument");
  hipStream_t s = ctx_.stream();
  const int N = ham.N(), nel = ham.nel();
  const size_t per = (size_t)nel * N;
  // determinant 0 is the hamiltonian's own
  std::vector<z64> a(per), b(per);
  AFQ_HIP(hipMemcpyAsync(a.data(), PsiT, sizeof(z64) * per, hipMemcpyDeviceToHost, s));
  AFQ_HIP(hipMemcpyAsync(b.data(), ham.PsiT(), sizeof(z64) * per, hipMemcpyDeviceToHost, s));
  AFQ_HIP(hipStreamSynchronize(s));
  for (size_t i = 0; i < per; ++i)
    AFQ_CHECK(a[i].re == b[i].re && a[i].im == b[i].im, kErrArg,
              "determinant 0 must be the hamiltonian's trial");
  ci_h_.assign(ci, ci + ndet);
  try {
    dets_.push_back(ham.self());
    for (int d = 1; d < ndet; ++d) dets_.push_back(ham.rotate(PsiT + (size_t)d * per));
    std::vector<z64> cc(ndet), lcc(ndet);
    for (int d = 0; d < ndet; ++d) {
      const z64 c = ci_h_[d];
      cc[d] = {c.re, -c.im};
      // log(conj(c)), with |c| < 1e-12 as 1e-12 (NOMSD.cpp:316)
      const double r = std::hypot(c.re, c.im);
      lcc[d] = r < 1e-12 ? z64{std::log(1e-12), 0.0} : z64{std::log(r), std::atan2(-c.im, c.re)};
    }
    cc_ = dnew<z64>(ndet);
    lcc_ = dnew<z64>(ndet);
    AFQ_HIP(hipMemcpyAsync(cc_, cc.data(), sizeof(z64) * ndet, hipMemcpyHostToDevice, s));
    AFQ_HIP(hipMemcpyAsync(lcc_, lcc.data(), sizeof(z64) * ndet, hipMemcpyHostToDevice, s));
    AFQ_HIP(hipStreamSynchronize(s));
  } catch (...) {
    for (size_t d = 1; d < dets_.size(); ++d) ham_.release(dets_[d]);
    if (cc_) (void)hipFree(cc_);
    if (lcc_) (void)hipFree(lcc_);
    throw;
  }
}

NomsdTrial::~NomsdTrial() {
  for (size_t d = 1; d < dets_.size(); ++d) ham_.release(dets_[d]);
  for (void* p : {(void*)cc_, (void*)lcc_, (void*)gmf_})
    if (p) (void)hipFree(p);
}

// Determinant d against the walkers: o[w] = log <D_d|Phi_w> (alpha doubled for CLOSED),
// accumulated into a zeroed o, and, with G, its density matrices (compact [nel][N] or full
// [nspin][N][N] per walker, gw entries apart)
void NomsdTrial::det_block(int d, const z64* Pa, const z64* Pb, int W, long bs_a, long bs_b,
                           z64* o, z64* G, bool compact) {
  hipStream_t s = ctx_.stream();
  const int N = ham_.N(), nup = ham_.nup(), ndown = ham_.ndown(), nel = ham_.nel();
  const bool closed = ham_.walker_type() == kClosed;
  const int nspin = closed ? 1 : 2;
  const long ka = bs_a ? bs_a : (long)N * nup, kb = bs_b ? bs_b : (long)N * ndown;
  const long gw = compact ? (long)nel * N : (long)nspin * N * N;
  const z64* P = dets_[d].PsiT;
  const double sa = closed ? 2.0 : 1.0;
  fill_z64(s, W, {0, 0}, o);
  if (G) {
    if (!closed && (nup == 0 || ndown == 0)) fill_z64(s, (long)W * gw, {0, 0}, G);
    if (nup > 0) det_ops::mixed_density_matrix(ctx_, P, nup, N, Pa, ka, W, G, gw, o, compact, sa);
    if (!closed && ndown > 0)
      det_ops::mixed_density_matrix(ctx_, P + (size_t)nup * N, ndown, N, Pb, kb, W,
                                    G + (compact ? (size_t)nup * N : (size_t)N * N), gw, o,
                                    compact, 1.0);
  } else {
    if (nup > 0) det_ops::log_overlap(ctx_, P, nup, N, Pa, ka, W, o, sa);
    if (!closed && ndown > 0)
      det_ops::log_overlap(ctx_, P + (size_t)nup * N, ndown, N, Pb, kb, W, o, 1.0);
  }
}

void NomsdTrial::log_overlap(const z64* Pa, const z64* Pb, int W, z64* Ov, long bs_a,
                             long bs_b) {
  if (W == 0) return;
  hipStream_t s = ctx_.stream();
  z64* lo = scratch_as<z64>(ctx_, "msd_lo", (size_t)ndet_ * W);
  for (int d = 0; d < ndet_; ++d) det_block(d, Pa, Pb, W, bs_a, bs_b, lo + (size_t)d * W, nullptr, true);
  // + log conj(c_d) (NOMSD.cpp:316-321), then the sum about the largest term (:324-331)
  hipLaunchKernelGGL(k_add_row_const, grid_of((long)ndet_ * W), dim3(kNT), 0, s, ndet_, W, lcc_, lo);
  AFQ_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_logsumexp, grid_of(W), dim3(kNT), 0, s, ndet_, W, lo, Ov);
  AFQ_LAUNCH_CHECK();
}

void NomsdTrial::mixed_density_matrix(const z64* Pa, const z64* Pb, int W, const z64* log_m,
                                      z64* G, z64* Ov, long bs_a, long bs_b) {
  if (W == 0) return;
  hipStream_t s = ctx_.stream();
  const int N = ham_.N();
  const int nspin = ham_.walker_type() == kClosed ? 1 : 2;
  const long K = (long)nspin * N * N;
  z64* ov = scratch_as<z64>(ctx_, "msd_ov", W);
  z64* Gd = scratch_as<z64>(ctx_, "msd_Gd", (size_t)W * K);
  z64* wd = scratch_as<z64>(ctx_, "msd_w", W);
  z64* sum = scratch_as<z64>(ctx_, "msd_sum", W);
  // the weights scale by the walkers' OVLP, so each determinant is added as it is computed
  for (int d = 0; d < ndet_; ++d) {
    const z64 c = ci_h_[d];
    det_block(d, Pa, Pb, W, bs_a, bs_b, ov, Gd, false);
    hipLaunchKernelGGL(k_det_weight, grid_of(W), dim3(kNT), 0, s, W, z64{c.re, -c.im}, ov, log_m,
                       wd, sum, d == 0 ? 1 : 0);
    AFQ_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_accumulate, grid_of((long)W * K), dim3(kNT), 0, s, (long)W, K, wd, Gd, G,
                       d == 0 ? 1 : 0);
    AFQ_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_normalize, grid_of((long)W * K), dim3(kNT), 0, s, (long)W, K, sum, log_m,
                     G, Ov);
  AFQ_LAUNCH_CHECK();
}

void NomsdTrial::vbias(const z64* Pa, const z64* Pb, int W, const z64* log_m, double dt, z64* v,
                       Precision prec, long bs_a, long bs_b) {
  if (W == 0) return;
  const int N = ham_.N();
  const int nspin = ham_.walker_type() == kClosed ? 1 : 2;
  z64* G = scratch_as<z64>(ctx_, "msd_G", (size_t)W * nspin * N * N);
  z64* Ov = scratch_as<z64>(ctx_, "msd_Ov", W);
  mixed_density_matrix(Pa, Pb, W, log_m, G, Ov, bs_a, bs_b);
  ham_.vbias_full(G, W, dt, v, prec);
}

void NomsdTrial::energy(const z64* Pa, const z64* Pb, int W, z64* E, z64* Ov, Precision prec,
                        long bs_a, long bs_b) {
  if (W == 0) return;
  hipStream_t s = ctx_.stream();
  const int N = ham_.N(), nel = ham_.nel();
  const long K = (long)nel * N;
  z64* ov = scratch_as<z64>(ctx_, "msde_ov", (size_t)ndet_ * W);
  z64* Gd = scratch_as<z64>(ctx_, "msde_Gd", (size_t)W * K);
  z64* Ed = scratch_as<z64>(ctx_, "msde_E", (size_t)ndet_ * W * 3);
  z64* wts = scratch_as<z64>(ctx_, "msde_w", (size_t)ndet_ * W);
  z64* sum = scratch_as<z64>(ctx_, "msde_sum", W);
  z64* m = scratch_as<z64>(ctx_, "msde_m", W);
  for (int d = 0; d < ndet_; ++d) {
    det_block(d, Pa, Pb, W, bs_a, bs_b, ov + (size_t)d * W, Gd, true);
    ham_.energy(Gd, W, Ed + (size_t)d * W * 3, true, true, true, prec, &dets_[d]);
  }
  // weights about the largest overlap of each walker (NOMSD.cpp:201-213)
  hipLaunchKernelGGL(k_weights, grid_of(W), dim3(kNT), 0, s, ndet_, W, ov, cc_, nullptr, wts, sum,
                     m);
  AFQ_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_combine_energy, grid_of(W), dim3(kNT), 0, s, ndet_, W, wts, Ed, sum, m, E,
                     Ov);
  AFQ_LAUNCH_CHECK();
}

void NomsdTrial::references(z64* rows0, z64* rows1, z64* cc) {
  hipStream_t s = ctx_.stream();
  const int N = ham_.N(), nup = ham_.nup(), ndown = ham_.ndown();
  for (int d = 0; d < ndet_; ++d) {
    const z64* P = dets_[d].PsiT;
    if (nup > 0)
      AFQ_HIP(hipMemcpyAsync(rows0 + (size_t)d * nup * N, P, sizeof(z64) * nup * N,
                             hipMemcpyDeviceToDevice, s));
    if (ham_.walker_type() == kCollinear && ndown > 0)
      AFQ_HIP(hipMemcpyAsync(rows1 + (size_t)d * ndown * N, P + (size_t)nup * N,
                             sizeof(z64) * ndown * N, hipMemcpyDeviceToDevice, s));
  }
  AFQ_HIP(hipMemcpyAsync(cc, cc_, sizeof(z64) * ndet_, hipMemcpyDeviceToDevice, s));
}

void NomsdTrial::vMF(double dt, z64* v) {
  hipStream_t s = ctx_.stream();
  const int N = ham_.N(), nup = ham_.nup(), ndown = ham_.ndown();
  const bool closed = ham_.walker_type() == kClosed;
  const int nspin = closed ? 1 : 2;
  const long K = (long)nspin * N * N;
  if (!gmf_) {
    // trial density over the pairs p <= q (NOMSD.cpp:467-511): determinant q against
    // determinant p as the walker, weight Re(conj(c_q) c_p e^{ov}), ov the overlap of the
    // blocks (not doubled for CLOSED)
    z64* g = dnew<z64>(K);
    try {
      fill_z64(s, K, {0, 0}, g);
      z64* Gpq = scratch_as<z64>(ctx_, "msdv_G", K);
      z64* ovd = scratch_as<z64>(ctx_, "msdv_ov", 1);
      z64* Pa = scratch_as<z64>(ctx_, "msdv_pa", (size_t)N * std::max(nup, 1));
      z64* Pb = scratch_as<z64>(ctx_, "msdv_pb", (size_t)N * std::max(ndown, 1));
      if (!closed && (nup == 0 || ndown == 0)) fill_z64(s, K, {0, 0}, Gpq);
      double osum = 0.0;
      for (int p = 0; p < ndet_; ++p) {
        const z64* P = dets_[p].PsiT;
        if (nup > 0) conj_transpose(s, nup, N, P, Pa);
        if (!closed && ndown > 0) conj_transpose(s, ndown, N, P + (size_t)nup * N, Pb);
        for (int q = p; q < ndet_; ++q) {
          const z64* Q = dets_[q].PsiT;
          fill_z64(s, 1, {0, 0}, ovd);
          if (nup > 0)
            det_ops::mixed_density_matrix(ctx_, Q, nup, N, Pa, (long)N * nup, 1, Gpq, K, ovd,
                                          false, 1.0);
          if (!closed && ndown > 0)
            det_ops::mixed_density_matrix(ctx_, Q + (size_t)nup * N, ndown, N, Pb,
                                          (long)N * ndown, 1, Gpq + (size_t)N * N, K, ovd,
                                          false, 1.0);
          z64 o;
          AFQ_HIP(hipMemcpyAsync(&o, ovd, sizeof(z64), hipMemcpyDeviceToHost, s));
          AFQ_HIP(hipStreamSynchronize(s));
          const z64 cq = ci_h_[q], cp = ci_h_[p];
          // conj(c_q) c_p e^{o}
          const z64 a = {cq.re * cp.re + cq.im * cp.im, cq.re * cp.im - cq.im * cp.re};
          const double e = std::exp(o.re);
          const double scl = e * (a.re * std::cos(o.im) - a.im * std::sin(o.im));
          osum += scl;
          hipLaunchKernelGGL(k_axpy, grid_of(K), dim3(kNT), 0, s, K, scl, Gpq, g);
          AFQ_LAUNCH_CHECK();
        }
      }
      AFQ_CHECK(osum != 0.0, kErrState, "the trial's weights sum to zero");
      z64* tmp = scratch_as<z64>(ctx_, "msdv_tmp", K);
      fill_z64(s, K, {0, 0}, tmp);
      hipLaunchKernelGGL(k_axpy, grid_of(K), dim3(kNT), 0, s, K, 1.0 / osum, g, tmp);
      AFQ_LAUNCH_CHECK();
      AFQ_HIP(hipMemcpyAsync(g, tmp, sizeof(z64) * K, hipMemcpyDeviceToDevice, s));
      AFQ_HIP(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipFree(g);
      throw;
    }
    gmf_ = g;
  }
  ham_.vbias_full(gmf_, 1, dt, v, kFP64);
  hipLaunchKernelGGL(k_zero_imag, grid_of(ham_.M()), dim3(kNT), 0, s, ham_.M(), v);
  AFQ_LAUNCH_CHECK();
}

}  // namespace afq
