// Auxiliary fields, walker weights and the growth estimator, all on the device.
//
//  * construct_x: assemble_X (AFQMCBasePropagator.cpp:686-733) with construct_X_impl
//    (construct_X_kernel.hpp:34-86) for continuous charge fields, plus the mean-field factor
//    (AFQMCBasePropagator.cpp:335-336).  One workgroup per walker; the per-walker sums HW and
//    sum_m X vMF are block reductions in a fixed order (deterministic, unlike the reference's
//    atomics).  Uniforms are either injected (parity runs, like the reference's
//    DEVICE_RNG_FROM_HOST build) or drawn in-kernel from Philox4x32-10, a counter-based
//    generator keyed by (seed, stream, step, walker, field) so every walker draws independently
//    of launch shape.
//  * walker_update: WalkerSetUpdate.hpp:70-191 (hybrid, phaseless) followed by
//    bound_walker_weights (:199-218), one thread per walker; the reference runs this loop on the
//    host after staging ten W-vectors across PCIe every step.
//  * average_pseudo_energy: averagePseudoEnergy (average_energy.hpp:54-63) and the Eshift
//    relaxation of run_afqmc.cpp:91, so the step never needs a host round trip.
#include "afq_internal.hpp"

namespace afq {

namespace {

__device__ inline double probit(double p) {
  // Acklam's rational approximation of the inverse normal CDF (probit.h:58-137)
  const double a0 = -3.969683028665376e+01, a1 = 2.209460984245205e+02, a2 = -2.759285104469687e+02,
               a3 = 1.383577518672690e+02, a4 = -3.066479806614716e+01, a5 = 2.506628277459239e+00;
  const double b0 = -5.447609879822406e+01, b1 = 1.615858368580409e+02, b2 = -1.556989798598866e+02,
               b3 = 6.680131188771972e+01, b4 = -1.328068155288572e+01;
  const double c0 = -7.784894002430293e-03, c1 = -3.223964580411365e-01, c2 = -2.400758277161838e+00,
               c3 = -2.549732539343734e+00, c4 = 4.374664141464968e+00, c5 = 2.938163982698783e+00;
  const double d0 = 7.784695709041462e-03, d1 = 3.224671290700398e-01, d2 = 2.445134137142996e+00,
               d3 = 3.754408661907416e+00;
  if (p < 0 || p > 1) return 0.0;
  if (p == 0) return -HUGE_VAL;
  if (p == 1) return HUGE_VAL;
  if (p < 0.02425) {
    double q = sqrt(-2 * log(p));
    return (((((c0 * q + c1) * q + c2) * q + c3) * q + c4) * q + c5) /
           ((((d0 * q + d1) * q + d2) * q + d3) * q + 1);
  }
  if (p > 0.97575) {
    double q = sqrt(-2 * log(1 - p));
    return -(((((c0 * q + c1) * q + c2) * q + c3) * q + c4) * q + c5) /
           ((((d0 * q + d1) * q + d2) * q + d3) * q + 1);
  }
  double q = p - 0.5, r = q * q;
  return (((((a0 * r + a1) * r + a2) * r + a3) * r + a4) * r + a5) * q /
         (((((b0 * r + b1) * r + b2) * r + b3) * r + b4) * r + 1);
}

// Philox4x32-10 (Salmon et al., SC'11)
__device__ inline void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__device__ inline double field_uniform(uint64_t seed, uint32_t stream, uint32_t step, uint32_t w,
                                       uint32_t m) {
  uint32_t c[4] = {m >> 1, w, step, stream};
  philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  uint64_t bits = (m & 1u) ? (((uint64_t)c[2] << 32) | c[3]) : (((uint64_t)c[0] << 32) | c[1]);
  return ((double)(bits >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

template <int NT>
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) t += red[i];
  return t;  // valid on thread 0
}

constexpr int kXThreads = 256;

__global__ void __launch_bounds__(kXThreads)
    k_construct_x(const double* __restrict__ u, const z64* __restrict__ vMF, int M, FieldParams p,
                  z64* __restrict__ X, z64* __restrict__ HW, z64* __restrict__ mf,
                  c32* __restrict__ X32) {
  __shared__ double red[kXThreads / 64];
  const int w = blockIdx.x;
  const uint32_t step = p.step_ptr ? *p.step_ptr : p.step;
  double hwr = 0, hwi = 0, xsr = 0, xsi = 0, vv = 0;
  int hits = 0;
  for (int m = threadIdx.x; m < M; m += kXThreads) {
    z64 vm = vMF[m];
    double am = hypot(vm.re, vm.im);
    z64 vmt = am > p.bound ? z64{vm.re / am * p.bound, vm.im / am * p.bound} : vm;
    long idx = (long)w * M + m;
    z64 x = X[idx];
    double ax = hypot(x.re, x.im);
    hits += ax > p.bound ? 1 : 0;
    z64 vbt = ax > p.bound ? z64{x.re / ax * p.bound, x.im / ax * p.bound} : x;
    if (p.project_vbias) vbt.im = 0.0;
    // vdiff = i (vb_t - vmf_t)
    z64 vd = p.free_projection ? z64{0.0, 0.0} : z64{-(vbt.im - vmt.im), vbt.re - vmt.re};
    double uu = u ? u[idx] : field_uniform(p.seed, p.stream, step, (uint32_t)w, (uint32_t)m);
    z64 xn = {probit(uu) + vd.re, vd.im};
    X[idx] = xn;
    if (X32) X32[idx] = {(float)xn.re, (float)xn.im};
    // HW += -vdiff * (X - vdiff/2)
    z64 t = {xn.re - 0.5 * vd.re, xn.im - 0.5 * vd.im};
    hwr -= vd.re * t.re - vd.im * t.im;
    hwi -= vd.re * t.im + vd.im * t.re;
    // sum_m X vMF
    xsr += xn.re * vm.re - xn.im * vm.im;
    xsi += xn.re * vm.im + xn.im * vm.re;
    vv += vm.re * vm.re - vm.im * vm.im;  // Re(vMF^2)
  }
  hwr = block_sum<kXThreads>(hwr, red);
  hwi = block_sum<kXThreads>(hwi, red);
  xsr = block_sum<kXThreads>(xsr, red);
  xsi = block_sum<kXThreads>(xsi, red);
  vv = block_sum<kXThreads>(vv, red);
  if (p.stats) {
    const double h = block_sum<kXThreads>((double)hits, red);
    if (threadIdx.x == 0) {
      atomicAdd(&p.stats[0], (unsigned long long)M);
      if (h > 0) atomicAdd(&p.stats[1], (unsigned long long)h);
    }
  }
  if (threadIdx.x == 0) {
    HW[w] = {hwr, hwi};
    // mf = i * sum X vMF - 0.5 Re(sum vMF^2)
    mf[w] = {-xsi - 0.5 * vv, xsr};
  }
}

__global__ void k_field_uniforms(uint64_t seed, uint32_t stream, uint32_t step, int W, int M,
                                 double* u) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)W * M) return;
  int w = (int)(i / M), m = (int)(i % M);
  u[i] = field_uniform(seed, stream, step, w, m);
}

__device__ inline z64 cexp(z64 z) {
  double e = exp(z.re);
  return {e * cos(z.im), e * sin(z.im)};
}

__global__ void k_walker_update(int W, UpdateParams p, const double* __restrict__ Eshift_p,
                                const z64* __restrict__ nov, const z64* __restrict__ mf,
                                const z64* __restrict__ HW, z64* __restrict__ weight,
                                z64* __restrict__ phase, z64* __restrict__ eloc_p,
                                z64* __restrict__ ovlp, z64* __restrict__ theta,
                                z64* __restrict__ wfac, int hl, const uint32_t* __restrict__ step_ptr,
                                const z64* __restrict__ E) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool e_counted = false, e_up = false, e_lo = false, w_up = false;
  if (i < W) {
  const double Eshift = *Eshift_p;
  const double hi = Eshift + p.upper_cut * sqrt(2.0 / p.dt);
  const double lo = Eshift - p.lower_cut * sqrt(2.0 / p.dt);
  z64 old_ov = ovlp[i], old_e = eloc_p[i], nv = nov[i], m = mf[i], hw = HW[i];
  double scale = 1.0, dth = 0.0;
  if (!p.free_projection) {
    dth = nv.im - old_ov.im - m.im;
    scale = p.use_cp ? (cos(dth) > 0.0 ? 1.0 : 0.0) : fmax(0.0, cos(dth));
  }
  theta[i] = {dth, 0.0};
  // hybrid local energy: (mf - HW - (ov_new - ov_old)) / dt + E0; or the local energy of the
  // propagated walker (E1 + EXX + EJ, E0 inside E1)
  z64 e = p.local_energy
              ? z64{E[3 * i].re + E[3 * i + 1].re + E[3 * i + 2].re,
                    E[3 * i].im + E[3 * i + 1].im + E[3 * i + 2].im}
              : z64{(m.re - hw.re - (nv.re - old_ov.re)) / p.dt + p.E0,
                    (m.im - hw.im - (nv.im - old_ov.im)) / p.dt};
  z64 w = weight[i], ph = phase[i];
  z64 fac = {0.0, 0.0};  // path-restoration factor of the step
  if (!isfinite(e.re)) {
    w = {0.0, 0.0};
    ph = {0.0, 0.0};
    e = old_e;
  } else {
    if (!p.free_projection) {
      e_counted = true;
      e_up = e.re > hi;
      e_lo = !e_up && e.re < lo;
      e.re = fmin(fmax(e.re, lo), hi);
    }
    // the hybrid expression is the log of the step factor itself; a local energy enters
    // through the midpoint of the old and the new time slice
    const z64 ee = p.local_energy ? z64{0.5 * (e.re + old_e.re), 0.5 * (e.im + old_e.im)} : e;
    z64 wf = cexp({-p.dt * (ee.re - Eshift), -p.dt * ee.im});
    double aw = hypot(wf.re, wf.im);
    if (scale > 2.2250738585072014e-308) fac = {wf.re / aw / scale, wf.im / aw / scale};
    if (p.free_projection) {
      w = cmul(w, z64{scale * wf.re, scale * wf.im});
    } else {
      w = {w.re * scale * aw, w.im * scale * aw};
    }
    ph = cmul(ph, z64{wf.re / aw, wf.im / aw});
  }
  // bound_walker_weights
  double maxw = fmax(p.weight_floor, p.weight_frac * p.target_pop);
  double a = hypot(w.re, w.im);
  w_up = a > maxw;
  if (a > maxw) w = {w.re * maxw / a, w.im * maxw / a};
  weight[i] = w;
  phase[i] = ph;
  eloc_p[i] = e;
  ovlp[i] = nv;
  if (wfac) wfac[(long)i * hl + (*step_ptr) % (uint32_t)hl] = fac;
  }
  if (p.stats) {  // one atomic per wavefront and counter
    const unsigned long long c[4] = {__ballot(e_counted), __ballot(e_up), __ballot(e_lo),
                                     __ballot(w_up)};
    const unsigned long long nw = __ballot(i < W);
    if ((threadIdx.x & 63) == 0) {
      if (c[0]) atomicAdd(&p.stats[3], (unsigned long long)__popcll(c[0]));
      if (c[1]) atomicAdd(&p.stats[4], (unsigned long long)__popcll(c[1]));
      if (c[2]) atomicAdd(&p.stats[5], (unsigned long long)__popcll(c[2]));
      if (nw) atomicAdd(&p.stats[6], (unsigned long long)__popcll(nw));
      if (c[3]) atomicAdd(&p.stats[7], (unsigned long long)__popcll(c[3]));
    }
  }
}

__global__ void k_store_energies(int W, const z64* __restrict__ E, z64* e1, z64* exx, z64* ej,
                                 long bs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W) return;
  if (e1) e1[(long)i * bs] = E[3 * i];
  if (exx) exx[(long)i * bs] = E[3 * i + 1];
  if (ej) ej[(long)i * bs] = E[3 * i + 2];
}

__global__ void k_store_fields(int W, int M, const z64* __restrict__ X, z64* __restrict__ ring,
                               int nring, const uint32_t* __restrict__ step_ptr) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)W * M) return;
  const int w = (int)(t / M), m = (int)(t % M);
  const int slot = (int)((*step_ptr) % (uint32_t)nring);
  ring[((long)w * nring + slot) * M + m] = X[t];
}

__global__ void k_gather_fields(int W, int M, const z64* __restrict__ ring, int nring, int slot,
                                z64* __restrict__ X) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)W * M) return;
  const int w = (int)(t / M), m = (int)(t % M);
  X[t] = ring[((long)w * nring + slot) * M + m];
}

__global__ void k_restore_weights(int W, const z64* __restrict__ weight, const z64* __restrict__ wfac,
                                  int hl, uint32_t pos, int n, z64* __restrict__ out) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  z64 x = weight[w];
  for (int k = 0; k < n; ++k) {
    const int slot = (int)(((long)pos - 1 - k) % hl + hl) % hl;
    x = cmul(x, wfac[(long)w * hl + slot]);
  }
  out[w] = x;
}

constexpr int kAvgThreads = 1024;

__global__ void __launch_bounds__(kAvgThreads)
    k_average_pseudo_energy(int W, const z64* __restrict__ weight, const z64* __restrict__ eloc,
                            double dt, double relax, double* out, double* Eshift) {
  __shared__ double red[kAvgThreads / 64];
  double s0r = 0, s0i = 0, s1r = 0, s1i = 0;
  for (int i = threadIdx.x; i < W; i += kAvgThreads) {
    z64 w = weight[i], e = eloc[i];
    z64 x = cmul(w, cexp({dt * e.re, dt * e.im}));
    s0r += x.re; s0i += x.im; s1r += w.re; s1i += w.im;
  }
  s0r = block_sum<kAvgThreads>(s0r, red);
  s0i = block_sum<kAvgThreads>(s0i, red);
  s1r = block_sum<kAvgThreads>(s1r, red);
  s1i = block_sum<kAvgThreads>(s1i, red);
  if (threadIdx.x == 0) {
    // (s0 / s1).real
    double den = s1r * s1r + s1i * s1i;
    double q = (s0r * s1r + s0i * s1i) / den;
    double e = log(q) / dt;
    if (out) *out = e;
    if (relax > 0.0 && Eshift) *Eshift += relax * (e - *Eshift);
  }
}

// comb (population_control.hpp:34-64) and the single-rank slot plan of swap_walkers_async
// (:140-209).  One workgroup: the cumulative |w| is summed left to right by one thread (the
// reference's std::partial_sum), every tooth (ic + r) * total / W finds its walker by binary
// search (the first cumulative weight above it, or the last walker with weight), and one
// thread lays out the copies: walkers without offspring free their slots, refilled in walker
// order.  src[d] = walker copied into slot d (d itself when kept).  status[0] = 1 when the
// total weight collapsed (<= 1e-6); the population is then left unchanged.
constexpr int kCombThreads = 1024;

__global__ void __launch_bounds__(kCombThreads)
    k_comb_plan(int W, const z64* __restrict__ weight, const double* r_in, uint64_t seed,
                uint32_t stream, const uint32_t* step_ptr, double* cum, int* src, int* status) {
  extern __shared__ int off[];  // [W]
  __shared__ double tot_s, r_s;
  __shared__ int last_s, ok_s;
  const int tid = threadIdx.x;
  for (int i = tid; i < W; i += kCombThreads) off[i] = 0;
  if (tid == 0) {
    double acc = 0.0;
    int last = W - 1;
    for (int i = 0; i < W; ++i) {
      acc += hypot(weight[i].re, weight[i].im);
      cum[i] = acc;
    }
    while (last > 0 && hypot(weight[last].re, weight[last].im) <= 0.0) --last;
    tot_s = acc;
    last_s = last;
    ok_s = acc > 1e-6 ? 1 : 0;
    double r;
    if (r_in) {
      r = *r_in;
    } else {
      uint32_t c[4] = {0u, 0u, step_ptr ? *step_ptr : 0u, stream};
      philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
      uint64_t bits = ((uint64_t)c[0] << 32) | c[1];
      r = ((double)(bits >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
    r_s = r;
    status[0] = ok_s ? 0 : 1;
  }
  __syncthreads();
  if (!ok_s) {
    for (int i = tid; i < W; i += kCombThreads) src[i] = i;
    return;
  }
  const double total = tot_s, r = r_s;
  const int last = last_s;
  for (int ic = tid; ic < W; ic += kCombThreads) {
    const double comb = (ic + r) * total / W;
    int lo = 0, hi = W;  // first j with comb < cum[j]
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (comb < cum[mid]) hi = mid; else lo = mid + 1;
    }
    int j = lo < last ? lo : last;
    atomicAdd(&off[j], 1);
  }
  __syncthreads();
  if (tid == 0) {
    int next = 0;  // scan position for free slots
    for (int i = 0; i < W; ++i) src[i] = i;
    for (int i = 0; i < W; ++i)
      for (int c = 1; c < off[i]; ++c) {
        while (off[next] != 0) ++next;
        src[next++] = i;
      }
  }
}

// copy walker rows into the freed slots: Slater blocks and all properties
__global__ void k_copy_walkers(int W, const int* __restrict__ src, long rowA, long rowB, long bsA,
                               long bsB, z64* PhiA, z64* PhiB, z64* weight, z64* phase, z64* eloc,
                               z64* ovlp, z64* theta, WalkerRows extra) {
  const int d = blockIdx.x;
  const int s = src[d];
  if (s == d) return;
  for (long e = threadIdx.x; e < rowA; e += blockDim.x) PhiA[d * bsA + e] = PhiA[s * bsA + e];
  if (PhiB)
    for (long e = threadIdx.x; e < rowB; e += blockDim.x) PhiB[d * bsB + e] = PhiB[s * bsB + e];
#pragma unroll
  for (int r = 0; r < WalkerRows::kMax; ++r) {
    z64* q = extra.p[r];
    const long n = extra.n[r], bs = extra.bs[r] ? extra.bs[r] : n;
    if (q)
      for (long e = threadIdx.x; e < n; e += blockDim.x) q[d * bs + e] = q[s * bs + e];
  }
  if (threadIdx.x == 0) {
    weight[d] = weight[s];
    phase[d] = phase[s];
    eloc[d] = eloc[s];
    ovlp[d] = ovlp[s];
    theta[d] = theta[s];
  }
}

// every copy carries its parent's weight; set |w| = 1 keeping the phase
__global__ void k_unit_weights(int W, const int* status, z64* weight) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W || status[0] != 0) return;
  z64 w = weight[i];
  double a = hypot(w.re, w.im);
  weight[i] = {w.re / a, w.im / a};
}

__global__ void k_increment(uint32_t* c) {
  if (threadIdx.x == 0) *c += 1u;
}

}  // namespace

void increment_counter(hipStream_t s, uint32_t* c) {
  hipLaunchKernelGGL(k_increment, dim3(1), dim3(64), 0, s, c);
  AFQ_LAUNCH_CHECK();
}

void population_control(hipStream_t s, int W, const double* r_dev, uint64_t seed, uint32_t stream,
                        const uint32_t* step_ptr, long rowA, long rowB, long bsA, long bsB,
                        z64* PhiA, z64* PhiB, z64* weight, z64* phase, z64* eloc, z64* ovlp,
                        z64* theta, double* cum, int* src, int* status, const WalkerRows& extra) {
  if (W <= 0) return;
  AFQ_CHECK((size_t)W * sizeof(int) <= 65536, kErrShape, "population control supports W <= 16384");
  hipLaunchKernelGGL(k_comb_plan, dim3(1), dim3(kCombThreads), sizeof(int) * W, s, W, weight, r_dev,
                     seed, stream, step_ptr, cum, src, status);
  AFQ_LAUNCH_CHECK();
  AFQ_CHECK(bsA >= rowA && (!PhiB || bsB >= rowB), kErrShape, "walker stride below the block size");
  hipLaunchKernelGGL(k_copy_walkers, dim3(W), dim3(256), 0, s, W, src, rowA, rowB, bsA, bsB, PhiA,
                     PhiB, weight, phase, eloc, ovlp, theta, extra);
  AFQ_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_unit_weights, dim3(ceil_div(W, 256)), dim3(256), 0, s, W, status, weight);
  AFQ_LAUNCH_CHECK();
}

void construct_x(hipStream_t s, const double* u, const z64* vMF, int W, int M,
                 const FieldParams& p, z64* X, z64* HW, z64* mf, c32* X32) {
  if (W <= 0) return;
  hipLaunchKernelGGL(k_construct_x, dim3(W), dim3(kXThreads), 0, s, u, vMF, M, p, X, HW, mf, X32);
  AFQ_LAUNCH_CHECK();
}

void field_uniforms(hipStream_t s, uint64_t seed, uint32_t stream, uint32_t step, int W, int M,
                    double* u) {
  long n = (long)W * M;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_field_uniforms, dim3(ceil_div(n, 256)), dim3(256), 0, s, seed, stream, step,
                     W, M, u);
  AFQ_LAUNCH_CHECK();
}

void walker_update(hipStream_t s, int W, const UpdateParams& p, const double* Eshift,
                   const z64* new_ovlp, const z64* mf, const z64* HW, z64* weight, z64* phase,
                   z64* eloc, z64* ovlp, z64* theta, z64* wfac, int hl, const uint32_t* step_ptr,
                   const z64* E) {
  if (W <= 0) return;
  AFQ_CHECK(!wfac || (hl > 0 && step_ptr), kErrArg, "weight-factor ring needs a length and a step");
  AFQ_CHECK(!p.local_energy || E, kErrArg, "local-energy update needs the energies");
  hipLaunchKernelGGL(k_walker_update, dim3(ceil_div(W, 256)), dim3(256), 0, s, W, p, Eshift,
                     new_ovlp, mf, HW, weight, phase, eloc, ovlp, theta, wfac, hl, step_ptr, E);
  AFQ_LAUNCH_CHECK();
}

void store_energy_components(hipStream_t s, int W, const z64* E, z64* e1, z64* exx, z64* ej,
                             long bs) {
  if (W <= 0 || !(e1 || exx || ej)) return;
  hipLaunchKernelGGL(k_store_energies, dim3(ceil_div(W, 256)), dim3(256), 0, s, W, E, e1, exx, ej,
                     bs);
  AFQ_LAUNCH_CHECK();
}

void store_fields(hipStream_t s, int W, int M, const z64* X, z64* ring, int nring,
                  const uint32_t* step_ptr) {
  const long n = (long)W * M;
  if (n <= 0) return;
  AFQ_CHECK(nring > 0 && step_ptr, kErrArg, "field ring needs a length and a step");
  hipLaunchKernelGGL(k_store_fields, dim3(ceil_div(n, 256)), dim3(256), 0, s, W, M, X, ring, nring,
                     step_ptr);
  AFQ_LAUNCH_CHECK();
}

void gather_fields(hipStream_t s, int W, int M, const z64* ring, int nring, int slot, z64* X) {
  const long n = (long)W * M;
  if (n <= 0) return;
  AFQ_CHECK(slot >= 0 && slot < nring, kErrArg, "field slot out of range");
  hipLaunchKernelGGL(k_gather_fields, dim3(ceil_div(n, 256)), dim3(256), 0, s, W, M, ring, nring,
                     slot, X);
  AFQ_LAUNCH_CHECK();
}

void restore_weights(hipStream_t s, int W, const z64* weight, const z64* wfac, int hl, uint32_t pos,
                     int n, z64* out) {
  if (W <= 0) return;
  AFQ_CHECK(n == 0 || (wfac && hl > 0 && n <= hl), kErrArg, "restoration longer than the history");
  hipLaunchKernelGGL(k_restore_weights, dim3(ceil_div(W, 256)), dim3(256), 0, s, W, weight, wfac,
                     hl > 0 ? hl : 1, pos, n, out);
  AFQ_LAUNCH_CHECK();
}

void average_pseudo_energy(hipStream_t s, int W, const z64* weight, const z64* eloc, double dt,
                           double relax, double* out, double* Eshift) {
  hipLaunchKernelGGL(k_average_pseudo_energy, dim3(1), dim3(kAvgThreads), 0, s, W, weight, eloc,
                     dt, relax, out, Eshift);
  AFQ_LAUNCH_CHECK();
}

}  // namespace afq
