// Walker-averaged density-matrix observables of one measurement, on the device.
//
//  * rdm_measure: OneRDM::measure (onerdm.hpp:43-70) and DiagonalTwoRDM::measure
//    (diagonal_twordm.hpp:40-81) over the per-walker density matrices G [W][nspin][N][N] that the
//    mixed estimator (full mixed density matrix of the current walkers) or back-propagation
//    (back-propagated trial against the anchors) produce, each divided by the sum of the
//    weights as MeasurementOutput does (Measurements.hpp:98-121).  The reference contracts with
//    gemv and tensor contractions; here a workgroup owns one output row (i, spin) of every
//    observable for one block of walkers (about 1024 workgroups in all), its four waves take
//    every fourth walker, and the walker blocks' partial sums are added in a fixed order by a
//    second kernel, so a measurement is deterministic.  Each walker's row i, column i and
//    diagonal are read once per workgroup (a few KB per walker, resident in L2 across rows).
//  * reference_coefficients: the per-walker share of each reference determinant in a
//    multi-reference measurement (MixedEstimator.hpp:61-103, BackPropEstimator.hpp:62-120).
#include "ops.hpp"

namespace afq {

namespace {

constexpr int kRdmT = 256;                   // 4 waves
constexpr int kRdmLanes = 64;                // output columns per pass
constexpr int kRdmGroups = kRdmT / kRdmLanes;

__device__ inline z64 cdiv(z64 a, z64 b) {
  const double d = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// den[0] = sum_w weight[w], one workgroup, fixed order
__global__ void __launch_bounds__(kRdmT) k_weight_sum(int W, const z64* __restrict__ weight,
                                                      z64* __restrict__ den) {
  __shared__ z64 part[kRdmT];
  z64 d = {0.0, 0.0};
  for (int w = threadIdx.x; w < W; w += kRdmT) d = cadd(d, weight[w]);
  part[threadIdx.x] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    z64 t = {0.0, 0.0};
    for (int k = 0; k < kRdmT; ++k) t = cadd(t, part[k]);
    den[0] = t;
  }
}

// grid (N, nsp, WB).  Workgroup (i, s, b) sums walkers [b wbs, min(W, (b + 1) wbs)) into row i
// of the partial block b: [one: nsp][N][N] then [two: nsp + 1][N][N], unnormalized.  The
// (aabb) block pairs spin 0 with spin `dn` (0 CLOSED, 1 COLLINEAR).
__global__ void __launch_bounds__(kRdmT)
    k_rdm_partial(int W, int wbs, int wmod, int nsp, int N, int dn, const z64* __restrict__ G,
                  long g_bs, const z64* __restrict__ weight, const z64* __restrict__ coeff,
                  int want_two, z64* __restrict__ part) {
  const int i = blockIdx.x, s = blockIdx.y, b = blockIdx.z;
  const int lane = threadIdx.x % kRdmLanes, grp = threadIdx.x / kRdmLanes;
  __shared__ z64 red[kRdmGroups][kRdmLanes][3];
  const long NN = (long)N * N;
  const long P = (long)(2 * nsp + 1) * NN;
  z64* out = part + (long)b * P;
  const int w0 = b * wbs, w1 = min(W, w0 + wbs);
  const bool pair = (s == 0) && want_two;  // this row also owns the (aabb) row
  for (int j0 = 0; j0 < N; j0 += kRdmLanes) {
    const int j = j0 + lane;
    z64 a1 = {0, 0}, a2 = {0, 0}, a3 = {0, 0};
    if (j < N) {
      for (int w = w0 + grp; w < w1; w += kRdmGroups) {
        z64 c = weight[w % wmod];
        if (coeff) c = cmul(c, coeff[w]);
        const z64* Gs = G + (long)w * g_bs + s * NN;
        const z64 gij = Gs[(long)i * N + j];
        a1 = cadd(a1, cmul(c, gij));
        if (want_two) {
          const z64 gii = Gs[(long)i * N + i], gjj = Gs[(long)j * N + j], gji = Gs[(long)j * N + i];
          const z64 ci = cmul(c, gii);
          const z64 cij = cmul(c, gij);
          // G_ii G_jj - G_ij G_ji
          const z64 t1 = cmul(ci, gjj), t2 = cmul(cij, gji);
          a2 = {a2.re + t1.re - t2.re, a2.im + t1.im - t2.im};
          if (pair) {
            const z64 gdn = G[(long)w * g_bs + dn * NN + (long)j * N + j];
            a3 = cadd(a3, cmul(ci, gdn));
          }
        }
      }
    }
    red[grp][lane][0] = a1;
    red[grp][lane][1] = a2;
    red[grp][lane][2] = a3;
    __syncthreads();
    if (grp == 0 && j < N) {
      z64 t1 = red[0][lane][0], t2 = red[0][lane][1], t3 = red[0][lane][2];
      for (int g = 1; g < kRdmGroups; ++g) {
        t1 = cadd(t1, red[g][lane][0]);
        t2 = cadd(t2, red[g][lane][1]);
        t3 = cadd(t3, red[g][lane][2]);
      }
      const long o = (long)i * N + j;
      out[s * NN + o] = t1;
      if (want_two) {
        out[(nsp + 2 * s) * NN + o] = t2;
        if (pair) out[(nsp + 1) * NN + o] = t3;
      }
    }
    __syncthreads();
  }
}

// element t of [one][two]: the walker blocks' partial sums added in order, divided by den;
// accumulate adds to the output (a sum over references, one measurement per reference)
__global__ void k_rdm_finish(long n_one, long n_all, int WB, const z64* __restrict__ part,
                             const z64* __restrict__ den, int accumulate, z64* __restrict__ one,
                             z64* __restrict__ two) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_all) return;
  z64* dst = t < n_one ? (one ? one + t : nullptr) : (two ? two + (t - n_one) : nullptr);
  if (!dst) return;
  z64 v = {0.0, 0.0};
  for (int b = 0; b < WB; ++b) v = cadd(v, part[(long)b * n_all + t]);
  v = cdiv(v, den[0]);
  *dst = accumulate ? cadd(*dst, v) : v;
}

// one thread per walker; the largest real part is taken out before exponentiating, which
// cancels in the ratio and keeps every term finite
__global__ void k_ref_coeff(int W, int R, const z64* __restrict__ lo, const z64* __restrict__ cc,
                            const z64* __restrict__ ad, z64* __restrict__ coeff, double scale) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  auto expo = [&](int d) {
    z64 e = lo[(long)d * W + w];
    if (ad) {
      const z64 a = ad[(long)w * R + d];
      e = {e.re + a.re, e.im - a.im};
    }
    return z64{scale * e.re, scale * e.im};
  };
  double m = -HUGE_VAL;
  for (int d = 0; d < R; ++d) m = fmax(m, expo(d).re);
  z64 tot = {0.0, 0.0};
  for (int d = 0; d < R; ++d) {
    const z64 e = expo(d);
    const double r = exp(e.re - m);
    const z64 x = cmul(cc[d], z64{r * cos(e.im), r * sin(e.im)});
    coeff[(long)d * W + w] = x;
    tot = cadd(tot, x);
  }
  for (int d = 0; d < R; ++d) coeff[(long)d * W + w] = cdiv(coeff[(long)d * W + w], tot);
}

}  // namespace

void reference_coefficients(hipStream_t s, int W, int R, const z64* lo, const z64* cc,
                            const z64* ad, z64* coeff, double scale) {
  AFQ_CHECK(W > 0 && R > 0, kErrShape, "reference coefficients shape");
  hipLaunchKernelGGL(k_ref_coeff, dim3(ceil_div(W, 256)), dim3(256), 0, s, W, R, lo, cc, ad, coeff,
                     scale);
  AFQ_LAUNCH_CHECK();
}

void rdm_measure(Context& c, int W, int nsp, int N, bool collinear, const z64* G, long g_bs,
                 const z64* weight, const z64* coeff, bool accumulate, z64* one, z64* two,
                 int wmod) {
  if (wmod <= 0) wmod = W;
  AFQ_CHECK(wmod <= W, kErrShape, "weights of more walkers than density matrices");
  AFQ_CHECK(W > 0 && N > 0 && (nsp == 1 || nsp == 2) && (nsp == 2) == collinear, kErrShape,
            "density-matrix measurement shape");
  AFQ_CHECK(g_bs >= (long)nsp * N * N, kErrShape, "density-matrix batch stride");
  if (!one && !two) return;
  hipStream_t s = c.stream();
  // walker blocks: about 1024 workgroups in the partial pass, at least 8 walkers per block
  int WB = std::max(1, std::min(ceil_div(1024, (long)N * nsp), ceil_div(W, 8)));
  const int wbs = ceil_div(W, WB);
  WB = ceil_div(W, wbs);
  const long NN = (long)N * N, n_one = nsp * NN, n_all = (long)(2 * nsp + 1) * NN;
  z64* part = scratch_as<z64>(c, "rdm_part", (size_t)WB * n_all);
  z64* den = scratch_as<z64>(c, "rdm_den", 1);
  hipLaunchKernelGGL(k_weight_sum, dim3(1), dim3(kRdmT), 0, s, wmod, weight, den);
  AFQ_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_rdm_partial, dim3(N, nsp, WB), dim3(kRdmT), 0, s, W, wbs, wmod, nsp, N,
                     collinear ? 1 : 0, G, g_bs, weight, coeff, two ? 1 : 0, part);
  AFQ_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_rdm_finish, dim3(ceil_div(n_all, 256)), dim3(256), 0, s, n_one, n_all, WB,
                     part, den, accumulate ? 1 : 0, one, two);
  AFQ_LAUNCH_CHECK();
}

}  // namespace afq
