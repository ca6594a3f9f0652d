// Walker-averaged density-matrix observables of one measurement, on the device.
//
//  * rdm_measure: OneRDM::measure (onerdm.hpp:43-70) and DiagonalTwoRDM::measure
//    (diagonal_twordm.hpp:40-81) over the per-walker density matrices G [W][nspin][N][N] that the
//    mixed estimator (full mixed density matrix of the current walkers) or back-propagation
//    (back-propagated trial against the anchors) produce, each divided by the sum of the
//    weights as MeasurementOutput does (Measurements.hpp:98-121).  The reference contracts with
//    gemv and tensor contractions; here one workgroup owns one output row (i, spin) of every
//    observable, its four waves take every fourth walker and the partial sums are combined in a
//    fixed order, so a measurement is deterministic.  Each walker's row i, column i and diagonal
//    are read once per workgroup (a few KB per walker, resident in L2 across the row blocks).
//  * reference_coefficients: the per-walker share of each reference determinant in a
//    multi-reference measurement (MixedEstimator.hpp:61-103, BackPropEstimator.hpp:62-120).
#include "afq_internal.hpp"

namespace afq {

namespace {

constexpr int kRdmT = 256;                   // 4 waves
constexpr int kRdmLanes = 64;                // output columns per pass
constexpr int kRdmGroups = kRdmT / kRdmLanes;

__device__ inline z64 cdiv(z64 a, z64 b) {
  const double d = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// grid (N, nsp).  one [nsp][N][N]; two [nblk][N][N] with the blocks (aaaa), (aabb) and, for
// COLLINEAR, (bbbb); the (aabb) block pairs spin 0 with spin `dn` (0 CLOSED, 1 COLLINEAR).
__global__ void __launch_bounds__(kRdmT)
    k_rdm_measure(int W, int nsp, int N, int dn, const z64* __restrict__ G, long g_bs,
                  const z64* __restrict__ weight, const z64* __restrict__ coeff, int accumulate,
                  z64* __restrict__ one, z64* __restrict__ two) {
  const int i = blockIdx.x, s = blockIdx.y;
  const int lane = threadIdx.x % kRdmLanes, grp = threadIdx.x / kRdmLanes;
  __shared__ z64 red[kRdmGroups][kRdmLanes][3];
  __shared__ z64 wsum[kRdmT];
  // denominator: sum of the weights (not of weight x coefficient), fixed order
  z64 d = {0.0, 0.0};
  for (int w = threadIdx.x; w < W; w += kRdmT) d = cadd(d, weight[w]);
  wsum[threadIdx.x] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    z64 t = {0.0, 0.0};
    for (int k = 0; k < kRdmT; ++k) t = cadd(t, wsum[k]);
    wsum[0] = t;
  }
  __syncthreads();
  const z64 den = wsum[0];
  const long NN = (long)N * N;
  const bool pair = (s == 0) && two;  // this row also owns the (aabb) row
  for (int j0 = 0; j0 < N; j0 += kRdmLanes) {
    const int j = j0 + lane;
    z64 a1 = {0, 0}, a2 = {0, 0}, a3 = {0, 0};
    if (j < N) {
      for (int w = grp; w < W; w += kRdmGroups) {
        z64 c = weight[w];
        if (coeff) c = cmul(c, coeff[w]);
        const z64* Gs = G + (long)w * g_bs + s * NN;
        const z64 gij = Gs[(long)i * N + j];
        a1 = cadd(a1, cmul(c, gij));
        if (two) {
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
      // accumulate: add to the outputs (a sum over references, one launch per reference)
      if (one) {
        const z64 v = cdiv(t1, den);
        one[s * NN + o] = accumulate ? cadd(one[s * NN + o], v) : v;
      }
      if (two) {
        const z64 v = cdiv(t2, den);
        two[(2 * s) * NN + o] = accumulate ? cadd(two[(2 * s) * NN + o], v) : v;
        if (pair) {
          const z64 u = cdiv(t3, den);
          two[NN + o] = accumulate ? cadd(two[NN + o], u) : u;
        }
      }
    }
    __syncthreads();
  }
}

// one thread per walker; the largest real part is taken out before exponentiating, which
// cancels in the ratio and keeps every term finite
__global__ void k_ref_coeff(int W, int R, const z64* __restrict__ lo, const z64* __restrict__ cc,
                            const z64* __restrict__ ad, z64* __restrict__ coeff) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  auto expo = [&](int d) {
    z64 e = lo[(long)d * W + w];
    if (ad) {
      const z64 a = ad[(long)w * R + d];
      e = {e.re + a.re, e.im - a.im};
    }
    return e;
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
                            const z64* ad, z64* coeff) {
  AFQ_CHECK(W > 0 && R > 0, kErrShape, "reference coefficients shape");
  hipLaunchKernelGGL(k_ref_coeff, dim3(ceil_div(W, 256)), dim3(256), 0, s, W, R, lo, cc, ad, coeff);
  AFQ_LAUNCH_CHECK();
}

void rdm_measure(hipStream_t s, int W, int nsp, int N, bool collinear, const z64* G, long g_bs,
                 const z64* weight, const z64* coeff, bool accumulate, z64* one, z64* two) {
  AFQ_CHECK(W > 0 && N > 0 && (nsp == 1 || nsp == 2) && (nsp == 2) == collinear, kErrShape,
            "density-matrix measurement shape");
  AFQ_CHECK(g_bs >= (long)nsp * N * N, kErrShape, "density-matrix batch stride");
  if (!one && !two) return;
  hipLaunchKernelGGL(k_rdm_measure, dim3(N, nsp), dim3(kRdmT), 0, s, W, nsp, N, collinear ? 1 : 0,
                     G, g_bs, weight, coeff, accumulate ? 1 : 0, one, two);
  AFQ_LAUNCH_CHECK();
}

}  // namespace afq
