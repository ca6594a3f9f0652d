// Batched small dense kernels: one wavefront per matrix.
//
//  * lu_inverse_logdet: Gauss-Jordan inversion with partial pivoting (pivot = first maximum of
//    |re| + |im|, the getrf choice), producing the inverse and log det of each n x n complex
//    matrix.  The pivots are exactly getrf's U diagonal, so the log det follows
//    log_determinant_from_getrf (determinants_impl.hpp:33-75): sum_i log(+-U_ii + DBL_MIN) with
//    the sign flipped on swapped rows, added to the running value, imaginary part wrapped into
//    [-pi, pi].  A singular matrix yields a zero inverse (getri_or_zero,
//    nda_functions.hpp:422-437).
//  * chol_upper_inverse: Cholesky S = R^H R (R upper, positive real diagonal) of a Hermitian
//    positive-definite Gram matrix and R^{-1}; the building block of the Cholesky-QR2
//    orthogonalization.
//
// n <= 64.  All LDS traffic of one workgroup is program-ordered, and explicit barriers separate
// the phases that read another lane's data.
#include "afq_internal.hpp"

namespace afq {

namespace {

constexpr double kPi = 3.14159265358979323846;

// log of a complex number (principal branch)
__device__ inline void clog(double re, double im, double& lr, double& li) {
  lr = log(hypot(re, im));
  li = atan2(im, re);
}

// lu_inverse_logdet: one wavefront per matrix, lane r holds row r in registers (n <= 64).
// Rows are never moved: each lane tracks the position its row would occupy in getrf's swapped
// order, the pivot of step j is the unused row of largest |re|+|im| in column j with ties going
// to the lowest position (getrf's first maximum), and the row at position j trades positions
// with it. Gauss-Jordan then eliminates column j from every other row with the pivot row read
// from the pivot lane's registers (readlane), so the only dynamically indexed register is
// column j. Without the inverse, columns <= j are dead after step j and are not updated. The
// log det
// takes log(+-U_jj + DBL_MIN) with the sign flipped exactly on getrf's swapped steps; the
// in-place result Y gives A^{-1}[k][c] = Y[p_k][s_c] (p_k = pivot row of step k, s_c = step at
// which row c pivoted), applied while storing.
__device__ inline double rdl(double v, int src) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_readlane(lo, src);
  hi = __builtin_amdgcn_readlane(hi, src);
  return __hiloint2double(hi, lo);
}

// ovlp[b] = scale * wrap(ovlp[b] + (lr, li)) + add, the imaginary part wrapped into [-pi, pi]
__device__ inline void ov_update(z64* ovlp, long b, double lr, double li, double ov_scale,
                                 double ov_add_re, double ov_add_im) {
  z64 o = ovlp[b];
  double re = o.re + lr, im = o.im + li;
  if (im > kPi) {
    while (im > kPi) im -= 2 * kPi;
  } else if (im < -kPi) {
    while (im < -kPi) im += 2 * kPi;
  }
  ovlp[b] = {ov_scale * re + ov_add_re, ov_scale * im + ov_add_im};
}

// keep_ldet != nullptr: the raw log det sum (before the update of ovlp) of matrix b is stored
// there.  reuse (INV only): keepT holds the matrix whose inverse (at Ainv) and raw log det (at
// keep_ldet) were kept; when matrix b matches it bit for bit (no NaN), the kept values are
// exact for it: ovlp is updated from keep_ldet and nothing else is written.  Otherwise keepT
// takes matrix b and the factors are recomputed into Ainv and keep_ldet.
// NOPIV: no pivoting, for upper triangular matrices only (the Cholesky factor's inverse): row j
// is zero left of column j at step j, so columns < j are not updated.
template <int NMAX, bool INV, bool NOPIV = false>
__global__ void __launch_bounds__(64)
    k_lu_rows(const z64* __restrict__ A, long a_bs, int n, z64* __restrict__ Ainv, long i_bs,
              z64* __restrict__ ovlp, double ov_scale, double ov_add_re, double ov_add_im,
              z64* __restrict__ keepT, z64* __restrict__ keep_ldet, int reuse) {
  __shared__ int pivrow[64];
  const int lane = threadIdx.x;
  const long b = blockIdx.x;
  const z64* Ab = A + b * a_bs;
  double xr[NMAX], xi[NMAX];
#pragma unroll
  for (int c = 0; c < NMAX; ++c) {
    z64 v = (lane < n && c < n) ? Ab[(long)lane * n + c] : z64{0.0, 0.0};
    xr[c] = v.re;
    xi[c] = v.im;
  }
  // (rows of 64 columns fill the register file: no kept factors there)
  if constexpr (INV && !NOPIV && NMAX <= 48) {
    if (reuse) {
      z64* Kb = keepT + b * a_bs;
      bool same = true;
      if (lane < n) {
#pragma unroll
        for (int c = 0; c < NMAX; ++c) {
          const z64 k = c < n ? Kb[(long)lane * n + c] : z64{xr[c], xi[c]};
          same = same & (xr[c] == xr[c]) & (xi[c] == xi[c]) &
                 (__double_as_longlong(k.re) == __double_as_longlong(xr[c])) &
                 (__double_as_longlong(k.im) == __double_as_longlong(xi[c]));
        }
      }
      if (__ballot(!same) == 0) {  // uniform
        if (ovlp != nullptr && lane == 0) {
          const z64 l = keep_ldet[b];
          ov_update(ovlp, b, l.re, l.im, ov_scale, ov_add_re, ov_add_im);
        }
        return;
      }
      if (lane < n) {
#pragma unroll
        for (int c = 0; c < NMAX; ++c)
          if (c < n) Kb[(long)lane * n + c] = z64{xr[c], xi[c]};
      }
    }
  }
  bool used = lane >= n;
  int pos = lane;         // position of my row in getrf's order
  int mystep = n;         // step at which my row pivoted
  double ur = 1.0, ui = 0.0;  // my pivot value with getrf's sign (+-U_jj)
  bool singular = false;
  // Steps j = CH * J + jj: the chunk J is a loop variable, jj is unrolled, so column j of the
  // register arrays is selected among NMAX / CH candidates and the column update needs a
  // select only on the NMAX / CH columns congruent to jj.
  constexpr int CH = NMAX < 16 ? NMAX : (NMAX % 16 == 0 ? 16 : 8);
  constexpr int NCH = NMAX / CH;
  static_assert(NMAX % CH == 0, "NMAX must be a multiple of the chunk width");
  for (int J = 0; J * CH < n; ++J) {
#pragma unroll
    for (int jj = 0; jj < CH; ++jj) {
      const int j = J * CH + jj;
      if (j >= n) break;
      double fr = 0.0, fi = 0.0;  // my row's entry in column j
#pragma unroll
      for (int JJ = 0; JJ < NCH; ++JJ)
        if (JJ == J) { fr = xr[JJ * CH + jj]; fi = xi[JJ * CH + jj]; }
      // pivot: max |re|+|im| over unused rows, ties to the lowest position
      // (a NaN entry ranks below every number but above used rows, so the reduction stays
      // well defined and uniform)
      double mag = fabs(fr) + fabs(fi);
      double best = used ? -1.0 : (mag == mag ? mag : -0.5);
      int bpos = used ? 1 << 20 : pos;
      if (NOPIV) {
        best = (pos == j) ? 1.0 : -1.0;
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        double ob = __shfl_xor(best, off, 64);
        int op = __shfl_xor(bpos, off, 64);
        if (ob > best || (ob == best && op < bpos)) { best = ob; bpos = op; }
      }
      // the pivot row is the lane at position bpos; the row at position j moves to bpos's slot
      const unsigned long long owner = __ballot(!used && pos == bpos);
      const int p = __builtin_amdgcn_readfirstlane(owner ? __ffsll(owner) - 1 : j);
      const bool swapped = bpos != j;
      if (lane == p) {
        pos = j;
      } else if (!used && pos == j) {
        pos = bpos;
      }
      const double dr = rdl(fr, p), di = rdl(fi, p);
      if (lane == p) {
        used = true;
        mystep = j;
        ur = swapped ? -dr : dr;
        ui = swapped ? -di : di;
      }
      if (lane == 0) pivrow[j] = p;
      if (dr == 0.0 && di == 0.0) {
        singular = true;
        continue;  // uniform
      }
      const double den = dr * dr + di * di;
      const double ir = dr / den, ii = -di / den;
      // h = g / d with g = d - 1 on the pivot row (which becomes the pivot row / d), f elsewhere
      const double gr = (lane == p) ? dr - 1.0 : fr, gi = (lane == p) ? di : fi;
      const double hr = gr * ir - gi * ii, hi = gr * ii + gi * ir;
      // column j of the result: 1/d on the pivot row, -f/d elsewhere
      const double njr = (lane == p) ? ir : -(fr * ir - fi * ii);
      const double nji = (lane == p) ? ii : -(fr * ii + fi * ir);
      // the pivot row is read from the pivot lane's registers (scalar reads of one lane, no
      // LDS round trip on the step's critical path)
#pragma unroll
      for (int c = 0; c < NMAX; ++c) {
        if (!INV && c <= j) continue;  // uniform: dead columns of the log det alone
        if (NOPIV && c < j) continue;  // uniform: zero in the pivot row
        const double pr = rdl(xr[c], p), pim = rdl(xi[c], p);
        // x - h p as four fused multiply-adds
        double nr = fma(hi, pim, fma(-hr, pr, xr[c]));
        double ni = fma(-hi, pr, fma(-hr, pim, xi[c]));
        if (c % CH == jj && c / CH == J) { nr = njr; ni = nji; }
        xr[c] = nr;
        xi[c] = ni;
      }
    }
  }
  if (ovlp != nullptr || keep_ldet != nullptr) {
    double lr = 0.0, li = 0.0;
    if (lane < n) clog(ur + 2.2250738585072014e-308, ui, lr, li);
    lr = wave_sum(lr);
    li = wave_sum(li);
    if (lane == 0) {
      if (keep_ldet != nullptr) keep_ldet[b] = {lr, li};
      if (ovlp != nullptr) ov_update(ovlp, b, lr, li, ov_scale, ov_add_re, ov_add_im);
    }
  }
  if constexpr (!INV) return;
  __syncthreads();
  z64* Ib = Ainv + b * i_bs;
  if (lane < n && mystep < n) {
#pragma unroll
    for (int m = 0; m < NMAX; ++m) {
      const int col = m < n ? pivrow[m] : -1;
      if (col >= 0 && col < n)
        Ib[(long)mystep * n + col] = singular ? z64{0.0, 0.0} : z64{xr[m], xi[m]};
    }
  }
}

constexpr int kLuThreads = 256;
// Workgroup Cholesky: S = R^H R with R upper triangular (positive real diagonal), written to R
// with zeros below the diagonal.
__global__ void __launch_bounds__(kLuThreads)
    k_chol_blk(const z64* __restrict__ S, int n, z64* __restrict__ R) {
  extern __shared__ double2 sm2[];
  const int ld = n + 1;
  double2* a = sm2;
  const int tid = threadIdx.x;
  const long b = blockIdx.x;
  const z64* Sb = S + b * (long)n * n;
  for (int e = tid; e < n * n; e += kLuThreads) {
    z64 v = Sb[e];
    a[(e / n) * ld + e % n] = make_double2(v.re, v.im);
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    const double d = sqrt(a[j * ld + j].x);
    __syncthreads();
    for (int c = tid; c < n; c += kLuThreads) {
      if (c == j) {
        a[j * ld + c] = make_double2(d, 0.0);
      } else if (c > j) {
        double2 x = a[j * ld + c];
        a[j * ld + c] = make_double2(x.x / d, x.y / d);
      }
    }
    __syncthreads();
    // A[i][c] -= conj(R[j][i]) R[j][c] for j < i <= c
    const int m = n - j - 1;
    for (int e = tid; e < m * m; e += kLuThreads) {
      const int i = j + 1 + e / m, c = j + 1 + e % m;
      if (c < i) continue;
      const double2 ri = a[j * ld + i], rc = a[j * ld + c];
      double2 x = a[i * ld + c];
      x.x -= ri.x * rc.x + ri.y * rc.y;
      x.y -= ri.x * rc.y - ri.y * rc.x;
      a[i * ld + c] = x;
    }
    __syncthreads();
  }
  z64* Rb = R + b * (long)n * n;
  for (int e = tid; e < n * n; e += kLuThreads) {
    const int r = e / n, c = e % n;
    double2 v = a[r * ld + c];
    Rb[e] = (c >= r) ? z64{v.x, v.y} : z64{0.0, 0.0};
  }
}

// Cholesky of a Hermitian positive-definite S (upper R, S = R^H R) and the inverse of R.
// Rinv is written as a full n x n matrix (zeros below the diagonal). ldet += sum_i log R_ii.
__global__ void __launch_bounds__(64)
    k_chol_upper_inverse(const z64* __restrict__ S, int n, z64* __restrict__ Rinv,
                         z64* __restrict__ ldet) {
  extern __shared__ double smem[];
  const int ld = n + 1;
  double* a = smem;  // [n][ld] complex
  const int lane = threadIdx.x;
  const long b = blockIdx.x;
  const z64* Sb = S + b * (long)n * n;
  for (int e = lane; e < n * n; e += 64) {
    int r = e / n, c = e % n;
    z64 v = Sb[e];
    a[2 * (r * ld + c)] = v.re;
    a[2 * (r * ld + c) + 1] = v.im;
  }
  __syncthreads();
  double lsum = 0.0;
  // right-looking upper Cholesky: row j of R
  for (int j = 0; j < n; ++j) {
    double d = sqrt(a[2 * (j * ld + j)]);
    lsum += log(d);
    __syncthreads();
    for (int c = lane; c < n; c += 64) {
      if (c == j) {
        a[2 * (j * ld + c)] = d;
        a[2 * (j * ld + c) + 1] = 0.0;
      } else if (c > j) {
        a[2 * (j * ld + c)] /= d;
        a[2 * (j * ld + c) + 1] /= d;
      }
    }
    __syncthreads();
    // trailing update: A[i][c] -= conj(R[j][i]) R[j][c] for j < i <= c
    for (int c = lane; c < n; c += 64) {
      if (c <= j) continue;
      double rcr = a[2 * (j * ld + c)], rci = a[2 * (j * ld + c) + 1];
      for (int i = j + 1; i <= c; ++i) {
        double rir = a[2 * (j * ld + i)], rii = -a[2 * (j * ld + i) + 1];  // conj(R[j][i])
        a[2 * (i * ld + c)] -= rir * rcr - rii * rci;
        a[2 * (i * ld + c) + 1] -= rir * rci + rii * rcr;
      }
    }
    __syncthreads();
  }
  // inverse of the upper triangle, lane = column c: X[c][c] = 1/R[c][c];
  // X[i][c] = -(sum_{k=i+1..c} R[i][k] X[k][c]) / R[i][i]
  // (each lane writes and re-reads only its own column of Rinv in global memory)
  z64* Rb = Rinv + b * (long)n * n;
  for (int c = lane; c < n; c += 64) {
    for (int i = c + 1; i < n; ++i) Rb[(long)i * n + c] = {0.0, 0.0};
    Rb[(long)c * n + c] = {1.0 / a[2 * (c * ld + c)], 0.0};
    for (int i = c - 1; i >= 0; --i) {
      double sr = 0.0, si = 0.0;
      for (int k = i + 1; k <= c; ++k) {
        double rr = a[2 * (i * ld + k)], ri = a[2 * (i * ld + k) + 1];
        z64 x = Rb[(long)k * n + c];
        sr += rr * x.re - ri * x.im;
        si += rr * x.im + ri * x.re;
      }
      double rd = a[2 * (i * ld + i)];
      Rb[(long)i * n + c] = {-sr / rd, -si / rd};
    }
  }
  if (lane == 0 && ldet != nullptr) {
    z64 o = ldet[b];
    ldet[b] = {o.re + lsum, o.im};
  }
}

}  // namespace

void lu_inverse_logdet(hipStream_t s, const z64* A, long a_bs, int nb, int n, z64* Ainv,
                       long i_bs, z64* ovlp, double ov_scale, z64 ov_add, bool want_inverse,
                       const LuKeep* keep, bool reuse) {
  AFQ_CHECK(n >= 1 && n <= 64, kErrShape, "n must be in [1, 64]");
  AFQ_CHECK(!keep || (want_inverse && n <= 48 && Ainv == keep->Tinv && a_bs == (long)n * n &&
                      i_bs == (long)n * n && (reuse || A == keep->T)),
            kErrArg, "kept factors: n <= 48, the inverse in the kept buffer, [nb][n][n] layouts");
  if (nb <= 0) return;
  const int wi = want_inverse ? 1 : 0;
  z64* kT = keep && reuse ? keep->T : nullptr;
  z64* kl = keep ? keep->ldet : nullptr;
  const int ru = keep && reuse ? 1 : 0;
#define AFQ_LU_ROWS(NM)                                                                       \
  do {                                                                                        \
    if (wi)                                                                                   \
      hipLaunchKernelGGL((k_lu_rows<NM, true>), dim3(nb), dim3(64), 0, s, A, a_bs, n, Ainv,   \
                         i_bs, ovlp, ov_scale, ov_add.re, ov_add.im, kT, kl, ru);             \
    else                                                                                      \
      hipLaunchKernelGGL((k_lu_rows<NM, false>), dim3(nb), dim3(64), 0, s, A, a_bs, n, Ainv,  \
                         i_bs, ovlp, ov_scale, ov_add.re, ov_add.im, kT, kl, ru);             \
  } while (0)
  if (n <= 8)
    AFQ_LU_ROWS(8);
  else if (n <= 16)
    AFQ_LU_ROWS(16);
  else if (n <= 32)
    AFQ_LU_ROWS(32);
  else if (n <= 40)
    AFQ_LU_ROWS(40);
  else if (n <= 48)
    AFQ_LU_ROWS(48);
  else
    AFQ_LU_ROWS(64);
  AFQ_LAUNCH_CHECK();
}

void chol_upper_inverse(hipStream_t s, const z64* S, int nb, int n, z64* Rinv, z64* ldet) {
  AFQ_CHECK(n >= 1 && n <= 64, kErrShape, "n must be in [1, 64]");
  if (nb <= 0) return;
  size_t lds = sizeof(double2) * ((size_t)n * (n + 1) + n);
  AFQ_CHECK(lds <= 160 * 1024, kErrShape, "matrix too large for one workgroup's LDS");
  if (lds > 65536)  // n = 64
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_chol_blk),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  // R is written into Rinv, then inverted in place (upper triangular: no pivoting needed);
  // the inversion's log det is sum_i log R_ii (R_ii real and positive).
  hipLaunchKernelGGL(k_chol_blk, dim3(nb), dim3(kLuThreads), lds, s, S, n, Rinv);
  AFQ_LAUNCH_CHECK();
  const long nn = (long)n * n;
  if (n <= 8)
    hipLaunchKernelGGL((k_lu_rows<8, true, true>), dim3(nb), dim3(64), 0, s, Rinv, nn, n, Rinv, nn,
                       ldet, 1.0, 0.0, 0.0, nullptr, nullptr, 0);
  else if (n <= 16)
    hipLaunchKernelGGL((k_lu_rows<16, true, true>), dim3(nb), dim3(64), 0, s, Rinv, nn, n, Rinv, nn,
                       ldet, 1.0, 0.0, 0.0, nullptr, nullptr, 0);
  else if (n <= 32)
    hipLaunchKernelGGL((k_lu_rows<32, true, true>), dim3(nb), dim3(64), 0, s, Rinv, nn, n, Rinv, nn,
                       ldet, 1.0, 0.0, 0.0, nullptr, nullptr, 0);
  else if (n <= 40)
    hipLaunchKernelGGL((k_lu_rows<40, true, true>), dim3(nb), dim3(64), 0, s, Rinv, nn, n, Rinv, nn,
                       ldet, 1.0, 0.0, 0.0, nullptr, nullptr, 0);
  else if (n <= 48)
    hipLaunchKernelGGL((k_lu_rows<48, true, true>), dim3(nb), dim3(64), 0, s, Rinv, nn, n, Rinv, nn,
                       ldet, 1.0, 0.0, 0.0, nullptr, nullptr, 0);
  else
    hipLaunchKernelGGL((k_lu_rows<64, true, true>), dim3(nb), dim3(64), 0, s, Rinv, nn, n, Rinv, nn,
                       ldet, 1.0, 0.0, 0.0, nullptr, nullptr, 0);
  AFQ_LAUNCH_CHECK();
}

}  // namespace afq
