// Batched small dense kernels: one wavefront per matrix, the matrix held in LDS.
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
// Lanes own columns (n <= 64).  All LDS traffic of one wavefront is program-ordered, and the
// explicit barriers separate the phases that read another lane's column.
#include "afq_internal.hpp"

namespace afq {

namespace {

constexpr double kPi = 3.14159265358979323846;

__device__ inline double2 ld2(const double* p) { return *reinterpret_cast<const double2*>(p); }

__device__ inline double cabs1(double re, double im) { return fabs(re) + fabs(im); }

// log of a complex number (principal branch)
__device__ inline void clog(double re, double im, double& lr, double& li) {
  lr = log(hypot(re, im));
  li = atan2(im, re);
}

__global__ void __launch_bounds__(64)
    k_lu_inverse_logdet(const z64* __restrict__ A, long a_bs, int n, z64* __restrict__ Ainv,
                        long i_bs, z64* __restrict__ ovlp, double ov_scale, double ov_add_re,
                        double ov_add_im, int want_inverse) {
  extern __shared__ double smem[];
  const int ld = n + 1;  // complex elements per LDS row
  double* a = smem;                 // [n][ld] complex, interleaved
  double* fcol = smem + 2 * n * ld;  // [n] complex
  int* piv = reinterpret_cast<int*>(fcol + 2 * n);
  const int lane = threadIdx.x;
  const long b = blockIdx.x;
  const z64* Ab = A + b * a_bs;

  for (int e = lane; e < n * n; e += 64) {
    int r = e / n, c = e % n;
    z64 v = Ab[e];
    a[2 * (r * ld + c)] = v.re;
    a[2 * (r * ld + c) + 1] = v.im;
  }
  __syncthreads();

  double ldr = 0.0, ldi = 0.0;  // log det accumulated on every lane (uniform values)
  bool singular = false;
  for (int j = 0; j < n; ++j) {
    // pivot: first maximum of |re|+|im| over rows j..n-1 of column j
    double best = -1.0;
    int p = n;
    for (int r = j + lane; r < n; r += 64) {
      double v = cabs1(a[2 * (r * ld + j)], a[2 * (r * ld + j) + 1]);
      if (v > best) { best = v; p = r; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      double ob = __shfl_xor(best, off, 64);
      int op = __shfl_xor(p, off, 64);
      if (ob > best || (ob == best && op < p)) { best = ob; p = op; }
    }
    if (lane == 0) piv[j] = p;
    // swap rows j and p (lane = column)
    if (p != j) {
      for (int c = lane; c < n; c += 64) {
        double2 x = ld2(a + 2 * (j * ld + c)), y = ld2(a + 2 * (p * ld + c));
        a[2 * (j * ld + c)] = y.x; a[2 * (j * ld + c) + 1] = y.y;
        a[2 * (p * ld + c)] = x.x; a[2 * (p * ld + c) + 1] = x.y;
      }
    }
    __syncthreads();
    double dr = a[2 * (j * ld + j)], di = a[2 * (j * ld + j) + 1];
    {
      double sr = (p == j) ? dr : -dr, si = (p == j) ? di : -di;
      double lr, li;
      clog(sr + 2.2250738585072014e-308, si, lr, li);
      ldr += lr;
      ldi += li;
    }
    if (dr == 0.0 && di == 0.0) {
      singular = true;
      continue;
    }
    // inverse of the pivot
    double den = dr * dr + di * di;
    double ir = dr / den, ii = -di / den;
    __syncthreads();
    // scale pivot row: a[j][c] *= 1/d (c != j), a[j][j] = 1/d
    for (int c = lane; c < n; c += 64) {
      if (c == j) {
        a[2 * (j * ld + c)] = ir;
        a[2 * (j * ld + c) + 1] = ii;
      } else {
        double xr = a[2 * (j * ld + c)], xi = a[2 * (j * ld + c) + 1];
        a[2 * (j * ld + c)] = xr * ir - xi * ii;
        a[2 * (j * ld + c) + 1] = xr * ii + xi * ir;
      }
    }
    // column j multipliers (old values)
    for (int r = lane; r < n; r += 64) {
      fcol[2 * r] = a[2 * (r * ld + j)];
      fcol[2 * r + 1] = a[2 * (r * ld + j) + 1];
    }
    __syncthreads();
    for (int c = lane; c < n; c += 64) {
      double pr = a[2 * (j * ld + c)], pi = a[2 * (j * ld + c) + 1];  // scaled pivot row
      for (int r = 0; r < n; ++r) {
        if (r == j) continue;
        double fr = fcol[2 * r], fi = fcol[2 * r + 1];
        if (c == j) {
          // new a[r][j] = -f / d = -f * (1/d)
          a[2 * (r * ld + c)] = -(fr * ir - fi * ii);
          a[2 * (r * ld + c) + 1] = -(fr * ii + fi * ir);
        } else {
          a[2 * (r * ld + c)] -= fr * pr - fi * pi;
          a[2 * (r * ld + c) + 1] -= fr * pi + fi * pr;
        }
      }
    }
    __syncthreads();
  }

  if (lane == 0 && ovlp != nullptr) {
    z64 o = ovlp[b];
    double re = o.re + ldr, im = o.im + ldi;
    if (im > kPi) {
      while (im > kPi) im -= 2 * kPi;
    } else if (im < -kPi) {
      while (im < -kPi) im += 2 * kPi;
    }
    ovlp[b] = {ov_scale * re + ov_add_re, ov_scale * im + ov_add_im};
  }
  if (!want_inverse) return;
  __syncthreads();
  // undo the row interchanges as column interchanges, last pivot first
  if (!singular) {
    for (int j = n - 1; j >= 0; --j) {
      int p = piv[j];
      if (p != j) {
        for (int r = lane; r < n; r += 64) {
          double2 x = ld2(a + 2 * (r * ld + j)), y = ld2(a + 2 * (r * ld + p));
          a[2 * (r * ld + j)] = y.x; a[2 * (r * ld + j) + 1] = y.y;
          a[2 * (r * ld + p)] = x.x; a[2 * (r * ld + p) + 1] = x.y;
        }
      }
      __syncthreads();
    }
  }
  z64* Ib = Ainv + b * i_bs;
  for (int e = lane; e < n * n; e += 64) {
    int r = e / n, c = e % n;
    Ib[e] = singular ? z64{0.0, 0.0} : z64{a[2 * (r * ld + c)], a[2 * (r * ld + c) + 1]};
  }
}

// Register-resident variant of the same Gauss-Jordan inversion: lane c keeps column c of the
// matrix in VGPRs (NMAX >= n rows), multipliers and pivots are broadcast with readlane, and the
// row interchange is a predicated select over the (unrolled) rows.  Same pivots, same log det,
// same inverse as k_lu_inverse_logdet, without LDS traffic in the elimination.
__device__ inline double bcast(double v, int src) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_readlane(lo, src);
  hi = __builtin_amdgcn_readlane(hi, src);
  return __hiloint2double(hi, lo);
}

template <int NMAX>
__global__ void __launch_bounds__(64)
    k_lu_reg(const z64* __restrict__ A, long a_bs, int n, z64* __restrict__ Ainv, long i_bs,
             z64* __restrict__ ovlp, double ov_scale, double ov_add_re, double ov_add_im,
             int want_inverse) {
  __shared__ int piv[NMAX];
  const int lane = threadIdx.x;
  const long b = blockIdx.x;
  const z64* Ab = A + b * a_bs;
  double xr[NMAX], xi[NMAX];
#pragma unroll
  for (int r = 0; r < NMAX; ++r) {
    z64 v = (lane < n && r < n) ? Ab[r * n + lane] : z64{0.0, 0.0};
    xr[r] = v.re;
    xi[r] = v.im;
  }
  double ur = 1.0, ui = 0.0;  // pivot of step `lane` with getrf's sign, logged after the loop
  bool singular = false;
  for (int j = 0; j < n; ++j) {
    double best = -1.0;
    int p = j;
#pragma unroll
    for (int r = 0; r < NMAX; ++r)
      if (r >= j && r < n) {
        double v = fabs(xr[r]) + fabs(xi[r]);
        if (v > best) { best = v; p = r; }
      }
    p = __builtin_amdgcn_readlane(p, j);
    if (lane == 0) piv[j] = p;
    if (p != j) {
      double tjr = 0, tji = 0, tpr = 0, tpi = 0;
#pragma unroll
      for (int r = 0; r < NMAX; ++r) {
        if (r == j) { tjr = xr[r]; tji = xi[r]; }
        if (r == p) { tpr = xr[r]; tpi = xi[r]; }
      }
#pragma unroll
      for (int r = 0; r < NMAX; ++r) {
        if (r == j) { xr[r] = tpr; xi[r] = tpi; }
        if (r == p) { xr[r] = tjr; xi[r] = tji; }
      }
    }
    double mjr = 0, mji = 0;  // my column's row-j entry
#pragma unroll
    for (int r = 0; r < NMAX; ++r)
      if (r == j) { mjr = xr[r]; mji = xi[r]; }
    const double dr = bcast(mjr, j), di = bcast(mji, j);
    if (lane == j) {
      ur = (p == j) ? dr : -dr;
      ui = (p == j) ? di : -di;
    }
    if (dr == 0.0 && di == 0.0) {
      singular = true;
      continue;
    }
    const double den = dr * dr + di * di;
    const double ir = dr / den, ii = -di / den;
    // scaled pivot-row entry of my column (1/d for column j itself)
    const double pjr = (lane == j) ? ir : mjr * ir - mji * ii;
    const double pji = (lane == j) ? ii : mjr * ii + mji * ir;
#pragma unroll
    for (int r = 0; r < NMAX; ++r) {
      if (r >= n) continue;
      const double fr = bcast(xr[r], j), fi = bcast(xi[r], j);
      if (r == j) {
        xr[r] = pjr;
        xi[r] = pji;
      } else if (lane == j) {
        xr[r] = -(fr * ir - fi * ii);
        xi[r] = -(fr * ii + fi * ir);
      } else {
        xr[r] -= fr * pjr - fi * pji;
        xi[r] -= fr * pji + fi * pjr;
      }
    }
  }
  if (ovlp != nullptr) {
    double ldr = 0.0, ldi = 0.0;
    if (lane < n) clog(ur + 2.2250738585072014e-308, ui, ldr, ldi);
    ldr = wave_sum(ldr);
    ldi = wave_sum(ldi);
    if (lane == 0) {
      z64 o = ovlp[b];
      double re = o.re + ldr, im = o.im + ldi;
      if (im > kPi) {
        while (im > kPi) im -= 2 * kPi;
      } else if (im < -kPi) {
        while (im < -kPi) im += 2 * kPi;
      }
      ovlp[b] = {ov_scale * re + ov_add_re, ov_scale * im + ov_add_im};
    }
  }
  if (!want_inverse) return;
  __syncthreads();
  if (!singular) {
    for (int j = n - 1; j >= 0; --j) {
      const int p = piv[j];
      if (p == j) continue;
      const int partner = (lane == j) ? p : ((lane == p) ? j : lane);
#pragma unroll
      for (int r = 0; r < NMAX; ++r) {
        xr[r] = __shfl(xr[r], partner, 64);
        xi[r] = __shfl(xi[r], partner, 64);
      }
    }
  }
  z64* Ib = Ainv + b * i_bs;
  if (lane < n) {
#pragma unroll
    for (int r = 0; r < NMAX; ++r)
      if (r < n) Ib[r * n + lane] = singular ? z64{0.0, 0.0} : z64{xr[r], xi[r]};
  }
}

// Register variant for 16 < n <= 64: one wavefront per matrix, lane r holds row r in registers.
// Rows are never moved: each lane tracks the position its row would occupy in getrf's swapped
// order, the pivot of step j is the unused row of largest |re|+|im| in column j with ties going
// to the lowest position (getrf's first maximum), and the row at position j trades positions
// with it. Gauss-Jordan then eliminates column j from every other row with the pivot row
// broadcast through LDS, so the only dynamically indexed register is column j. The log det
// takes log(+-U_jj + DBL_MIN) with the sign flipped exactly on getrf's swapped steps; the
// in-place result Y gives A^{-1}[k][c] = Y[p_k][s_c] (p_k = pivot row of step k, s_c = step at
// which row c pivoted), applied while storing.
__device__ inline double rdl(double v, int src) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_readlane(lo, src);
  hi = __builtin_amdgcn_readlane(hi, src);
  return __hiloint2double(hi, lo);
}

template <int NMAX>
__global__ void __launch_bounds__(64)
    k_lu_rows(const z64* __restrict__ A, long a_bs, int n, z64* __restrict__ Ainv, long i_bs,
              z64* __restrict__ ovlp, double ov_scale, double ov_add_re, double ov_add_im,
              int want_inverse, int no_pivot) {
  __shared__ int pivrow[64];
  __shared__ double2 prow[NMAX];
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
  bool used = lane >= n;
  int pos = lane;         // position of my row in getrf's order
  int mystep = n;         // step at which my row pivoted
  double ur = 1.0, ui = 0.0;  // my pivot value with getrf's sign (+-U_jj)
  bool singular = false;
  for (int j = 0; j < n; ++j) {
    double fr = 0.0, fi = 0.0;  // my row's entry in column j
#pragma unroll
    for (int c = 0; c < NMAX; ++c)
      if (c == j) { fr = xr[c]; fi = xi[c]; }
    // pivot: max |re|+|im| over unused rows, ties to the lowest position
    // (a NaN entry ranks below every number but above used rows, so the reduction stays
    // well defined and uniform)
    double mag = fabs(fr) + fabs(fi);
    double best = used ? -1.0 : (mag == mag ? mag : -0.5);
    int bpos = used ? 1 << 20 : pos;
    if (no_pivot) {
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
    // h = g / d with g = d - 1 on the pivot row (which becomes prow / d), f elsewhere
    const double gr = (lane == p) ? dr - 1.0 : fr, gi = (lane == p) ? di : fi;
    const double hr = gr * ir - gi * ii, hi = gr * ii + gi * ir;
    // column j of the result: 1/d on the pivot row, -f/d elsewhere
    const double njr = (lane == p) ? ir : -(fr * ir - fi * ii);
    const double nji = (lane == p) ? ii : -(fr * ii + fi * ir);
    // the pivot row goes through LDS (one writer lane, broadcast reads)
    if (lane == p) {
#pragma unroll
      for (int c = 0; c < NMAX; ++c) prow[c] = make_double2(xr[c], xi[c]);
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NMAX; ++c) {
      const double2 pv = prow[c];
      double nr = xr[c] - (hr * pv.x - hi * pv.y);
      double ni = xi[c] - (hr * pv.y + hi * pv.x);
      if (c == j) { nr = njr; ni = nji; }
      xr[c] = nr;
      xi[c] = ni;
    }
    __syncthreads();
  }
  if (ovlp != nullptr) {
    double lr = 0.0, li = 0.0;
    if (lane < n) clog(ur + 2.2250738585072014e-308, ui, lr, li);
    lr = wave_sum(lr);
    li = wave_sum(li);
    if (lane == 0) {
      z64 o = ovlp[b];
      double re = o.re + lr, im = o.im + li;
      if (im > kPi) {
        while (im > kPi) im -= 2 * kPi;
      } else if (im < -kPi) {
        while (im < -kPi) im += 2 * kPi;
      }
      ovlp[b] = {ov_scale * re + ov_add_re, ov_scale * im + ov_add_im};
    }
  }
  if (!want_inverse) return;
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

// Workgroup variant (256 threads): the matrix in LDS, every elimination step spread over all
// n^2 entries (pivot search by wavefront 0, row swap, pivot-row scaling, rank-1 update), one
// barrier between phases.  Same pivot rule and log det as k_lu_inverse_logdet.
constexpr int kLuThreads = 256;

__global__ void __launch_bounds__(kLuThreads)
    k_lu_blk(const z64* __restrict__ A, long a_bs, int n, z64* __restrict__ Ainv, long i_bs,
             z64* __restrict__ ovlp, double ov_scale, double ov_add_re, double ov_add_im,
             int want_inverse, int no_pivot) {
  extern __shared__ double2 sm2[];
  const int ld = n + 1;
  double2* a = sm2;               // [n][ld]
  double2* fcol = sm2 + n * ld;   // [n]
  __shared__ int piv_s[64];
  __shared__ int p_s;
  __shared__ double ld_s[2];
  __shared__ int sing_s;
  const int tid = threadIdx.x;
  const long b = blockIdx.x;
  const z64* Ab = A + b * a_bs;
  for (int e = tid; e < n * n; e += kLuThreads) {
    z64 v = Ab[e];
    a[(e / n) * ld + e % n] = make_double2(v.re, v.im);
  }
  if (tid == 0) {
    ld_s[0] = 0.0;
    ld_s[1] = 0.0;
    sing_s = 0;
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (tid < 64) {
      double best = -1.0;
      int p = n;
      if (no_pivot) {
        p = j;
      } else {
        for (int r = j + tid; r < n; r += 64) {
          double2 v = a[r * ld + j];
          double c = fabs(v.x) + fabs(v.y);
          if (c > best) { best = c; p = r; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          double ob = __shfl_xor(best, off, 64);
          int op = __shfl_xor(p, off, 64);
          if (ob > best || (ob == best && op < p)) { best = ob; p = op; }
        }
      }
      if (tid == 0) {
        p_s = p;
        piv_s[j] = p;
      }
    }
    __syncthreads();
    const int p = p_s;
    if (p != j)
      for (int c = tid; c < n; c += kLuThreads) {
        double2 x = a[j * ld + c];
        a[j * ld + c] = a[p * ld + c];
        a[p * ld + c] = x;
      }
    __syncthreads();
    const double2 d = a[j * ld + j];
    if (tid == 0) {
      double sr = (p == j) ? d.x : -d.x, si = (p == j) ? d.y : -d.y;
      double lr, li;
      clog(sr + 2.2250738585072014e-308, si, lr, li);
      ld_s[0] += lr;
      ld_s[1] += li;
    }
    if (d.x == 0.0 && d.y == 0.0) {
      if (tid == 0) sing_s = 1;
      continue;  // uniform: every thread read the same pivot
    }
    const double den = d.x * d.x + d.y * d.y;
    const double ir = d.x / den, ii = -d.y / den;
    for (int r = tid; r < n; r += kLuThreads)
      if (r != j) fcol[r] = a[r * ld + j];
    __syncthreads();
    for (int c = tid; c < n; c += kLuThreads) {
      if (c == j) {
        a[j * ld + c] = make_double2(ir, ii);
      } else {
        double2 x = a[j * ld + c];
        a[j * ld + c] = make_double2(x.x * ir - x.y * ii, x.x * ii + x.y * ir);
      }
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += kLuThreads) {
      const int r = e / n, c = e % n;
      if (r == j) continue;
      const double2 f = fcol[r];
      if (c == j) {
        a[r * ld + c] = make_double2(-(f.x * ir - f.y * ii), -(f.x * ii + f.y * ir));
      } else {
        const double2 pj = a[j * ld + c];
        double2 x = a[r * ld + c];
        x.x -= f.x * pj.x - f.y * pj.y;
        x.y -= f.x * pj.y + f.y * pj.x;
        a[r * ld + c] = x;
      }
    }
    __syncthreads();
  }
  if (tid == 0 && ovlp != nullptr) {
    z64 o = ovlp[b];
    double re = o.re + ld_s[0], im = o.im + ld_s[1];
    if (im > kPi) {
      while (im > kPi) im -= 2 * kPi;
    } else if (im < -kPi) {
      while (im < -kPi) im += 2 * kPi;
    }
    ovlp[b] = {ov_scale * re + ov_add_re, ov_scale * im + ov_add_im};
  }
  if (!want_inverse) return;
  __syncthreads();
  const bool singular = sing_s != 0;
  if (!singular) {
    for (int j = n - 1; j >= 0; --j) {
      const int p = piv_s[j];
      if (p != j)
        for (int r = tid; r < n; r += kLuThreads) {
          double2 x = a[r * ld + j];
          a[r * ld + j] = a[r * ld + p];
          a[r * ld + p] = x;
        }
      __syncthreads();
    }
  }
  z64* Ib = Ainv + b * i_bs;
  for (int e = tid; e < n * n; e += kLuThreads) {
    double2 v = a[(e / n) * ld + e % n];
    Ib[e] = singular ? z64{0.0, 0.0} : z64{v.x, v.y};
  }
}

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
                       long i_bs, z64* ovlp, double ov_scale, z64 ov_add, bool want_inverse) {
  AFQ_CHECK(n >= 1 && n <= 64, kErrShape, "n must be in [1, 64]");
  if (nb <= 0) return;
  const int wi = want_inverse ? 1 : 0;
#define AFQ_LU_REG(NM)                                                                        \
  hipLaunchKernelGGL(k_lu_reg<NM>, dim3(nb), dim3(64), 0, s, A, a_bs, n, Ainv, i_bs, ovlp,     \
                     ov_scale, ov_add.re, ov_add.im, wi)
#define AFQ_LU_ROWS(NM, NP)                                                                    \
  hipLaunchKernelGGL(k_lu_rows<NM>, dim3(nb), dim3(64), 0, s, A, a_bs, n, Ainv, i_bs, ovlp,    \
                     ov_scale, ov_add.re, ov_add.im, wi, NP)
  if (n <= 8)
    AFQ_LU_REG(8);
  else if (n <= 16)
    AFQ_LU_REG(16);
  else if (n <= 32)
    AFQ_LU_ROWS(32, 0);
  else if (n <= 48)
    AFQ_LU_ROWS(48, 0);
  else
    AFQ_LU_ROWS(64, 0);
#undef AFQ_LU_REG
  AFQ_LAUNCH_CHECK();
}

void chol_upper_inverse(hipStream_t s, const z64* S, int nb, int n, z64* Rinv, z64* ldet) {
  AFQ_CHECK(n >= 1 && n <= 64, kErrShape, "n must be in [1, 64]");
  if (nb <= 0) return;
  size_t lds = sizeof(double2) * ((size_t)n * (n + 1) + n);
  AFQ_CHECK(lds <= 65536, kErrShape, "matrix too large for one workgroup's LDS");
  // R is written into Rinv, then inverted in place (upper triangular: no pivoting needed);
  // the inversion's log det is sum_i log R_ii (R_ii real and positive).
  hipLaunchKernelGGL(k_chol_blk, dim3(nb), dim3(kLuThreads), lds, s, S, n, Rinv);
  AFQ_LAUNCH_CHECK();
  const long nn = (long)n * n;
  if (n <= 32)
    hipLaunchKernelGGL(k_lu_rows<32>, dim3(nb), dim3(64), 0, s, Rinv, nn, n, Rinv, nn, ldet, 1.0,
                       0.0, 0.0, 1, 1);
  else if (n <= 48)
    hipLaunchKernelGGL(k_lu_rows<48>, dim3(nb), dim3(64), 0, s, Rinv, nn, n, Rinv, nn, ldet, 1.0,
                       0.0, 0.0, 1, 1);
  else
    hipLaunchKernelGGL(k_lu_rows<64>, dim3(nb), dim3(64), 0, s, Rinv, nn, n, Rinv, nn, ldet, 1.0,
                       0.0, 0.0, 1, 1);
  AFQ_LAUNCH_CHECK();
}

}  // namespace afq
