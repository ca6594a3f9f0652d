// Elementwise and reduction helpers (casts, fills, hermitization, the energy reductions).
// The reference's equivalents are the copy/accumulate/add_scalar/zero_imag device kernels
// (numerics/device_kernels/cuda/*.cu) and the host loops of energy_impl
// (Real3IndexFactorization.hpp:529-587).
#include <algorithm>

#include "afq_internal.hpp"

namespace afq {

namespace {

constexpr int kT = 256;

__global__ void k_scale(long n, const z64* x, double ar, double ai, z64* y) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) {
    z64 v = x[i];
    y[i] = {ar * v.re - ai * v.im, ar * v.im + ai * v.re};
  }
}
__global__ void k_axpy(long n, double ar, double ai, const z64* x, z64* y) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) {
    z64 v = x[i], o = y[i];
    y[i] = {o.re + ar * v.re - ai * v.im, o.im + ar * v.im + ai * v.re};
  }
}
__global__ void k_z2c(long n, const z64* x, c32* y) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) y[i] = {(float)x[i].re, (float)x[i].im};
}
__global__ void k_c2z(long n, const c32* x, z64* y) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) y[i] = {(double)x[i].re, (double)x[i].im};
}
__global__ void k_d2f(long n, const double* x, float* y) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}
__global__ void k_fill(long n, z64 v, z64* y) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) y[i] = v;
}
// y[b * stride + i] = v for i < n (one block of n elements in every stride)
__global__ void k_fill_blocks(long B, long n, long stride, z64 v, z64* y) {
  long t = (long)blockIdx.x * kT + threadIdx.x;
  if (t < B * n) y[(t / n) * stride + t % n] = v;
}
// out[b][i][j] = in[b][j][i] for B square N x N blocks
__global__ void k_transpose_blocks(long B, int N, const z64* __restrict__ in, z64* __restrict__ out) {
  long t = (long)blockIdx.x * kT + threadIdx.x;
  const long NN = (long)N * N;
  if (t >= B * NN) return;
  const long b = t / NN, e = t % NN;
  out[t] = in[b * NN + (e % N) * N + e / N];
}
__global__ void k_zero_imag(long n, z64* y) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) y[i].im = 0.0;
}
__global__ void k_truncate(long n, double cut, z64* y) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) {
    z64 v = y[i];
    if (hypot(v.re, v.im) <= cut) y[i] = {0.0, 0.0};
  }
}
// hermitize (tensor.hpp:75-94) then scale: one thread per (i, j >= i)
__global__ void k_hermitize_scale(int N, double scale, z64* A) {
  long t = (long)blockIdx.x * kT + threadIdx.x;
  if (t >= (long)N * N) return;
  int i = (int)(t / N), j = (int)(t % N);
  if (j < i) return;
  if (i == j) {
    A[(long)i * N + i] = {scale * A[(long)i * N + i].re, 0.0};
  } else {
    z64 a = A[(long)i * N + j], b = A[(long)j * N + i];
    z64 h = {0.5 * a.re + 0.5 * b.re, 0.5 * a.im - 0.5 * b.im};
    A[(long)i * N + j] = {scale * h.re, scale * h.im};
    A[(long)j * N + i] = {scale * h.re, -scale * h.im};
  }
}
// H1 = dt*h + dt*vexx + shift (add_one_body_shifts, one_body.hpp)
__global__ void k_add_one_body(long n, double dt, const double* h, const double* vexx,
                               const z64* shift, z64* H1) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) H1[i] = {dt * h[i] + dt * vexx[i] + shift[i].re, shift[i].im};
}
__global__ void k_identity(int N, z64* A) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < (long)N * N) A[i] = {(i / N == i % N) ? 1.0 : 0.0, 0.0};
}
// out[m][a][k] = in[a][k][m]   (half-rotated Cholesky layout change)
__global__ void k_permute_akm_mak(int A, int K, int M, const z64* in, z64* out) {
  long t = (long)blockIdx.x * kT + threadIdx.x;
  if (t >= (long)A * K * M) return;
  int m = (int)(t % M);
  long r = t / M;
  int k = (int)(r % K), a = (int)(r / K);
  out[((long)m * A + a) * K + k] = in[t];
}
// flag = 1 when L[i][k][m] != L[k][i][m] for some i < k, m
__global__ void k_check_sym(int N, int M, const double* L, int* flag) {
  long t = (long)blockIdx.x * kT + threadIdx.x;
  if (t >= (long)N * N * M) return;
  int m = (int)(t % M);
  long ik = t / M;
  int i = (int)(ik / N), k = (int)(ik % N);
  if (i < k && L[((long)i * N + k) * M + m] != L[((long)k * N + i) * M + m]) *flag = 1;
}
// Psi[i][a] = conj(PsiT[a][i])
__global__ void k_conj_transpose(int R, int C, const z64* in, z64* out) {
  long t = (long)blockIdx.x * kT + threadIdx.x;
  if (t >= (long)R * C) return;
  int r = (int)(t / C), c = (int)(t % C);
  z64 v = in[t];
  out[(long)c * R + r] = {v.re, -v.im};
}
__global__ void k_broadcast(long W, long n, const z64* src, z64* dst) {
  long t = (long)blockIdx.x * kT + threadIdx.x;
  if (t >= W * n) return;
  dst[t] = src[t % n];
}
// dst[r][:] = src[idx[r]][:] (rows of length C)
__global__ void k_gather_rows(long R, int C, const int* idx, const z64* src, z64* dst) {
  long t = (long)blockIdx.x * kT + threadIdx.x;
  if (t >= R * C) return;
  const long r = t / C;
  dst[t] = src[(long)idx[r] * C + t % C];
}
// columns cols.p[k][w * stride] <-> packed[k][w] for the K property columns of W walkers
__global__ void k_columns(int K, int W, long stride, ColumnSet cols, z64* packed, int to_packed) {
  long t = (long)blockIdx.x * kT + threadIdx.x;
  if (t >= (long)K * W) return;
  const int k = (int)(t / W), w = (int)(t % W);
  z64* c = cols.p[k] + (long)w * stride;
  if (to_packed) packed[t] = *c;
  else *c = packed[t];
}
// out[b][j][i][:] = in[b][i][j][:]   (in [B][D0][D1][K], out [B][D1][D0][K])
__global__ void k_swap_mid(long B, int D0, int D1, int K, const z64* in, z64* out) {
  long t = (long)blockIdx.x * kT + threadIdx.x;
  if (t >= B * D0 * D1 * K) return;
  const int k = (int)(t % K);
  long r = t / K;
  const int j = (int)(r % D1);
  r /= D1;
  const int i = (int)(r % D0);
  const long b = r / D0;
  out[((b * D1 + j) * D0 + i) * K + k] = in[t];
}
__global__ void k_real_part(long n, const z64* x, double* y) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) y[i] = x[i].re;
}
__global__ void k_real_part_f(long n, const z64* x, float* y) {
  long i = (long)blockIdx.x * kT + threadIdx.x;
  if (i < n) y[i] = (float)x[i].re;
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
  return t;
}

// out[w] += c * sum_k G[w][k] H[k] + add   (the E1 contraction, energy_impl :504-508)
__global__ void __launch_bounds__(kT)
    k_dot_rows(long K, const z64* G, const z64* H, double cr, double ci, double add, z64* out,
               long os) {
  __shared__ double red[kT / 64];
  const int w = blockIdx.x;
  double sr = 0, si = 0;
  for (long k = threadIdx.x; k < K; k += kT) {
    z64 g = G[(long)w * K + k], h = H[k];
    sr += g.re * h.re - g.im * h.im;
    si += g.re * h.im + g.im * h.re;
  }
  sr = block_sum<kT>(sr, red);
  si = block_sum<kT>(si, red);
  if (threadIdx.x == 0) {
    z64 o = out[(long)w * os];
    out[(long)w * os] = {o.re + cr * sr - ci * si + add, o.im + cr * si + ci * sr};
  }
}

// Exchange/Coulomb reduction of one Cholesky chunk.  T holds, for every walker w and vector
// m in [m0, m0 + MC), T[(w, b)][(m, a)] = sum_k G[w][b][k] Lnak[m][a][k]  ([W*n][MC*n]).
//   exx_part[w][chunk] = sum_{m,a,b} T[w,a,m,b] T[w,b,m,a];   Kl[w][m] += sum_a T[w,a,m,a]
// One workgroup per walker; fixed summation order.
__global__ void __launch_bounds__(kT)
    k_exchange_reduce(int n, int MC, int mcount, int m0, int M, const z64* T, z64* exx_part,
                      int nchunk, int chunk_id, z64* Kl) {
  __shared__ double red[kT / 64];
  const int w = blockIdx.x;
  const long ldt = (long)MC * n;
  const z64* Tw = T + (long)w * n * ldt;
  double er = 0, ei = 0;
  const long tot = (long)mcount * n * n;
  for (long t = threadIdx.x; t < tot; t += kT) {
    int m = (int)(t / ((long)n * n));
    int rem = (int)(t % ((long)n * n));
    int a = rem / n, b = rem % n;
    z64 x = Tw[(long)a * ldt + (long)m * n + b];   // T[w,a,m,b]
    z64 y = Tw[(long)b * ldt + (long)m * n + a];   // T[w,b,m,a]
    er += x.re * y.re - x.im * y.im;
    ei += x.re * y.im + x.im * y.re;
  }
  er = block_sum<kT>(er, red);
  ei = block_sum<kT>(ei, red);
  if (threadIdx.x == 0) exx_part[(long)w * nchunk + chunk_id] = {er, ei};
  if (Kl) {
    for (int m = threadIdx.x; m < mcount; m += kT) {
      double kr = 0, ki = 0;
      for (int a = 0; a < n; ++a) {
        z64 x = Tw[(long)a * ldt + (long)m * n + a];
        kr += x.re;
        ki += x.im;
      }
      z64 o = Kl[(long)w * M + m0 + m];
      Kl[(long)w * M + m0 + m] = {o.re + kr, o.im + ki};
    }
  }
}

// Exchange and Coulomb vectors of a group of mg Cholesky vectors per workgroup.  T holds, per
// walker, rows b (electrons) of ldt = MC n elements: T[w][b][m][a] = sum_k G[w][b][k] L[m][a][k]
// for the chunk's m (local index).  Workgroup (w, g) copies rows b of its mg blocks into LDS
// (mg n contiguous elements per row, one coalesced segment each), then
//   exx_part[w][g0 + g] = sum_{m, a, b} T^m[b][a] T^m[a][b]   (fixed order, FP64)
//   Kl[w][m0 + m] += sum_a T^m[a][a]
template <typename TT>
__global__ void __launch_bounds__(kT)
    k_exchange_blocks(int n, int MC, int mcount, int mg, int m0, int M, const cplx<TT>* T,
                      z64* exx_part, int ng_total, int g0, z64* Kl) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TT* tr = reinterpret_cast<TT*>(smem_raw);
  const int w = blockIdx.x, g = blockIdx.y;
  const int mb = g * mg, me = min(mcount, mb + mg), cnt = me - mb;
  const int rl = cnt * n, ld = rl + 1;  // row length in LDS (+1 against bank conflicts)
  TT* ti = tr + (long)n * ld;
  const long ldt = (long)MC * n;
  const cplx<TT>* Tw = T + (long)w * n * ldt + (long)mb * n;
  for (int t = threadIdx.x; t < n * rl; t += kT) {
    const int b = t / rl, c = t % rl;
    const cplx<TT> v = Tw[(long)b * ldt + c];
    tr[b * ld + c] = v.re;
    ti[b * ld + c] = v.im;
  }
  __syncthreads();
  __shared__ double red[kT / 64];
  double er = 0, ei = 0;
  const int nn = n * n;
  for (int t = threadIdx.x; t < cnt * nn; t += kT) {
    const int m = t / nn, r = t % nn, b = r / n, a = r % n;
    const int i1 = b * ld + m * n + a, i2 = a * ld + m * n + b;  // T^m[b][a], T^m[a][b]
    const double xr = tr[i1], xi = ti[i1], yr = tr[i2], yi = ti[i2];
    er += xr * yr - xi * yi;
    ei += xr * yi + xi * yr;
  }
  er = block_sum<kT>(er, red);
  ei = block_sum<kT>(ei, red);
  if (threadIdx.x == 0) exx_part[(long)w * ng_total + g0 + g] = {er, ei};
  if (Kl) {
    for (int m = threadIdx.x; m < cnt; m += kT) {
      double kr = 0, ki = 0;
      for (int a = 0; a < n; ++a) {
        kr += tr[a * ld + m * n + a];
        ki += ti[a * ld + m * n + a];
      }
      z64 o = Kl[(long)w * M + m0 + mb + m];
      Kl[(long)w * M + m0 + mb + m] = {o.re + kr, o.im + ki};
    }
  }
}

// E[w][1] += cx * sum_chunks exx;  E[w][2] += cj * sum_m Kl^2
__global__ void __launch_bounds__(kT)
    k_energy_finish(int M, int nchunk, const z64* exx_part, const z64* Kl, double cx, double cj,
                    int addEXX, int addEJ, z64* E) {
  __shared__ double red[kT / 64];
  const int w = blockIdx.x;
  double jr = 0, ji = 0;
  if (addEJ)
    for (int m = threadIdx.x; m < M; m += kT) {
      z64 k = Kl[(long)w * M + m];
      jr += k.re * k.re - k.im * k.im;
      ji += 2 * k.re * k.im;
    }
  jr = block_sum<kT>(jr, red);
  ji = block_sum<kT>(ji, red);
  if (threadIdx.x == 0) {
    if (addEXX) {
      double xr = 0, xi = 0;
      for (int c = 0; c < nchunk; ++c) {
        xr += exx_part[(long)w * nchunk + c].re;
        xi += exx_part[(long)w * nchunk + c].im;
      }
      E[(long)w * 3 + 1] = {E[(long)w * 3 + 1].re + cx * xr, E[(long)w * 3 + 1].im + cx * xi};
    }
    if (addEJ) E[(long)w * 3 + 2] = {E[(long)w * 3 + 2].re + cj * jr, E[(long)w * 3 + 2].im + cj * ji};
  }
}

inline dim3 grid1(long n) { return dim3((unsigned)ceil_div(n, kT)); }

}  // namespace

void scale_add_cplx(hipStream_t s, long n, const z64* x, double ar, double ai, z64* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale, grid1(n), dim3(kT), 0, s, n, x, ar, ai, y);
  AFQ_LAUNCH_CHECK();
}
void axpy_z64(hipStream_t s, long n, double ar, double ai, const z64* x, z64* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy, grid1(n), dim3(kT), 0, s, n, ar, ai, x, y);
  AFQ_LAUNCH_CHECK();
}
void cast_z64_c32(hipStream_t s, long n, const z64* x, c32* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_z2c, grid1(n), dim3(kT), 0, s, n, x, y);
  AFQ_LAUNCH_CHECK();
}
void cast_c32_z64(hipStream_t s, long n, const c32* x, z64* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_c2z, grid1(n), dim3(kT), 0, s, n, x, y);
  AFQ_LAUNCH_CHECK();
}
void cast_f64_f32(hipStream_t s, long n, const double* x, float* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_d2f, grid1(n), dim3(kT), 0, s, n, x, y);
  AFQ_LAUNCH_CHECK();
}
void fill_z64(hipStream_t s, long n, z64 v, z64* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, grid1(n), dim3(kT), 0, s, n, v, y);
  AFQ_LAUNCH_CHECK();
}
void fill_blocks_z64(hipStream_t s, long B, long n, long stride, z64 v, z64* y) {
  if (B * n <= 0) return;
  hipLaunchKernelGGL(k_fill_blocks, grid1(B * n), dim3(kT), 0, s, B, n, stride, v, y);
  AFQ_LAUNCH_CHECK();
}
void transpose_blocks(hipStream_t s, long B, int N, const z64* in, z64* out) {
  if (B * N <= 0) return;
  hipLaunchKernelGGL(k_transpose_blocks, grid1(B * N * N), dim3(kT), 0, s, B, N, in, out);
  AFQ_LAUNCH_CHECK();
}
void zero_imag(hipStream_t s, long n, z64* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_imag, grid1(n), dim3(kT), 0, s, n, y);
  AFQ_LAUNCH_CHECK();
}
void truncate_small(hipStream_t s, long n, double cut, z64* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_truncate, grid1(n), dim3(kT), 0, s, n, cut, y);
  AFQ_LAUNCH_CHECK();
}
void hermitize_scale(hipStream_t s, int N, double scale, z64* A) {
  hipLaunchKernelGGL(k_hermitize_scale, grid1((long)N * N), dim3(kT), 0, s, N, scale, A);
  AFQ_LAUNCH_CHECK();
}
void add_one_body(hipStream_t s, int N, double dt, const double* h, const double* vexx,
                  const z64* shift, z64* H1) {
  long n = (long)N * N;
  hipLaunchKernelGGL(k_add_one_body, grid1(n), dim3(kT), 0, s, n, dt, h, vexx, shift, H1);
  AFQ_LAUNCH_CHECK();
}
void identity(hipStream_t s, int N, z64* A) {
  hipLaunchKernelGGL(k_identity, grid1((long)N * N), dim3(kT), 0, s, N, A);
  AFQ_LAUNCH_CHECK();
}
void permute_akm_mak(hipStream_t s, int A, int K, int M, const z64* in, z64* out) {
  long n = (long)A * K * M;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_permute_akm_mak, grid1(n), dim3(kT), 0, s, A, K, M, in, out);
  AFQ_LAUNCH_CHECK();
}
bool symmetric_cholesky(hipStream_t s, int N, int M, const double* L) {
  int* f = nullptr;
  AFQ_HIP(hipMalloc(&f, sizeof(int)));
  int h = 0;
  AFQ_HIP(hipMemcpyAsync(f, &h, sizeof(int), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_check_sym, grid1((long)N * N * M), dim3(kT), 0, s, N, M, L, f);
  AFQ_LAUNCH_CHECK();
  AFQ_HIP(hipMemcpyAsync(&h, f, sizeof(int), hipMemcpyDeviceToHost, s));
  AFQ_HIP(hipStreamSynchronize(s));
  AFQ_HIP(hipFree(f));
  return h == 0;
}
void conj_transpose(hipStream_t s, int R, int C, const z64* in, z64* out) {
  long n = (long)R * C;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_conj_transpose, grid1(n), dim3(kT), 0, s, R, C, in, out);
  AFQ_LAUNCH_CHECK();
}
void broadcast_z64(hipStream_t s, long W, long n, const z64* src, z64* dst) {
  if (W * n <= 0) return;
  hipLaunchKernelGGL(k_broadcast, grid1(W * n), dim3(kT), 0, s, W, n, src, dst);
  AFQ_LAUNCH_CHECK();
}
void gather_rows(hipStream_t s, long R, int C, const int* idx, const z64* src, z64* dst) {
  if (R * C <= 0) return;
  hipLaunchKernelGGL(k_gather_rows, grid1(R * C), dim3(kT), 0, s, R, C, idx, src, dst);
  AFQ_LAUNCH_CHECK();
}
void copy_columns(hipStream_t s, int K, int W, long stride, const ColumnSet& cols, z64* packed,
                  bool to_packed) {
  AFQ_CHECK(K >= 0 && K <= ColumnSet::kMax, kErrArg, "too many columns");
  if ((long)K * W <= 0) return;
  hipLaunchKernelGGL(k_columns, grid1((long)K * W), dim3(kT), 0, s, K, W, stride, cols, packed,
                     to_packed ? 1 : 0);
  AFQ_LAUNCH_CHECK();
}
void swap_mid(hipStream_t s, long B, int D0, int D1, int K, const z64* in, z64* out) {
  const long n = B * D0 * D1 * K;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_swap_mid, grid1(n), dim3(kT), 0, s, B, D0, D1, K, in, out);
  AFQ_LAUNCH_CHECK();
}
void real_part(hipStream_t s, long n, const z64* x, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_real_part, grid1(n), dim3(kT), 0, s, n, x, y);
  AFQ_LAUNCH_CHECK();
}
void real_part_f32(hipStream_t s, long n, const z64* x, float* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_real_part_f, grid1(n), dim3(kT), 0, s, n, x, y);
  AFQ_LAUNCH_CHECK();
}
void dot_rows(hipStream_t s, int W, long K, const z64* G, const z64* H, double cr, double ci,
              double add, z64* out, long out_stride) {
  if (W <= 0) return;
  hipLaunchKernelGGL(k_dot_rows, dim3(W), dim3(kT), 0, s, K, G, H, cr, ci, add, out, out_stride);
  AFQ_LAUNCH_CHECK();
}
void exchange_reduce(hipStream_t s, int W, int n, int MC, int m0, int M, const z64* T,
                     z64* exx_part, int nchunk, int chunk_id, z64* Kl) {
  if (W <= 0) return;
  int mcount = (m0 + MC <= M) ? MC : M - m0;
  hipLaunchKernelGGL(k_exchange_reduce, dim3(W), dim3(kT), 0, s, n, MC, mcount, m0, M, T, exx_part,
                     nchunk, chunk_id, Kl);
  AFQ_LAUNCH_CHECK();
}
int exchange_group(int n) {
  // Cholesky vectors per workgroup: the blocks' FP64 re/im planes within 64 KB of LDS
  const long per = 2L * sizeof(double) * n * (n + 1);
  return (int)std::max<long>(1, std::min<long>(16, 65536 / per));
}

template <typename TT>
void exchange_blocks(hipStream_t s, int W, int n, int mg, int MC, int m0, int mcount, int M,
                     const cplx<TT>* T, z64* exx_part, int ng_total, int g0, z64* Kl) {
  if (W <= 0 || mcount <= 0) return;
  AFQ_CHECK(mg >= 1 && m0 % mg == 0, kErrArg, "exchange group size");
  const int ng = ceil_div(mcount, mg);
  const size_t lds = 2 * sizeof(TT) * (size_t)n * (mg * n + 1);
  AFQ_CHECK(lds <= 160 * 1024 && g0 + ng <= ng_total, kErrShape, "exchange block layout");
  if (lds > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_exchange_blocks<TT>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k_exchange_blocks<TT>, dim3(W, ng), dim3(kT), lds, s, n, MC, mcount, mg, m0,
                     M, T, exx_part, ng_total, g0, Kl);
  AFQ_LAUNCH_CHECK();
}
template void exchange_blocks<double>(hipStream_t, int, int, int, int, int, int, int, const z64*,
                                      z64*, int, int, z64*);
template void exchange_blocks<float>(hipStream_t, int, int, int, int, int, int, int, const c32*,
                                     z64*, int, int, z64*);

void energy_finish(hipStream_t s, int W, int M, int nchunk, const z64* exx_part, const z64* Kl,
                   double cx, double cj, int addEXX, int addEJ, z64* E) {
  if (W <= 0) return;
  hipLaunchKernelGGL(k_energy_finish, dim3(W), dim3(kT), 0, s, M, nchunk, exx_part, Kl, cx, cj,
                     addEXX, addEJ, E);
  AFQ_LAUNCH_CHECK();
}

}  // namespace afq
