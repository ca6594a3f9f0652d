// Particle-hole multi-determinant trial (PHMSD) for COLLINEAR walkers on the device.
//
//   woodbury         det_ops::MixedDensityMatrixForWoodbury (density_matrix.hpp:651-717):
//                    TAB = orbs Phi, LU/log det/inverse of its reference rows, QQ0 = TAB TNN^{-1}
//                    (only the particle rows are kept), GA = TNN^{-T} Phi^T
//   ratios           calculate_overlaps (phmsd_helpers.hpp:112-198) and the block inverses of
//                    get_compact_ph_R_matrices (:312-390): one thread per (walker, excitation)
//   weights          the spin-coupling products of PHMSD::MixedDensityMatrix (PHMSD.cpp:712-760)
//   R, G             calculate_R (phmsd_helpers.hpp:201-290) and G = R^T GA / Ov
//   energy           PHMSD::energy_alg1 (PHMSD.cpp:376-544) with ph_reference_energy and
//                    ph_excited_energy (Real3IndexFactorization.hpp:212-273,
//                    detail/phmsd_impl.hpp:33-190)
//   vMF              PHMSD::vMF (PHMSD.cpp:1329-1573)
//
// Energy formulation. For an excitation u of rank r (holes h_p, particles x_p) the reference
// evaluates, per Cholesky slice n, the compact R_u (r x nact) contracted with
// T[a][l] = sum_k L[n][a][k] GA[l][k]: delta_p[l] = sum_a R_u[p][a] T[a][l]. With
// R_u = Minv (E - Q P_u) that is delta = Minv Vz,
//   Vz[q][l] = T[x_q][l] - B[x_q][l] - sum_p' M[q][p'] Delta_p'[l],
//   B[x][l] = sum_i QQ0[x][i] T[i][l]  (shared by every excitation of the walker and slice),
//   Delta_p[l] = T[x_p][l] - T[h_p][l],   M[q][p] = QQ0[x_q][h_p],
// so the per-excitation work is O(r^2 nel) instead of O(r nact nel). The reference's
// exchange and Coulomb differences then read
//   Kn = Jref + sum_p (delta_p[h_p] + Delta_p[h_p])
//   eX = sum_pq delta_p[h_q] delta_q[h_p] + 2 sum_pl delta_p[l] T[O_l][h_p]
//        + 2 sum_pl Delta_p[l] T[l][h_p] + sum_pq Delta_p[h_q] Delta_q[h_p]
// (O_l = orbital of electron l in u), which is exactly tr(Y^2) - tr(Zref^2) of its
// ph_excited_2body_energy_dense_cholesky; the one-body slice (haj appended to L) gives the
// ph_excited_1body_energy term Kn - Jref.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gemm.hpp"
#include "ops.hpp"
#include "tuner.hpp"

namespace afq {

namespace {

constexpr int kXS = 2 * PhmsdTrial::kMaxRank;

template <typename T>
T* dalloc(size_t count) {
  void* p = nullptr;
  if (hipMalloc(&p, (count ? count : 1) * sizeof(T)) != hipSuccess)
    throw Error(kErrAlloc, "device allocation failed");
  return reinterpret_cast<T*>(p);
}
template <typename T>
T* dcopy(hipStream_t s, const std::vector<T>& h) {
  T* d = dalloc<T>(h.size());
  if (!h.empty()) AFQ_HIP(hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
  return d;
}

__device__ inline z64 zadd(z64 a, z64 b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline z64 zsub(z64 a, z64 b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline z64 zmul(z64 a, z64 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline z64 zfma(z64 a, z64 b, z64 c) {  // c + a b
  return {fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im))};
}
__device__ inline z64 zconj(z64 a) { return {a.re, -a.im}; }
__device__ inline double zabs1(z64 a) { return fabs(a.re) + fabs(a.im); }
__device__ inline bool zzero(z64 a) { return a.re == 0.0 && a.im == 0.0; }
__device__ inline z64 zrecip(z64 b) {  // 1 / b, scaled (Smith)
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return {1.0 / d, -r / d};
  }
  double r = b.re / b.im, d = b.re * r + b.im;
  return {r / d, -1.0 / d};
}
__device__ inline z64 zlog(z64 a) { return {log(hypot(a.re, a.im)), atan2(a.im, a.re)}; }

// Block-wide sum of complex values (deterministic tree over a fixed thread order).
template <int NT>
__device__ inline z64 block_sum(z64 v, z64* red) {
  red[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int off = NT / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] = zadd(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  z64 r = red[0];
  __syncthreads();
  return r;
}

// ---- excitation determinants and inverses -----------------------------------------------------

// LU with partial pivoting (first maximum of |re|+|im|, zgetf2 semantics) of an R x R matrix in
// registers; every index is a compile-time constant so nothing spills to scratch. det receives
// the determinant; inv (when requested) the inverse, zero when a pivot vanished.
template <int R>
__device__ inline void small_lu(z64 (&a)[R][R], z64& det, bool want_inv, z64 (&inv)[R][R]) {
  int perm[R];
#pragma unroll
  for (int i = 0; i < R; ++i) perm[i] = i;
  bool sing = false;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    int p = j;
    double best = zabs1(a[j][j]);
#pragma unroll
    for (int i = j + 1; i < R; ++i) {
      double v = zabs1(a[i][j]);
      if (v > best) { best = v; p = i; }
    }
#pragma unroll
    for (int i = j + 1; i < R; ++i)
      if (i == p) {
#pragma unroll
        for (int c = 0; c < R; ++c) {
          z64 t = a[j][c];
          a[j][c] = a[i][c];
          a[i][c] = t;
        }
        int t = perm[j];
        perm[j] = perm[i];
        perm[i] = t;
      }
    bool swapped = p != j;
    if (!zzero(a[j][j])) {
      z64 r = zrecip(a[j][j]);
#pragma unroll
      for (int i = j + 1; i < R; ++i) a[i][j] = zmul(a[i][j], r);
    } else {
      sing = true;
    }
#pragma unroll
    for (int i = j + 1; i < R; ++i)
#pragma unroll
      for (int c = j + 1; c < R; ++c) a[i][c] = zsub(a[i][c], zmul(a[i][j], a[j][c]));
    z64 d = swapped ? z64{-a[j][j].re, -a[j][j].im} : a[j][j];
    det = j == 0 ? d : zmul(det, d);
  }
  if (!want_inv) return;
  // solve L U X = P: column c of P has its one in row k where perm[k] == c
#pragma unroll
  for (int c = 0; c < R; ++c) {
    z64 x[R];
#pragma unroll
    for (int k = 0; k < R; ++k) x[k] = perm[k] == c ? z64{1.0, 0.0} : z64{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int k = 0; k < i; ++k) x[i] = zsub(x[i], zmul(a[i][k], x[k]));
#pragma unroll
    for (int i = R - 1; i >= 0; --i) {
#pragma unroll
      for (int k = i + 1; k < R; ++k) x[i] = zsub(x[i], zmul(a[i][k], x[k]));
      x[i] = zmul(x[i], zrecip(a[i][i]));
    }
#pragma unroll
    for (int i = 0; i < R; ++i) inv[i][c] = sing ? z64{0.0, 0.0} : x[i];
  }
}

// ratio[w][u] = det M_u, M_u[p][q] = QQ0[x_p][h_q]; Minv of every excitation (woodbury R).
// With want_inv also the compact R rows of get_compact_ph_R_matrices (phmsd_helpers.hpp:312-390),
// Rc[p][a] = sum_q Minv[p][q] ([a == x_q] - QQ0[x_q][pos(a)]) with pos(a) the electron holding
// active orbital a in the excitation (none for an emptied hole or an unoccupied orbital),
// written to Rc[w][rstart[u] + p][0..nact).
template <int R>
__global__ void __launch_bounds__(256)
    k_ph_ratios(int W, int nel, int nvir, int u0, int cnt, int nconf, const int* __restrict__ exc,
                const int* __restrict__ mstart, int P2, const z64* __restrict__ Qv,
                z64* __restrict__ ovr, z64* __restrict__ Minv, int want_inv,
                const int* __restrict__ rstart, int P, z64* __restrict__ Rc) {
  long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)W * cnt) return;
  const int w = (int)(idx / cnt), u = u0 + (int)(idx % cnt);
  const int* e = exc + (long)u * kXS;
  const z64* Q = Qv + (long)w * nvir * nel;
  z64 a[R][R], inv[R][R];
#pragma unroll
  for (int p = 0; p < R; ++p)
#pragma unroll
    for (int q = 0; q < R; ++q) a[p][q] = Q[(e[R + p] - nel) * nel + e[q]];
  z64 det;
  small_lu<R>(a, det, want_inv != 0, inv);
  ovr[(long)w * nconf + u] = det;
  if (want_inv) {
    z64* out = Minv + (long)w * P2 + mstart[u];
#pragma unroll
    for (int p = 0; p < R; ++p)
#pragma unroll
      for (int q = 0; q < R; ++q) out[p * R + q] = inv[p][q];
    const int nact = nel + nvir;
    z64* rc = Rc + ((long)w * P + rstart[u]) * nact;
    for (int a = 0; a < nact; ++a) {
      int pos = a < nel ? a : -1;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (e[q] == a) pos = -1;
        if (e[R + q] == a) pos = e[q];
      }
      z64 coef[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        coef[q] = {e[R + q] == a ? 1.0 : 0.0, 0.0};
        if (pos >= 0) coef[q] = zsub(coef[q], Q[(e[R + q] - nel) * nel + pos]);
      }
#pragma unroll
      for (int p = 0; p < R; ++p) {
        z64 v = {0.0, 0.0};
#pragma unroll
        for (int q = 0; q < R; ++q) v = zfma(inv[p][q], coef[q], v);
        rc[(long)p * nact + a] = v;
      }
    }
  }
}

// wa[u] = ova[u] sum_v C0[u][v] ovb[v], Ov = sum_u wa[u]; wb[v] = ovb[v] sum_u C1[v][u] ova[u].
// One wavefront per coupling row (rows differ widely in length: the reference configuration
// couples to every determinant that leaves its spin unexcited); fixed summation order.
__global__ void __launch_bounds__(256)
    k_ph_weights(int na, int nb, const z64* __restrict__ ova, const z64* __restrict__ ovb,
                 const int* __restrict__ cpa, const int* __restrict__ cca, const z64* __restrict__ cva,
                 const int* __restrict__ cpb, const int* __restrict__ ccb, const z64* __restrict__ cvb,
                 z64* __restrict__ wa, z64* __restrict__ wb, z64* __restrict__ Ov) {
  __shared__ z64 part[4];
  const int w = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const z64* oa = ova + (long)w * na;
  const z64* ob = ovb + (long)w * nb;
  z64 acc = {0.0, 0.0};
  for (int u = wave; u < na; u += 4) {
    z64 s = {0.0, 0.0};
    for (int k = cpa[u] + lane; k < cpa[u + 1]; k += 64) s = zfma(cva[k], ob[cca[k]], s);
    s.re = wave_sum(s.re);
    s.im = wave_sum(s.im);
    const z64 x = zmul(oa[u], s);
    acc = zadd(acc, x);
    if (lane == 0) wa[(long)w * na + u] = x;
  }
  for (int v = wave; v < nb; v += 4) {
    z64 s = {0.0, 0.0};
    for (int k = cpb[v] + lane; k < cpb[v + 1]; k += 64) s = zfma(cvb[k], oa[ccb[k]], s);
    s.re = wave_sum(s.re);
    s.im = wave_sum(s.im);
    if (lane == 0) wb[(long)w * nb + v] = zmul(ob[v], s);
  }
  if (lane == 0) part[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) Ov[w] = zadd(zadd(part[0], part[1]), zadd(part[2], part[3]));
}

// The same weights from coupling chunks (at most kChunk entries of one row each): one thread
// per chunk with both spins' overlaps in LDS, then one thread per row adds its chunks in order,
// and Ov adds the alpha weights in a fixed tree.  chunks [nka + nkb] = (row, begin, end); chunk
// ranges per row in kpa / kpb.
constexpr int kChunk = 16;
__global__ void __launch_bounds__(256)
    k_ph_weights_chunked(int na, int nb, const z64* __restrict__ ova, const z64* __restrict__ ovb,
                         const int* __restrict__ cca, const z64* __restrict__ cva,
                         const int* __restrict__ ccb, const z64* __restrict__ cvb,
                         const int4* __restrict__ chunks, int nka, int nkb,
                         const int* __restrict__ kpa, const int* __restrict__ kpb,
                         z64* __restrict__ wa, z64* __restrict__ wb, z64* __restrict__ Ov) {
  extern __shared__ z64 sm[];
  z64* soa = sm;                  // [na]
  z64* sob = soa + na;            // [nb]
  z64* part = sob + nb;           // [nka + nkb]
  __shared__ z64 red[256];
  const int w = blockIdx.x, tid = threadIdx.x;
  for (int u = tid; u < na; u += 256) soa[u] = ova[(long)w * na + u];
  for (int v = tid; v < nb; v += 256) sob[v] = ovb[(long)w * nb + v];
  __syncthreads();
  for (int t = tid; t < nka + nkb; t += 256) {
    const bool isa = t < nka;
    const int4 c = chunks[t];
    const int* cc = isa ? cca : ccb;
    const z64* cv = isa ? cva : cvb;
    const z64* other = isa ? sob : soa;
    z64 acc = {0.0, 0.0};
    for (int k = c.y; k < c.z; ++k) acc = zfma(cv[k], other[cc[k]], acc);
    part[t] = acc;
  }
  __syncthreads();
  z64 osum = {0.0, 0.0};
  for (int r = tid; r < na + nb; r += 256) {
    const bool isa = r < na;
    const int u = isa ? r : r - na;
    const int* kp = isa ? kpa : kpb;
    const z64* pp = isa ? part : part + nka;
    z64 acc = {0.0, 0.0};
    for (int i = kp[u]; i < kp[u + 1]; ++i) acc = zadd(acc, pp[i]);
    if (isa) {
      const z64 x = zmul(soa[u], acc);
      wa[(long)w * na + u] = x;
      osum = zadd(osum, x);
    } else {
      wb[(long)w * nb + u] = zmul(sob[u], acc);
    }
  }
  osum = block_sum<256>(osum, red);
  if (tid == 0) Ov[w] = osum;
}

// R[w][j][a] / Ov with R = sum_u w_u (P_u + sum_p e_{h_p} Rc_u[p]) (calculate_R,
// phmsd_helpers.hpp:201-290), walked by hole position j over the excitation rows that empty it:
// hent = (configuration, row of Rc, particle orbital).  S = blockDim / (nel * nact) threads
// share one (j, a): thread `sub` takes every S-th entry, and the S partial sums are added in
// order.  The weights are staged in LDS when they fit (wlds).
__global__ void __launch_bounds__(1024)
    k_ph_R(int nel, int nact, int nconf, int P, const int* __restrict__ hptr,
           const int4* __restrict__ hent, const z64* __restrict__ Rc,
           const z64* __restrict__ wgt, const z64* __restrict__ Ov, z64* __restrict__ Rout,
           int wlds) {
  extern __shared__ z64 sm[];
  const int w = blockIdx.x, tid = threadIdx.x;
  const int NQ = nel * nact;
  const int S = max(1, (int)blockDim.x / NQ);
  z64* part = sm;                    // [S][NQ]
  z64* swg = sm + (long)S * NQ;      // [nconf] when wlds
  const z64* wg = wgt + (long)w * nconf;
  if (wlds) {
    for (int u = tid; u < nconf; u += blockDim.x) swg[u] = wg[u];
    __syncthreads();
    wg = swg;
  }
  const z64* rcw = Rc + (long)w * P * nact;
  for (int t = tid; t < S * NQ; t += blockDim.x) {
    const int q = t % NQ, sub = t / NQ;
    const int j = q / nact, a = q % nact;
    z64 val = {0.0, 0.0};
    for (int k = hptr[j] + sub; k < hptr[j + 1]; k += S) {
      const int4 e = hent[k];
      const z64 wu = wg[e.x];
      z64 rc = rcw[(long)e.y * nact + a];
      if (e.z == a) rc.re += 1.0;
      if (a == j) rc.re -= 1.0;
      val = zfma(wu, rc, val);
    }
    part[(long)sub * NQ + q] = val;
  }
  __syncthreads();
  const z64 ov = Ov[w];
  const z64 inv = zrecip(ov);
  for (int q = tid; q < NQ; q += blockDim.x) {
    const int j = q / nact, a = q % nact;
    z64 val = a == j ? ov : z64{0.0, 0.0};
    for (int sub = 0; sub < S; ++sub) val = zadd(val, part[(long)sub * NQ + q]);
    Rout[((long)w * nel + j) * nact + a] = zmul(val, inv);
  }
}

// Ov[w] = log(Ov[w]) + log_ov[w]
__global__ void k_ph_log(int W, const z64* __restrict__ sum, const z64* __restrict__ log_ov,
                         z64* __restrict__ Ov) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < W) Ov[w] = zadd(zlog(sum[w]), log_ov[w]);
}

__global__ void k_ph_fill_one(int W, int nconf, z64* __restrict__ ovr) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < W) ovr[(long)w * nconf] = {1.0, 0.0};
}

// ---- energy -----------------------------------------------------------------------------------

// T3 [W][NS][nact][nel] <- T3[w][n][a][l] = sum_k Lext[n][a][k] GA[w][l][k]: one GEMM over all
// walkers, rows (n, a), columns (w, l), stored walker-major.
struct EpiStoreWL {
  z64* t3;
  int nel;
  long wstride;  // NS * nact * nel
  __device__ inline void operator()(int, int m, int c, double re, double im) const {
    const int w = c / nel, l = c % nel;
    t3[(long)w * wstride + (long)m * nel + l] = {re, im};
  }
};

// jr[w][n] = sum_i T[i][i] (Jref), tz[w][n] = sum_ik T[k][i] T[i][k] over the reference rows.
__global__ void k_ph_energy_prep(int W, int NS, int nel, int nact, const z64* __restrict__ T3,
                                 z64* __restrict__ jr, z64* __restrict__ tz) {
  const long slice = (long)nact * nel;
  const long wn = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (wn >= (long)W * NS) return;
  const z64* T = T3 + wn * slice;
  z64 j = {0.0, 0.0}, z = {0.0, 0.0};
  for (int i = 0; i < nel; ++i) {
    j = zadd(j, T[i * nel + i]);
    for (int k = 0; k < nel; ++k) z = zfma(T[k * nel + i], T[i * nel + k], z);
  }
  jr[wn] = j;
  tz[wn] = z;
}

// Excitations of rank R, configurations [u0, u1), slices [n0, n0 + nc) of walker w. With
// g_p[l] = sum_q Minv[p][q] V[x_q][l] the excited Y differs from the reference Zref = T[0:nel]
// only in the hole rows, Y[h_p] = Zref[h_p] + g_p, hence
//   Kn = Jref + sum_p g_p[h_p],   eX = sum_pq g_p[h_q] g_q[h_p] + 2 sum_pl g_p[l] T[l][h_p],
// which are tr(Y) and tr(Y^2) - tr(Zref^2) of the reference's per-determinant evaluation.
// When there are fewer excitations than threads, G threads share an excitation and take every
// G-th slice; their sums are combined in a fixed order.
// part [W][nch][nconf][3] <- (e1 on the haj slice, sum eX, sum Kn^2 - Jref^2); KE [W][nconf][M].
template <int R>
__device__ inline void ph_slice_terms(const z64* __restrict__ Ts, int nel, const int (&h)[R],
                                      const int (&x)[R], const z64 (&Mi)[R][R], z64& k, z64& ex) {
  z64 gh[R][R];
  z64 acc = {0.0, 0.0};
  for (int l = 0; l < nel; ++l) {
    z64 V[R];
#pragma unroll
    for (int q = 0; q < R; ++q) V[q] = Ts[x[q] * nel + l];
#pragma unroll
    for (int p = 0; p < R; ++p) {
      z64 g = {0.0, 0.0};
#pragma unroll
      for (int q = 0; q < R; ++q) g = zfma(Mi[p][q], V[q], g);
      acc = zfma(g, Ts[l * nel + h[p]], acc);
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (h[q] == l) gh[p][q] = g;
    }
  }
  k = {0.0, 0.0};
  ex = {2.0 * acc.re, 2.0 * acc.im};
#pragma unroll
  for (int p = 0; p < R; ++p) {
    k = zadd(k, gh[p][p]);
#pragma unroll
    for (int q = 0; q < R; ++q) ex = zfma(gh[p][q], gh[q][p], ex);
  }
}

template <int R>
__global__ void __launch_bounds__(256)
    k_ph_energy_rank(int nel, int nact, int nconf, int M, int NS, int NC, int u0, int u1,
                     const int* __restrict__ exc, const int* __restrict__ mstart, int P2,
                     const z64* __restrict__ T3, const z64* __restrict__ Qv,
                     const z64* __restrict__ jr, const z64* __restrict__ Minv,
                     z64* __restrict__ part, z64* __restrict__ KE) {
  extern __shared__ z64 smem[];
  const int slice = nact * nel, nvir = nact - nel;
  const int chunk = blockIdx.x, w = blockIdx.y, nch = gridDim.x;
  const int n0 = chunk * NC, nc = min(NC, NS - n0);
  z64* red = smem + (long)NC * slice;   // [256][3]
  z64* sq = red + 3 * 256;              // [nvir][nel] QQ0 rows of the particles
  const z64* src = T3 + ((long)w * NS + n0) * slice;
  for (int t = threadIdx.x; t < nc * slice; t += 256) smem[t] = src[t];
  for (int t = threadIdx.x; t < nvir * nel; t += 256) sq[t] = Qv[(long)w * nvir * nel + t];
  __syncthreads();
  // particle rows x >= nel become V[x][l] = T[x][l] - sum_i QQ0[x][i] T[i][l] (the reference
  // rows are only read)
  for (int t = threadIdx.x; t < nc * nvir * nel; t += 256) {
    const int l = t % nel, x = (t / nel) % nvir, c = t / (nvir * nel);
    z64* T = smem + (long)c * slice;
    z64 b = {0.0, 0.0};
    for (int i = 0; i < nel; ++i) b = zfma(sq[x * nel + i], T[i * nel + l], b);
    T[(nel + x) * nel + l] = zsub(T[(nel + x) * nel + l], b);
  }
  __syncthreads();
  const z64* jw = jr + (long)w * NS + n0;
  const int cnt = u1 - u0;
  const int G = cnt >= 256 ? 1 : max(1, min(nc, 256 / cnt));
  const int nitem = cnt >= 256 ? cnt : cnt * G;
  for (int it0 = 0; it0 < nitem; it0 += 256) {
    const int it = it0 + threadIdx.x;
    z64 a0 = {0.0, 0.0}, a1 = {0.0, 0.0}, a2 = {0.0, 0.0};
    if (it < nitem) {
      const int u = u0 + it / G, g = it % G;
      const int* e = exc + (long)u * kXS;
      const z64* mi = Minv + (long)w * P2 + mstart[u];
      int h[R], x[R];
      z64 Mi[R][R];
#pragma unroll
      for (int p = 0; p < R; ++p) {
        h[p] = e[p];
        x[p] = e[R + p];
#pragma unroll
        for (int q = 0; q < R; ++q) Mi[p][q] = mi[p * R + q];
      }
      z64* ke = KE + ((long)w * nconf + u) * M;
      for (int c = g; c < nc; c += G) {
        z64 k, ex;
        ph_slice_terms<R>(smem + (long)c * slice, nel, h, x, Mi, k, ex);
        const int n = n0 + c;
        if (n < M) {
          const z64 kn = zadd(jw[c], k);
          ke[n] = kn;
          a1 = zadd(a1, ex);
          a2 = zadd(a2, zsub(zmul(kn, kn), zmul(jw[c], jw[c])));
        } else {
          a0 = zadd(a0, k);
        }
      }
    }
    if (G == 1) {
      if (it < nitem) {
        z64* pw = part + (((long)w * nch + chunk) * nconf + u0 + it) * 3;
        pw[0] = a0;
        pw[1] = a1;
        pw[2] = a2;
      }
    } else {
      red[3 * threadIdx.x + 0] = a0;
      red[3 * threadIdx.x + 1] = a1;
      red[3 * threadIdx.x + 2] = a2;
      __syncthreads();
      if ((int)threadIdx.x < cnt) {
        z64 s0 = {0.0, 0.0}, s1 = {0.0, 0.0}, s2 = {0.0, 0.0};
        for (int g = 0; g < G; ++g) {
          const z64* r = red + 3 * (threadIdx.x * G + g);
          s0 = zadd(s0, r[0]);
          s1 = zadd(s1, r[1]);
          s2 = zadd(s2, r[2]);
        }
        z64* pw = part + (((long)w * nch + chunk) * nconf + u0 + threadIdx.x) * 3;
        pw[0] = s0;
        pw[1] = s1;
        pw[2] = s2;
      }
    }
  }
}

// KE[w][0][n] = Jref for n < M (the reference configuration's Coulomb vector)
__global__ void k_ph_ke_ref(int W, int M, int NS, int nconf, const z64* __restrict__ jr,
                            z64* __restrict__ KE) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)W * M) return;
  const int w = (int)(t / M), n = (int)(t % M);
  KE[(long)w * nconf * M + n] = jr[(long)w * NS + n];
}

// Opposite-spin Coulomb coupling over Cholesky slices [n0, n0 + CH) of walker w:
//   out[w][chunk] = sum_d conj(ci_d) ova[a_d] ovb[b_d] sum_n Ka[a_d][n] Kb[b_d][n]
// with both spins' slices staged in LDS; determinants in alpha-row order (crow, cca, cva);
// per-thread sums over a fixed set of determinants, then a fixed tree.
__global__ void __launch_bounds__(256)
    k_ph_cross(int na, int nb, int M, int CH, int lds, int ndet, const int* __restrict__ crow,
               const int* __restrict__ ccol, const z64* __restrict__ cval,
               const z64* __restrict__ ova, const z64* __restrict__ ovb,
               const z64* __restrict__ KEa, const z64* __restrict__ KEb, z64* __restrict__ out) {
  extern __shared__ z64 sk[];
  __shared__ z64 red[256];
  const int chunk = blockIdx.x, w = blockIdx.y, nch = gridDim.x, tid = threadIdx.x;
  const int n0 = chunk * CH, cn = min(CH, M - n0);
  const z64* Ka = KEa + (long)w * na * M;
  const z64* Kb = KEb + (long)w * nb * M;
  // slices in LDS ([na][CH] then [nb][CH]) when they fit, else read in place
  const z64* sa = Ka + n0;
  const z64* sb = Kb + n0;
  long ld = M;
  if (lds) {
    z64* da = sk;
    z64* db = sk + (long)na * CH;
    for (int t = tid; t < na * CH; t += 256) {
      const int u = t / CH, j = t % CH;
      da[t] = j < cn ? Ka[(long)u * M + n0 + j] : z64{0.0, 0.0};
    }
    for (int t = tid; t < nb * CH; t += 256) {
      const int v = t / CH, j = t % CH;
      db[t] = j < cn ? Kb[(long)v * M + n0 + j] : z64{0.0, 0.0};
    }
    __syncthreads();
    sa = da;
    sb = db;
    ld = CH;
  }
  const z64* oa = ova + (long)w * na;
  const z64* ob = ovb + (long)w * nb;
  z64 acc = {0.0, 0.0};
  for (int d = tid; d < ndet; d += 256) {
    const int u = crow[d], v = ccol[d];
    z64 dot = {0.0, 0.0};
    for (int j = 0; j < cn; ++j) dot = zfma(sa[u * ld + j], sb[v * ld + j], dot);
    acc = zfma(zmul(cval[d], zmul(oa[u], ob[v])), dot, acc);
  }
  acc = block_sum<256>(acc, red);
  if (tid == 0) out[(long)w * nch + chunk] = acc;
}

// energy_alg1 assembly per walker: reference terms times Ov, weighted excitation terms of both
// spins, the opposite-spin Coulomb coupling sum_d conj(ci) ova ovb Ka . Kb (partial sums over
// Cholesky chunks in cross[w][0..ncross)), then / Ov.
__global__ void __launch_bounds__(256)
    k_ph_energy_combine(int na, int nb, int ncha, int nchb, int M, int NS, double E0,
                        const z64* __restrict__ pa, const z64* __restrict__ pb,
                        const z64* __restrict__ jra, const z64* __restrict__ tza,
                        const z64* __restrict__ jrb, const z64* __restrict__ tzb,
                        const z64* __restrict__ wa, const z64* __restrict__ wb,
                        const z64* __restrict__ Osum, const z64* __restrict__ log_ov,
                        const z64* __restrict__ cross, int ncross, z64* __restrict__ E,
                        z64* __restrict__ Ov) {
  __shared__ z64 red[256];
  const int w = blockIdx.x;
  const z64 ov = Osum[w];
  z64 e1 = {0.0, 0.0}, ex = {0.0, 0.0}, ej = {0.0, 0.0};
  for (int s = 0; s < 2; ++s) {
    const int nconf = s == 0 ? na : nb, nch = s == 0 ? ncha : nchb;
    const z64* p = (s == 0 ? pa : pb) + (long)w * nch * nconf * 3;
    const z64* wg = (s == 0 ? wa : wb) + (long)w * nconf;
    for (int u = 1 + threadIdx.x; u < nconf; u += 256) {
      z64 a0 = {0.0, 0.0}, a1 = {0.0, 0.0}, a2 = {0.0, 0.0};
      for (int c = 0; c < nch; ++c) {
        a0 = zadd(a0, p[((long)c * nconf + u) * 3 + 0]);
        a1 = zadd(a1, p[((long)c * nconf + u) * 3 + 1]);
        a2 = zadd(a2, p[((long)c * nconf + u) * 3 + 2]);
      }
      const z64 f = wg[u];
      e1 = zfma(f, a0, e1);
      ex = zfma(f, {-0.5 * a1.re, -0.5 * a1.im}, ex);
      ej = zfma(f, {0.5 * a2.re, 0.5 * a2.im}, ej);
    }
    // reference configuration: (E1ref, EXref, EJref) of this spin times Ov
    const z64* jr = (s == 0 ? jra : jrb) + (long)w * NS;
    const z64* tz = (s == 0 ? tza : tzb) + (long)w * NS;
    z64 r1 = {0.0, 0.0}, r2 = {0.0, 0.0};
    for (int n = threadIdx.x; n < M; n += 256) {
      r1 = zadd(r1, tz[n]);
      r2 = zfma(jr[n], jr[n], r2);
    }
    ex = zfma(ov, {-0.5 * r1.re, -0.5 * r1.im}, ex);
    ej = zfma(ov, {0.5 * r2.re, 0.5 * r2.im}, ej);
    if (threadIdx.x == 0) e1 = zfma(ov, jr[M], e1);
  }
  e1 = block_sum<256>(e1, red);
  ex = block_sum<256>(ex, red);
  ej = block_sum<256>(ej, red);
  if (threadIdx.x == 0) {
    for (int c = 0; c < ncross; ++c) ej = zadd(ej, cross[(long)w * ncross + c]);
    const z64 inv = zrecip(ov);
    z64 a = zmul(e1, inv);
    E[(long)w * 3 + 0] = {a.re + E0, a.im};
    E[(long)w * 3 + 1] = zmul(ex, inv);
    E[(long)w * 3 + 2] = zmul(ej, inv);
    Ov[w] = zadd(zlog(ov), log_ov[w]);
  }
}

// ---- trial density for vMF --------------------------------------------------------------------

// A[u][i][k] = orbs[orb_u(i)][k]
__global__ void k_ph_gather_rows(int nconf, int nel, int N, const int* __restrict__ occ,
                                 const z64* __restrict__ orbs, z64* __restrict__ A) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nconf * nel * N) return;
  const int k = (int)(t % N);
  const long ui = t / N;
  A[t] = orbs[(long)occ[ui] * N + k];
}

// r[u] = exp(l[u] - l[0]) for u >= 1, r[0] = 1   (PHMSD.cpp:1384-1389)
__global__ void k_ph_ratio_from_log(int nconf, const z64* __restrict__ l, z64* __restrict__ r) {
  int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nconf) return;
  if (u == 0) {
    r[0] = {1.0, 0.0};
    return;
  }
  z64 d = zsub(l[u], l[0]);
  double m = exp(d.re);
  r[u] = {m * cos(d.im), m * sin(d.im)};
}

// ov = sum_d |ci|^2 ra[a_d] rb[b_d]; wd_s[u] = r_s[u] sum_{row u of C_s} r_other[col] |c|^2
__global__ void __launch_bounds__(256)
    k_ph_mf_weights(int ndet, const int* __restrict__ det_ab, const z64* __restrict__ det_c,
                    const z64* __restrict__ ra, const z64* __restrict__ rb, int na, int nb,
                    const int* __restrict__ cpa, const int* __restrict__ cca,
                    const z64* __restrict__ cva, const int* __restrict__ cpb,
                    const int* __restrict__ ccb, const z64* __restrict__ cvb,
                    z64* __restrict__ wda, z64* __restrict__ wdb, z64* __restrict__ ov) {
  __shared__ z64 red[256];
  z64 s = {0.0, 0.0};
  for (int d = threadIdx.x; d < ndet; d += 256) {
    z64 c = det_c[d];
    double n2 = c.re * c.re + c.im * c.im;
    z64 t = zmul(ra[det_ab[2 * d]], rb[det_ab[2 * d + 1]]);
    s = zadd(s, {n2 * t.re, n2 * t.im});
  }
  for (int u = threadIdx.x; u < na; u += 256) {
    z64 a = {0.0, 0.0};
    for (int k = cpa[u]; k < cpa[u + 1]; ++k) {
      z64 c = cva[k];
      double n2 = c.re * c.re + c.im * c.im;
      z64 r = rb[cca[k]];
      a = zadd(a, {n2 * r.re, n2 * r.im});
    }
    wda[u] = zmul(a, ra[u]);
  }
  for (int v = threadIdx.x; v < nb; v += 256) {
    z64 a = {0.0, 0.0};
    for (int k = cpb[v]; k < cpb[v + 1]; ++k) {
      z64 c = cvb[k];
      double n2 = c.re * c.re + c.im * c.im;
      z64 r = ra[ccb[k]];
      a = zadd(a, {n2 * r.re, n2 * r.im});
    }
    wdb[v] = zmul(a, rb[v]);
  }
  z64 tot = block_sum<256>(s, red);
  if (threadIdx.x == 0) ov[0] = tot;
}

// Gs[a][k] = (sum_u wd[u] sum_{i: occ_u(i) = a} G2[u][i][k] + single-excitation couplings) / ov
// pairs: (hole orbital, particle orbital, d1, d2) with weight conj(ci_d1) ci_d2.
__global__ void k_ph_mf_density(int nconf, int nel, int nact, int N, const int* __restrict__ occ,
                                const z64* __restrict__ wd, const z64* __restrict__ G2, int npair,
                                const int* __restrict__ pairs, const z64* __restrict__ det_c,
                                const z64* __restrict__ orbs, const z64* __restrict__ ov,
                                z64* __restrict__ Gs) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nact * N) return;
  const int a = (int)(t / N), k = (int)(t % N);
  z64 g = {0.0, 0.0};
  for (int u = 0; u < nconf; ++u)
    for (int i = 0; i < nel; ++i)
      if (occ[(long)u * nel + i] == a) g = zfma(wd[u], G2[((long)u * nel + i) * N + k], g);
  for (int q = 0; q < npair; ++q) {
    const int hole = pairs[4 * q], part = pairs[4 * q + 1];
    // det_c holds conj(ci): wgt = val(n1) conj(val(n2)) = det_c[d1] conj(det_c[d2])
    const z64 wgt = zmul(det_c[pairs[4 * q + 2]], zconj(det_c[pairs[4 * q + 3]]));
    if (part == a) g = zfma(wgt, orbs[(long)hole * N + k], g);
    if (hole == a) g = zfma(zconj(wgt), orbs[(long)part * N + k], g);
  }
  Gs[t] = zmul(g, zrecip(ov[0]));
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// PhmsdTrial
// ---------------------------------------------------------------------------------------------

PhmsdTrial::PhmsdTrial(CholeskyOps& ham, const PhExcitations& ex) : ham_(ham), ctx_(ham.ctx()) {
  AFQ_CHECK(ham.walker_type() == kCollinear, kErrArg, "PHMSD requires a COLLINEAR hamiltonian");
  AFQ_CHECK(ex.ndet > 0 && ex.det_a && ex.det_b && ex.ci, kErrArg, "empty expansion");
  hipStream_t s = ctx_.stream();
  const int N = ham.N(), M = ham.M();
  const int nact[2] = {ham.nup(), ham.ndown()};
  try {
    for (int sp = 0; sp < 2; ++sp) {
      SpinDev& d = sp_[sp];
      d.nel = ex.nel[sp];
      d.nact = nact[sp];
      d.nvir = d.nact - d.nel;
      d.nconf = ex.nconf[sp];
      AFQ_CHECK(d.nel > 0 && d.nel <= d.nact, kErrShape, "electrons must be in 1..nact");
      AFQ_CHECK(d.nconf >= 1 && ex.rank[sp] && ex.exc_start[sp] && ex.exc[sp] && ex.refc[sp],
                kErrArg, "excitation arrays missing");
      for (int i = 0; i < d.nel; ++i)
        AFQ_CHECK(ex.refc[sp][i] == i, kErrArg, "the woodbury algorithm requires refc[i] == i");
      AFQ_CHECK(ex.rank[sp][0] == 0, kErrArg, "configuration 0 must be the reference");
      d.rank_h.assign(ex.rank[sp], ex.rank[sp] + d.nconf);
      d.exc_h.assign((size_t)d.nconf * kXS, 0);
      std::vector<int> mstart(d.nconf, 0), rstart(d.nconf, 0);
      std::vector<std::vector<int>> byhole(d.nel);
      d.rank_off.assign(kMaxRank + 2, d.nconf);
      d.rank_off[0] = 0;
      int prev = 0;
      for (int u = 1; u < d.nconf; ++u) {
        int r = d.rank_h[u];
        AFQ_CHECK(r >= 1 && r <= kMaxRank, kErrShape,
                  "excitation rank must be in 1.." + std::to_string(kMaxRank));
        AFQ_CHECK(r >= prev, kErrArg, "configurations must be ordered by excitation rank");
        for (int q = prev + 1; q <= r; ++q) d.rank_off[q] = u;
        prev = r;
        d.rmax = std::max(d.rmax, r);
        const int* e = ex.exc[sp] + ex.exc_start[sp][u];
        for (int p = 0; p < r; ++p) {
          AFQ_CHECK(e[p] >= 0 && e[p] < d.nel, kErrArg, "hole position out of range");
          AFQ_CHECK(e[r + p] >= d.nel && e[r + p] < d.nact, kErrArg,
                    "particle must be an active orbital outside the reference");
          d.exc_h[(size_t)u * kXS + p] = e[p];
          d.exc_h[(size_t)u * kXS + r + p] = e[r + p];
          byhole[e[p]].insert(byhole[e[p]].end(), {u, d.P + p, e[r + p], 0});
        }
        mstart[u] = d.P2;
        rstart[u] = d.P;
        d.P += r;
        d.P2 += r * r;
      }
      for (int q = prev + 1; q <= kMaxRank + 1; ++q) d.rank_off[q] = d.nconf;
      std::vector<int> hptr(d.nel + 1, 0), hent;   // hent: (configuration, Rc row, particle, 0)
      for (int j = 0; j < d.nel; ++j) {
        hptr[j] = (int)hent.size() / 4;
        hent.insert(hent.end(), byhole[j].begin(), byhole[j].end());
      }
      hptr[d.nel] = (int)hent.size() / 4;
      for (int j = 0; j < d.nel; ++j) d.hlen_max = std::max(d.hlen_max, hptr[j + 1] - hptr[j]);
      if (hent.empty()) hent.assign(4, 0);
      d.exc = dcopy(s, d.exc_h);
      d.rank = dcopy(s, d.rank_h);
      d.mstart = dcopy(s, mstart);
      d.rstart = dcopy(s, rstart);
      d.hptr = dcopy(s, hptr);
      d.hent = dcopy(s, hent);
      d.orbs = ham.PsiT() + (sp == 0 ? 0 : (size_t)nact[0] * N);
      // Lext = [Lnak of this spin; haj of this spin]
      d.Lext = dalloc<z64>((size_t)(M + 1) * d.nact * N);
      const z64* Ls = ham.Lnak() + (sp == 0 ? 0 : (size_t)M * nact[0] * N);
      AFQ_HIP(hipMemcpyAsync(d.Lext, Ls, sizeof(z64) * M * d.nact * N, hipMemcpyDeviceToDevice, s));
      AFQ_HIP(hipMemcpyAsync(d.Lext + (size_t)M * d.nact * N,
                             ham.haj() + (sp == 0 ? 0 : (size_t)nact[0] * N),
                             sizeof(z64) * d.nact * N, hipMemcpyDeviceToDevice, s));
    }
    // couplings (Wavefunction.cpp:425-451): row u of spin s lists the other spin's
    // configurations with conj(ci), in determinant order
    ndet_ = ex.ndet;
    std::vector<std::vector<std::pair<int, z64>>> rows[2];
    std::vector<int> chunk_h;   // (row, begin, end, 0) per chunk, alpha rows then beta rows
    rows[0].resize(sp_[0].nconf);
    rows[1].resize(sp_[1].nconf);
    det_ab_h_.resize(2 * (size_t)ndet_);
    det_c_h_.resize(ndet_);
    for (int dd = 0; dd < ndet_; ++dd) {
      int a = ex.det_a[dd], b = ex.det_b[dd];
      AFQ_CHECK(a >= 0 && a < sp_[0].nconf && b >= 0 && b < sp_[1].nconf, kErrArg,
                "determinant refers to an unknown configuration");
      z64 c = {ex.ci[dd].re, -ex.ci[dd].im};
      rows[0][a].push_back({b, c});
      rows[1][b].push_back({a, c});
      det_ab_h_[2 * dd] = a;
      det_ab_h_[2 * dd + 1] = b;
      det_c_h_[dd] = c;
    }
    for (int sp = 0; sp < 2; ++sp) {
      std::vector<int> ptr(1, 0), col;
      std::vector<z64> val;
      for (auto& r : rows[sp]) {
        for (auto& kv : r) {
          col.push_back(kv.first);
          val.push_back(kv.second);
        }
        ptr.push_back((int)col.size());
      }
      sp_[sp].cptr = dcopy(s, ptr);
      sp_[sp].ccol = dcopy(s, col);
      sp_[sp].cval = dcopy(s, val);
      if (sp == 0) {  // alpha row of every coupling (determinants in alpha-row order)
        std::vector<int> crow(col.size() ? col.size() : 1, 0);
        for (int u = 0; u < sp_[0].nconf; ++u)
          for (int k = ptr[u]; k < ptr[u + 1]; ++k) crow[k] = u;
        crow_ = dcopy(s, crow);
      }
      // chunks of at most kChunk couplings per row, rows in order
      std::vector<int> kp(1, 0);
      for (int u = 0; u < sp_[sp].nconf; ++u) {
        for (int k0 = ptr[u]; k0 < ptr[u + 1]; k0 += kChunk)
          chunk_h.insert(chunk_h.end(), {u, k0, std::min(k0 + kChunk, ptr[u + 1]), 0});
        kp.push_back((int)chunk_h.size() / 4 - (sp == 0 ? 0 : (int)sp_[0].nchunk));
      }
      sp_[sp].nchunk = kp.back();
      sp_[sp].kptr = dcopy(s, kp);
    }
    if (chunk_h.empty()) chunk_h.assign(4, 0);
    chunks_ = reinterpret_cast<int4*>(dcopy(s, chunk_h));
    det_ab_ = dcopy(s, det_ab_h_);
    det_c_ = dcopy(s, det_c_h_);
    AFQ_HIP(hipStreamSynchronize(s));
  } catch (...) {
    free_all();
    throw;
  }
}

PhmsdTrial::~PhmsdTrial() {
  (void)hipStreamSynchronize(ctx_.stream());
  free_all();
}

void PhmsdTrial::free_all() {
  for (auto& d : sp_)
    for (void* p : {(void*)d.exc, (void*)d.rank, (void*)d.mstart, (void*)d.rstart, (void*)d.hptr, (void*)d.hent,
                    (void*)d.cptr, (void*)d.ccol, (void*)d.cval, (void*)d.kptr, (void*)d.Lext})
      if (p) (void)hipFree(p);
  for (void* p : {(void*)det_ab_, (void*)det_c_, (void*)chunks_, (void*)crow_})
    if (p) (void)hipFree(p);
  for (auto& d : sp_) d = SpinDev{};
  det_ab_ = nullptr;
  det_c_ = nullptr;
  chunks_ = nullptr;
  crow_ = nullptr;
}

static std::string tag(const char* base, int s) { return std::string(base) + (s ? "1" : "0"); }

void PhmsdTrial::woodbury(int s, const z64* Phi, long phi_bs, int W, bool want_G, z64* log_ov) {
  using gemm::EpiStore;
  using gemm::Opnd;
  const SpinDev& d = sp_[s];
  const int N = ham_.N(), n = d.nel, na = d.nact;
  if (phi_bs == 0) phi_bs = (long)N * n;
  AFQ_CHECK(phi_bs >= (long)N * n, kErrShape, "walker stride below the Slater block size");
  z64* TAB = scratch_as<z64>(ctx_, "ph_TAB", (size_t)W * na * n);
  z64* Ti = scratch_as<z64>(ctx_, tag("ph_Ti", s), (size_t)W * n * n);
  z64* Qv = scratch_as<z64>(ctx_, tag("ph_Qv", s), (size_t)W * std::max(d.nvir, 1) * n);
  {  // TAB = orbs * Phi
    Opnd<double, true, true> A{&d.orbs->re, N, 0};
    Opnd<double, true, false> B{&Phi->re, n, phi_bs};
    zgemm_small(ctx_, "ph_TAB", W, na, n, N, A, B,
                       EpiStore<double, double>{TAB, n, (long)na * n, 1, 0, 0});
  }
  lu_inverse_logdet(ctx_.stream(), TAB, (long)na * n, W, n, Ti, (long)n * n, log_ov, 1.0, {0, 0},
                    true);
  if (d.nvir > 0) {  // QQ0 particle rows = TAB[nel:] * TNN^{-1}
    Opnd<double, true, true> A{&(TAB + (size_t)n * n)->re, n, (long)na * n};
    Opnd<double, true, false> B{&Ti->re, n, (long)n * n};
    zgemm_small(ctx_, "ph_QQ0", W, d.nvir, n, n, A, B,
                       EpiStore<double, double>{Qv, n, (long)d.nvir * n, 1, 0, 0});
  }
  if (want_G) {  // GA = TNN^{-T} Phi^T
    z64* GA = scratch_as<z64>(ctx_, tag("ph_GA", s), (size_t)W * n * N);
    Opnd<double, true, false> A{&Ti->re, n, (long)n * n};
    Opnd<double, true, true> B{&Phi->re, n, phi_bs};
    zgemm_small(ctx_, "mdm_G", W, n, N, n, A, B,
                       EpiStore<double, double>{GA, N, (long)n * N, 1, 0, 0});
  }
}

void PhmsdTrial::ratios(int s, int W, bool want_inverse) {
  const SpinDev& d = sp_[s];
  hipStream_t st = ctx_.stream();
  z64* ovr = scratch_as<z64>(ctx_, tag("ph_ovr", s), (size_t)W * d.nconf);
  z64* Mi = scratch_as<z64>(ctx_, tag("ph_Minv", s), (size_t)W * std::max(d.P2, 1));
  const z64* Qv = scratch_as<z64>(ctx_, tag("ph_Qv", s), (size_t)W * std::max(d.nvir, 1) * d.nel);
  z64* Rc = want_inverse ? scratch_as<z64>(ctx_, tag("ph_Rc", s), (size_t)W * std::max(d.P, 1) * d.nact)
                         : nullptr;
  hipLaunchKernelGGL(k_ph_fill_one, dim3(ceil_div(W, 256)), dim3(256), 0, st, W, d.nconf, ovr);
  AFQ_LAUNCH_CHECK();
  for (int r = 1; r <= d.rmax; ++r) {
    const int u0 = d.rank_off[r], cnt = d.rank_off[r + 1] - u0;
    if (cnt <= 0) continue;
    dim3 g(ceil_div((long)W * cnt, 256));
    const int wi = want_inverse ? 1 : 0;
#define AFQ_RATIOS(R)                                                                          \
  hipLaunchKernelGGL(k_ph_ratios<R>, g, dim3(256), 0, st, W, d.nel, d.nvir, u0, cnt, d.nconf, \
                     d.exc, d.mstart, d.P2, Qv, ovr, Mi, wi, d.rstart, d.P, Rc)
    switch (r) {
      case 1: AFQ_RATIOS(1); break;
      case 2: AFQ_RATIOS(2); break;
      case 3: AFQ_RATIOS(3); break;
      case 4: AFQ_RATIOS(4); break;
      case 5: AFQ_RATIOS(5); break;
      default: AFQ_RATIOS(6); break;
    }
#undef AFQ_RATIOS
    AFQ_LAUNCH_CHECK();
  }
}

void PhmsdTrial::weights(int W, z64* Ov) {
  const SpinDev &a = sp_[0], &b = sp_[1];
  z64* ova = scratch_as<z64>(ctx_, "ph_ovr0", (size_t)W * a.nconf);
  z64* ovb = scratch_as<z64>(ctx_, "ph_ovr1", (size_t)W * b.nconf);
  z64* wa = scratch_as<z64>(ctx_, "ph_w0", (size_t)W * a.nconf);
  z64* wb = scratch_as<z64>(ctx_, "ph_w1", (size_t)W * b.nconf);
  const int nka = (int)a.nchunk, nkb = (int)b.nchunk;
  const size_t lds = sizeof(z64) * ((size_t)a.nconf + b.nconf + nka + nkb);
  if (lds <= 48 * 1024) {
    hipLaunchKernelGGL(k_ph_weights_chunked, dim3(W), dim3(256), lds, ctx_.stream(), a.nconf,
                       b.nconf, ova, ovb, a.ccol, a.cval, b.ccol, b.cval, chunks_, nka, nkb,
                       a.kptr, b.kptr, wa, wb, Ov);
  } else {
    hipLaunchKernelGGL(k_ph_weights, dim3(W), dim3(256), 0, ctx_.stream(), a.nconf, b.nconf, ova,
                       ovb, a.cptr, a.ccol, a.cval, b.cptr, b.ccol, b.cval, wa, wb, Ov);
  }
  AFQ_LAUNCH_CHECK();
}

void PhmsdTrial::mixed_density_matrix(const z64* Pa, const z64* Pb, int W, z64* G, z64* Ov,
                                      long bs_a, long bs_b) {
  using gemm::EpiStore;
  using gemm::Opnd;
  if (W <= 0) return;
  hipStream_t s = ctx_.stream();
  const int N = ham_.N(), ntot = sp_[0].nact + sp_[1].nact;
  z64* log_ov = scratch_as<z64>(ctx_, "ph_logov", W);
  z64* Osum = scratch_as<z64>(ctx_, "ph_Osum", W);
  fill_z64(s, W, {0, 0}, log_ov);
  woodbury(0, Pa, bs_a, W, true, log_ov);
  ratios(0, W, true);
  woodbury(1, Pb, bs_b, W, true, log_ov);
  ratios(1, W, true);
  weights(W, Osum);
  for (int sp = 0; sp < 2; ++sp) {
    const SpinDev& d = sp_[sp];
    z64* R = scratch_as<z64>(ctx_, "ph_R", (size_t)W * d.nel * d.nact);
    {
      const int NQ = d.nel * d.nact;
      // enough threads per (j, a) to cover the longest hole list in about 8 passes; the kernel
      // derives the same S = threads / NQ
      const int S0 = std::max(1, std::min({1024 / NQ, ceil_div(d.hlen_max, 8), 32}));
      const int threads = NQ >= 1024 ? 1024 : std::min(1024, ((S0 * NQ + 63) / 64) * 64);
      const int S = std::max(1, threads / NQ);
      const size_t base = sizeof(z64) * (size_t)S * NQ;
      const int wlds = base + sizeof(z64) * (size_t)d.nconf <= 48 * 1024 ? 1 : 0;
      AFQ_CHECK(base <= 48 * 1024, kErrShape, "active space too large for the R assembly kernel");
      const size_t lds = base + (wlds ? sizeof(z64) * (size_t)d.nconf : 0);
      hipLaunchKernelGGL(k_ph_R, dim3(W), dim3(threads), lds, s, d.nel, d.nact, d.nconf, d.P,
                         d.hptr, reinterpret_cast<const int4*>(d.hent),
                         scratch_as<z64>(ctx_, tag("ph_Rc", sp), (size_t)W * std::max(d.P, 1) * d.nact),
                         scratch_as<z64>(ctx_, tag("ph_w", sp), (size_t)W * d.nconf), Osum, R,
                         wlds);
      AFQ_LAUNCH_CHECK();
    }
    // G rows of this spin: G[a][k] = sum_j R[j][a] GA[j][k]
    const z64* GA = scratch_as<z64>(ctx_, tag("ph_GA", sp), (size_t)W * d.nel * N);
    Opnd<double, true, false> A{&R->re, d.nact, (long)d.nel * d.nact};
    Opnd<double, true, false> B{&GA->re, N, (long)d.nel * N};
    zgemm_small(ctx_, "ph_G", W, d.nact, N, d.nel, A, B,
                       EpiStore<double, double>{G + (sp == 0 ? 0 : (size_t)sp_[0].nact * N), N,
                                                (long)ntot * N, 1, 0, 0});
  }
  hipLaunchKernelGGL(k_ph_log, dim3(ceil_div(W, 256)), dim3(256), 0, s, W, Osum, log_ov, Ov);
  AFQ_LAUNCH_CHECK();
}

void PhmsdTrial::log_overlap(const z64* Pa, const z64* Pb, int W, z64* Ov, long bs_a, long bs_b) {
  if (W <= 0) return;
  hipStream_t s = ctx_.stream();
  z64* log_ov = scratch_as<z64>(ctx_, "ph_logov", W);
  z64* Osum = scratch_as<z64>(ctx_, "ph_Osum", W);
  fill_z64(s, W, {0, 0}, log_ov);
  woodbury(0, Pa, bs_a, W, false, log_ov);
  ratios(0, W, false);
  woodbury(1, Pb, bs_b, W, false, log_ov);
  ratios(1, W, false);
  weights(W, Osum);
  hipLaunchKernelGGL(k_ph_log, dim3(ceil_div(W, 256)), dim3(256), 0, s, W, Osum, log_ov, Ov);
  AFQ_LAUNCH_CHECK();
}

void PhmsdTrial::energy(const z64* Pa, const z64* Pb, int W, z64* E, z64* Ov, long bs_a,
                        long bs_b) {
  using gemm::EpiStore;
  using gemm::Opnd;
  if (W <= 0) return;
  hipStream_t s = ctx_.stream();
  const int N = ham_.N(), M = ham_.M(), NS = M + 1;
  z64* log_ov = scratch_as<z64>(ctx_, "ph_logov", W);
  z64* Osum = scratch_as<z64>(ctx_, "ph_Osum", W);
  fill_z64(s, W, {0, 0}, log_ov);
  woodbury(0, Pa, bs_a, W, true, log_ov);
  ratios(0, W, true);
  woodbury(1, Pb, bs_b, W, true, log_ov);
  ratios(1, W, true);
  weights(W, Osum);
  int nch[2];
  z64* part[2];
  z64* KE[2];
  z64* jr[2];
  z64* tz[2];
  for (int sp = 0; sp < 2; ++sp) {
    const SpinDev& d = sp_[sp];
    const long slice = (long)d.nact * d.nel;
    // LDS: NC slices, the [256][3] reduction buffer and the particle rows of QQ0 within 64 KB
    const long budget = 64L * 1024 - (long)sizeof(z64) * (3 * 256 + (long)d.nvir * d.nel);
    const int NC = (int)std::min<long>(32, budget / (slice * (long)sizeof(z64)));
    AFQ_CHECK(NC >= 1, kErrShape, "active space too large for the PHMSD energy kernel");
    nch[sp] = ceil_div(NS, NC);
    z64* T3 = scratch_as<z64>(ctx_, "ph_T3", (size_t)W * NS * slice);
    {  // T3[w][n][a][l] = sum_k Lext[n][a][k] GA[w][l][k]: rows (n, a), columns (w, l)
      const z64* GA = scratch_as<z64>(ctx_, tag("ph_GA", sp), (size_t)W * d.nel * N);
      Opnd<double, true, true> A{&d.Lext->re, N, 0};
      Opnd<double, true, true> B{&GA->re, N, 0};
      gemm_large<double>(ctx_, "ph_T3", 1, NS * d.nact, W * d.nel, N, A, B,
                         EpiStoreWL{T3, d.nel, (long)NS * slice});
    }
    jr[sp] = scratch_as<z64>(ctx_, tag("ph_jr", sp), (size_t)W * NS);
    tz[sp] = scratch_as<z64>(ctx_, tag("ph_tz", sp), (size_t)W * NS);
    const z64* Qv = scratch_as<z64>(ctx_, tag("ph_Qv", sp), (size_t)W * std::max(d.nvir, 1) * d.nel);
    hipLaunchKernelGGL(k_ph_energy_prep, dim3(ceil_div((long)W * NS, 256)), dim3(256), 0, s, W, NS,
                       d.nel, d.nact, T3, jr[sp], tz[sp]);
    AFQ_LAUNCH_CHECK();
    part[sp] = scratch_as<z64>(ctx_, tag("ph_part", sp), (size_t)W * nch[sp] * d.nconf * 3);
    KE[sp] = scratch_as<z64>(ctx_, tag("ph_KE", sp), (size_t)W * M * d.nconf);
    hipLaunchKernelGGL(k_ph_ke_ref, dim3(ceil_div((long)W * M, 256)), dim3(256), 0, s, W, M, NS,
                       d.nconf, jr[sp], KE[sp]);
    AFQ_LAUNCH_CHECK();
    const z64* Mi = scratch_as<z64>(ctx_, tag("ph_Minv", sp), (size_t)W * std::max(d.P2, 1));
    const size_t shm = sizeof(z64) * ((size_t)NC * slice + 3 * 256 + (size_t)d.nvir * d.nel);
    for (int r = 1; r <= d.rmax; ++r) {
      const int u0 = d.rank_off[r], u1 = d.rank_off[r + 1];
      if (u1 <= u0) continue;
      dim3 g(nch[sp], W);
#define AFQ_ERANK(R)                                                                           \
  hipLaunchKernelGGL(k_ph_energy_rank<R>, g, dim3(256), shm, s, d.nel, d.nact, d.nconf, M, NS, \
                     NC, u0, u1, d.exc, d.mstart, d.P2, T3, Qv, jr[sp], Mi, part[sp], KE[sp])
      switch (r) {
        case 1: AFQ_ERANK(1); break;
        case 2: AFQ_ERANK(2); break;
        case 3: AFQ_ERANK(3); break;
        case 4: AFQ_ERANK(4); break;
        case 5: AFQ_ERANK(5); break;
        default: AFQ_ERANK(6); break;
      }
#undef AFQ_ERANK
      AFQ_LAUNCH_CHECK();
    }
  }
  // opposite-spin Coulomb term over Cholesky chunks of CH slices
  const int na = sp_[0].nconf, nb = sp_[1].nconf;
  // (both spins' slices of a chunk in 64 KB of LDS; read in place for larger expansions)
  const int CHl = (int)std::min<long>(16, (64L * 1024) / ((long)sizeof(z64) * (na + nb)));
  const int use_lds = CHl >= 1 ? 1 : 0;
  const int CH = use_lds ? CHl : 16;
  const int ncross = ceil_div(M, CH);
  z64* cross = scratch_as<z64>(ctx_, "ph_cross", (size_t)W * ncross);
  hipLaunchKernelGGL(k_ph_cross, dim3(ncross, W), dim3(256),
                     use_lds ? sizeof(z64) * (size_t)(na + nb) * CH : 0, s, na, nb, M, CH,
                     use_lds, ndet_, crow_, sp_[0].ccol, sp_[0].cval,
                     scratch_as<z64>(ctx_, "ph_ovr0", (size_t)W * na),
                     scratch_as<z64>(ctx_, "ph_ovr1", (size_t)W * nb), KE[0], KE[1], cross);
  AFQ_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ph_energy_combine, dim3(W), dim3(256), 0, s, sp_[0].nconf, sp_[1].nconf,
                     nch[0], nch[1], M, NS, ham_.E0(), part[0], part[1], jr[0], tz[0],
                     jr[1], tz[1], scratch_as<z64>(ctx_, "ph_w0", (size_t)W * sp_[0].nconf),
                     scratch_as<z64>(ctx_, "ph_w1", (size_t)W * sp_[1].nconf), Osum, log_ov,
                     cross, ncross, E, Ov);
  AFQ_LAUNCH_CHECK();
}

std::vector<int> PhmsdTrial::occupations(int sp) const {
  // configurations in place (ph_excitations::get_configuration, Excitations.hpp:442-483)
  const SpinDev& d = sp_[sp];
  std::vector<int> occ((size_t)d.nconf * d.nel);
  for (int u = 0; u < d.nconf; ++u) {
    for (int i = 0; i < d.nel; ++i) occ[(size_t)u * d.nel + i] = i;
    const int r = d.rank_h[u];
    for (int p = 0; p < r; ++p)
      occ[(size_t)u * d.nel + d.exc_h[(size_t)u * kXS + p]] = d.exc_h[(size_t)u * kXS + r + p];
  }
  return occ;
}

std::vector<int> PhmsdTrial::reference_orbitals(int sp) const {
  const SpinDev& d = sp_[sp];
  const std::vector<int> occ = occupations(sp);
  std::vector<int> idx((size_t)ndet_ * d.nel);
  for (int dd = 0; dd < ndet_; ++dd) {
    const int u = det_ab_h_[2 * dd + sp];
    std::copy_n(occ.begin() + (size_t)u * d.nel, d.nel, idx.begin() + (size_t)dd * d.nel);
  }
  return idx;
}

void PhmsdTrial::references(z64* rows0, z64* rows1, z64* cc) {
  hipStream_t s = ctx_.stream();
  const int N = ham_.N();
  z64* rows[2] = {rows0, rows1};
  for (int sp = 0; sp < 2; ++sp) {
    const SpinDev& d = sp_[sp];
    const std::vector<int> idx = reference_orbitals(sp);
    int* idx_d = dcopy(s, idx);
    gather_rows(s, (long)ndet_ * d.nel, N, idx_d, d.orbs, rows[sp]);
    AFQ_HIP(hipStreamSynchronize(s));
    (void)hipFree(idx_d);
  }
  AFQ_HIP(hipMemcpyAsync(cc, det_c_h_.data(), sizeof(z64) * ndet_, hipMemcpyHostToDevice, s));
  AFQ_HIP(hipStreamSynchronize(s));
}

void PhmsdTrial::vMF(double dt, z64* v) {
  using gemm::EpiStore;
  using gemm::Opnd;
  hipStream_t s = ctx_.stream();
  const int N = ham_.N(), ntot = sp_[0].nact + sp_[1].nact;
  z64* Gfull = scratch_as<z64>(ctx_, "phmf_G", (size_t)ntot * N);
  z64* rat[2];
  z64* wd[2];
  std::vector<int> occ_h[2] = {occupations(0), occupations(1)};
  std::vector<void*> tmp;
  auto talloc = [&](size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) throw Error(kErrAlloc, "vMF scratch");
    tmp.push_back(p);
    return p;
  };
  try {
    int* occ_d[2];
    z64* G2[2];
    for (int sp = 0; sp < 2; ++sp) {
      const SpinDev& d = sp_[sp];
      const int n = d.nel, nc = d.nconf;
      occ_d[sp] = (int*)talloc(sizeof(int) * occ_h[sp].size());
      AFQ_HIP(hipMemcpyAsync(occ_d[sp], occ_h[sp].data(), sizeof(int) * occ_h[sp].size(),
                             hipMemcpyHostToDevice, s));
      z64* A = (z64*)talloc(sizeof(z64) * nc * n * N);
      hipLaunchKernelGGL(k_ph_gather_rows, dim3(ceil_div((long)nc * n * N, 256)), dim3(256), 0, s,
                         nc, n, N, occ_d[sp], d.orbs, A);
      AFQ_LAUNCH_CHECK();
      // TNN = Orbs_u^H Orbs_u = A A^H (herm = false overlaps, density_matrix.hpp:141-145)
      z64* T = (z64*)talloc(sizeof(z64) * nc * n * n);
      {
        Opnd<double, true, true> Aop{&A->re, N, (long)n * N};
        Opnd<double, true, true, true> Bop{&A->re, N, (long)n * N};
        gemm_small<double>(ctx_, "phmf_T", nc, n, n, N, Aop, Bop,
                           EpiStore<double, double>{T, n, (long)n * n, 1, 0, 0});
      }
      z64* lg = (z64*)talloc(sizeof(z64) * nc);
      fill_z64(s, nc, {0, 0}, lg);
      lu_inverse_logdet(s, T, (long)n * n, nc, n, T, (long)n * n, lg, 1.0, {0, 0}, true);
      rat[sp] = (z64*)talloc(sizeof(z64) * nc);
      hipLaunchKernelGGL(k_ph_ratio_from_log, dim3(ceil_div(nc, 256)), dim3(256), 0, s, nc, lg,
                         rat[sp]);
      AFQ_LAUNCH_CHECK();
      // G2[u][i][k] = sum_j Tinv[j][i] conj(orbs[occ_u(j)][k])
      G2[sp] = (z64*)talloc(sizeof(z64) * nc * n * N);
      {
        Opnd<double, true, false> Aop{&T->re, n, (long)n * n};
        Opnd<double, true, false, true> Bop{&A->re, N, (long)n * N};
        gemm_small<double>(ctx_, "phmf_G2", nc, n, N, n, Aop, Bop,
                           EpiStore<double, double>{G2[sp], N, (long)n * N, 1, 0, 0});
      }
      wd[sp] = (z64*)talloc(sizeof(z64) * nc);
    }
    z64* ov = (z64*)talloc(sizeof(z64));
    hipLaunchKernelGGL(k_ph_mf_weights, dim3(1), dim3(256), 0, s, ndet_, det_ab_, det_c_, rat[0],
                       rat[1], sp_[0].nconf, sp_[1].nconf, sp_[0].cptr, sp_[0].ccol, sp_[0].cval,
                       sp_[1].cptr, sp_[1].ccol, sp_[1].cval, wd[0], wd[1], ov);
    AFQ_LAUNCH_CHECK();
    for (int sp = 0; sp < 2; ++sp) {
      const SpinDev& d = sp_[sp];
      const int other = 1 - sp;
      // single excitations between configurations of this spin that share the other spin's
      // configuration (PHMSD.cpp:1443-1517): integer bookkeeping on the host
      std::vector<std::vector<int>> rows(sp_[other].nconf);   // determinant indices per row
      for (int dd = 0; dd < ndet_; ++dd) rows[det_ab_h_[2 * dd + other]].push_back(dd);
      std::vector<int> pairs;
      std::vector<int> c1(d.nel), c2(d.nel);
      for (auto& r : rows)
        for (size_t n1 = 0; n1 < r.size(); ++n1) {
          const int u1 = det_ab_h_[2 * r[n1] + sp];
          std::copy(occ_h[sp].begin() + (size_t)u1 * d.nel, occ_h[sp].begin() + (size_t)(u1 + 1) * d.nel,
                    c1.begin());
          std::sort(c1.begin(), c1.end());
          for (size_t n2 = n1 + 1; n2 < r.size(); ++n2) {
            const int u2 = det_ab_h_[2 * r[n2] + sp];
            std::copy(occ_h[sp].begin() + (size_t)u2 * d.nel,
                      occ_h[sp].begin() + (size_t)(u2 + 1) * d.nel, c2.begin());
            std::sort(c2.begin(), c2.end());
            int np = 0, hole = -1, part = -1;
            for (int i = 0; i < d.nel; ++i)
              if (!std::binary_search(c2.begin(), c2.end(), c1[i])) { ++np; hole = c1[i]; }
            for (int i = 0; i < d.nel; ++i)
              if (!std::binary_search(c1.begin(), c1.end(), c2[i])) part = c2[i];
            if (np != 1) continue;
            pairs.insert(pairs.end(), {hole, part, r[n1], r[n2]});
          }
        }
      int* pairs_d = (int*)talloc(sizeof(int) * std::max<size_t>(pairs.size(), 1));
      if (!pairs.empty())
        AFQ_HIP(hipMemcpyAsync(pairs_d, pairs.data(), sizeof(int) * pairs.size(),
                               hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_ph_mf_density, dim3(ceil_div((long)d.nact * N, 256)), dim3(256), 0, s,
                         d.nconf, d.nel, d.nact, N, occ_d[sp], wd[sp], G2[sp],
                         (int)(pairs.size() / 4), pairs_d, det_c_, d.orbs, ov,
                         Gfull + (sp == 0 ? 0 : (size_t)sp_[0].nact * N));
      AFQ_LAUNCH_CHECK();
      // the pair list must outlive the kernel: synchronize before it goes out of scope
      AFQ_HIP(hipStreamSynchronize(s));
    }
    ham_.vbias(Gfull, 1, dt, v, kFP64);
    zero_imag(s, ham_.M(), v);
    AFQ_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(s);
    for (void* p : tmp) (void)hipFree(p);
    throw;
  }
  for (void* p : tmp) (void)hipFree(p);
  (void)ntot;
}

}  // namespace afq
