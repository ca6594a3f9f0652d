// Fused exchange and Coulomb contraction of the local energy (Real3IndexFactorization::energy,
// energy_impl, Real3IndexFactorization.hpp:476-590) for a real half-rotated tensor in FP32.
//
// For walker w and Cholesky vector m the n x n matrix T^m_ba = sum_k G[w][b][k] L[m][a][k] is
// formed on the matrix cores and reduced on chip to sum_ab T^m_ab T^m_ba (exchange) and
// sum_a T^m_aa (Coulomb); no T reaches memory.  A wavefront owns a tile of two walkers x two
// Cholesky vectors: rows (walker, b) and columns (vector, a), each padded to nsp = 8 NB >= n,
// so the tile is 16 NB x 16 NB and every (walker, vector) block of it lies inside the tile.
//   * The operands are streamed from memory straight into MFMA fragments: lane (fr, fk) of k
//     group g feeds MFMA step j with k = 16 g + 4 fk + j, four consecutive k values per lane.
//   * v_mfma_f32_16x16x4f32 with FP32 accumulation; G complex x L real = 2 MFMAs per block.
//   * The next k group's fragments are loaded while this group's MFMAs run (a full group loads
//     unconditionally from clamped rows, so the compiler's wait lands after the MFMAs).
//   * The pairing T_ab T_ba crosses lanes, so the accumulators go through LDS: the columns of
//     one vector at a time, column-major ([nsp][2 nsp + 8] per wavefront), then each lane adds
//     its share of the products in FP64 over 8 x 8 blocks of (b, a), using the symmetry of the
//     sum to visit only the blocks on and above the diagonal.  Each lane keeps its own FP64
//     sums over all its tiles; a fixed-order wavefront reduction at the end gives one exchange
//     partial per (walker, wavefront).
//   * Coulomb traces sum_a T^m_aa go to Kl[w][m] (added, so both spin blocks of a COLLINEAR
//     walker accumulate), each (w, m) owned by exactly one wavefront.
// A workgroup (four wavefronts) takes one walker pair and a run of 8 VPW Cholesky vectors;
// workgroups are renumbered so that each XCD processes a contiguous run with the walker pair
// fastest, so the workgroups resident on one XCD share their Cholesky vectors through its L2.
#include "gemm.hpp"
#include "ops.hpp"

namespace afq {

namespace {

using gemm::f32x4;

constexpr int kXWaves = 4;

template <int NB>
struct XFrag {
  float gr[NB][4], gi[NB][4];  // G rows of the tile's NB row blocks, four k values each
  float l[NB][4];              // L rows (columns of the tile), four k values each
};

// Byte offsets of a lane's G and L rows (at its first k value) from the operands' bases.
// Rows and columns past the walkers, vectors or orbitals point at a clamped, valid row: they only
// feed tile rows and columns that are never read, and finite data keeps them harmless.  The
// offsets are 32-bit (the host splits G by walkers and checks L's size), so every load is a
// uniform base address plus one offset register.
template <int NB>
struct XOfs {
  uint32_t g[NB], l[NB];
};

template <int NB>
__device__ inline XOfs<NB> x_offsets(int W, int ns, int N, int w0, int m0, int M, int fr, int fk) {
  constexpr int NSP = 8 * NB;
  XOfs<NB> o;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int r = 16 * i + fr, b = min(r % NSP, ns - 1), w = min(w0 + r / NSP, W - 1);
    o.g[i] = (uint32_t)((((long)w * ns + b) * N + 4 * fk) * (long)sizeof(c32));
  }
#pragma unroll
  for (int jb = 0; jb < NB; ++jb) {
    const int c = 16 * jb + fr, a = min(c % NSP, ns - 1), m = min(m0 + c / NSP, M - 1);
    o.l[jb] = (uint32_t)((((long)m * ns + a) * N + 4 * fk) * (long)sizeof(float));
  }
  return o;
}

// The fragments of one k group: gb and lb are the operands' bases advanced to the group's first
// k value.  VEC: N % 4 == 0 and 16-byte aligned operands (vector loads).  FULL groups load
// unconditionally, so the compiler counts the loads in flight and the next group's loads overlap
// this group's MFMAs; the ragged last group passes rem = N - k (the lane's k values left) and
// zero-fills k >= N.
template <int NB, bool VEC, bool FULL>
__device__ inline void x_load(XFrag<NB>& f, const char* gb, const char* lb, const XOfs<NB>& o,
                              int rem) {
  const float4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const char* p = gb + o.g[i];
    float4 x = z, y = z;
    if constexpr (VEC) {
      if (FULL || rem > 0) {
        x = *reinterpret_cast<const float4*>(p);
        y = *reinterpret_cast<const float4*>(p + 16);
      }
    } else {
      float2 t[4] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (FULL || rem > j) t[j] = *reinterpret_cast<const float2*>(p + 8 * j);
      x = float4{t[0].x, t[0].y, t[1].x, t[1].y};
      y = float4{t[2].x, t[2].y, t[3].x, t[3].y};
    }
    f.gr[i][0] = x.x; f.gi[i][0] = x.y; f.gr[i][1] = x.z; f.gi[i][1] = x.w;
    f.gr[i][2] = y.x; f.gi[i][2] = y.y; f.gr[i][3] = y.z; f.gi[i][3] = y.w;
  }
#pragma unroll
  for (int jb = 0; jb < NB; ++jb) {
    const char* p = lb + o.l[jb];
    float4 x = z;
    if constexpr (VEC) {
      if (FULL || rem > 0) x = *reinterpret_cast<const float4*>(p);
    } else {
      float t[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (FULL || rem > j) t[j] = *reinterpret_cast<const float*>(p + 4 * j);
      x = float4{t[0], t[1], t[2], t[3]};
    }
    f.l[jb][0] = x.x; f.l[jb][1] = x.y; f.l[jb][2] = x.z; f.l[jb][3] = x.w;
  }
}

template <int NB>
__device__ inline void x_mma(const XFrag<NB>& f, f32x4 (&ar)[NB][NB], f32x4 (&ai)[NB][NB]) {
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int jb = 0; jb < NB; ++jb) {
        ar[i][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(f.gr[i][j], f.l[jb][j], ar[i][jb], 0, 0, 0);
        ai[i][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(f.gi[i][j], f.l[jb][j], ai[i][jb], 0, 0, 0);
      }
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int NB, bool VEC>
__global__ void __launch_bounds__(64 * kXWaves)
    k_exx_fused(int W, int ns, int N, int M, const c32* __restrict__ G,
                const float* __restrict__ L, int vpw, int npairs, int nchunk, z64* __restrict__ part,
                int nparts, z64* __restrict__ Kl) {
  // LDS pitch of a column of T: the rows of both walkers plus 8, which keeps the 8 x 8 lane
  // blocks of the exchange reads on distinct banks
  constexpr int NSP = 8 * NB, R = 2 * NSP, LDC = R + 8;
  extern __shared__ __align__(16) float smem_x[];  // [kXWaves][2][NSP * LDC]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int fr = lane & 15, fk = lane >> 4;
  const int xb = lane & 7, xa = lane >> 3;
  const long total = gridDim.x, xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const long q = xcd * (total / 8) + min(xcd, total % 8) + slot;
  const int pair = (int)(q % npairs), chunk = (int)(q / npairs);
  const int w0 = 2 * pair;
  const int NGF = N / 16;  // full k groups
  float* tre = smem_x + wave * 2 * NSP * LDC;
  float* tim = tre + NSP * LDC;
  double xr[2] = {0.0, 0.0}, xi[2] = {0.0, 0.0};

  for (int vp = 0; vp < vpw; ++vp) {
    const int m0 = (chunk * kXWaves * vpw + wave * vpw + vp) * 2;
    if (m0 >= M) break;  // uniform per wavefront; no workgroup barrier below
    f32x4 ar[NB][NB], ai[NB][NB];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int jb = 0; jb < NB; ++jb) {
        ar[i][jb] = f32x4{0.f, 0.f, 0.f, 0.f};
        ai[i][jb] = f32x4{0.f, 0.f, 0.f, 0.f};
      }
    const XOfs<NB> o = x_offsets<NB>(W, ns, N, w0, m0, M, fr, fk);
    const char* gb = reinterpret_cast<const char*>(G);
    const char* lb = reinterpret_cast<const char*>(L);
    constexpr long GS = 16 * sizeof(c32), LS = 16 * sizeof(float);  // bytes per k group
    XFrag<NB> cur, nxt;
    if (NGF > 0) {
      x_load<NB, VEC, true>(cur, gb, lb, o, 0);
      for (int g = 0; g < NGF; ++g) {
        // the last full group reloads itself rather than branching around the prefetch
        const int gn = min(g + 1, NGF - 1);
        x_load<NB, VEC, true>(nxt, gb + gn * GS, lb + gn * LS, o, 0);
        // the prefetch stays ahead of the MFMAs and the copy behind them, so the loads have a
        // whole group of MFMAs to land
        __builtin_amdgcn_sched_barrier(0);
        x_mma<NB>(cur, ar, ai);
        __builtin_amdgcn_sched_barrier(0);
        cur = nxt;
      }
    }
    if (NGF * 16 < N) {
      x_load<NB, VEC, false>(cur, gb + NGF * GS, lb + NGF * LS, o, N - 16 * NGF - 4 * fk);
      x_mma<NB>(cur, ar, ai);
    }
    // one Cholesky vector (tile columns mj NSP .. mj NSP + NSP - 1) at a time through LDS,
    // column-major: T[row][c] at c LDC + row, so each lane stores its four accumulator rows as
    // one 16-byte write
#pragma unroll
    for (int mj = 0; mj < 2; ++mj) {
      // the previous pass's reads of this wavefront's LDS are complete before it is rewritten
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int jb = 0; jb < NB; ++jb) {
        const int c = 16 * jb + fr - mj * NSP;
        if (c < 0 || c >= NSP) continue;
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          *reinterpret_cast<f32x4*>(tre + c * LDC + 16 * i + 4 * fk) = ar[i][jb];
          *reinterpret_cast<f32x4*>(tim + c * LDC + 16 * i + 4 * fk) = ai[i][jb];
        }
      }
      // the stores are complete (LDS operations of a wavefront complete in order) before any
      // lane reads another lane's entries
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (m0 + mj >= M) continue;
#pragma unroll
      for (int wi = 0; wi < 2; ++wi) {
        if (w0 + wi >= W) continue;
        const float* ur = tre + wi * NSP;
        const float* ui = tim + wi * NSP;
        // exchange: sum_ab T_ab T_ba over 8 x 8 blocks (ib, ia) of (b, a), lane (lb, la) taking
        // b = 8 ib + lb, a = 8 ia + la; the sum is symmetric in (a, b), so the blocks above the
        // diagonal count twice and those below are skipped
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
          for (int ia = ib; ia < NB; ++ia) {
            const int b = 8 * ib + xb, a = 8 * ia + xa;
            const double pr = ur[a * LDC + b], pi = ui[a * LDC + b];  // T_ba
            const double qr = ur[b * LDC + a], qi = ui[b * LDC + a];  // T_ab
            const double wgt = (b < ns && a < ns) ? (ia == ib ? 1.0 : 2.0) : 0.0;
            xr[wi] += wgt * (pr * qr - pi * qi);
            xi[wi] += wgt * (pr * qi + pi * qr);
          }
        // Coulomb trace of this (walker, vector)
        double kr = 0.0, ki = 0.0;
        if (lane < ns) {
          kr = ur[lane * LDC + lane];
          ki = ui[lane * LDC + lane];
        }
        kr = wave_sum(kr);
        ki = wave_sum(ki);
        if (Kl && lane == 0) {
          z64* o = Kl + (long)(w0 + wi) * M + m0 + mj;
          *o = z64{o->re + kr, o->im + ki};
        }
      }
    }
  }
#pragma unroll
  for (int wi = 0; wi < 2; ++wi) {
    const double er = wave_sum(xr[wi]), ei = wave_sum(xi[wi]);
    if (lane == 0 && w0 + wi < W)
      part[(long)(w0 + wi) * nparts + chunk * kXWaves + wave] = z64{er, ei};
  }
}

template <int NB>
void launch_exx(hipStream_t s, int W, int ns, int N, int M, const c32* G, const float* L,
                z64* part, int nparts, z64* Kl) {
  const int vpw = exx_fused_vpw();
  const int npairs = ceil_div(W, 2);
  const int nchunk = ceil_div(M, 2 * kXWaves * vpw);
  AFQ_CHECK(nparts == nchunk * kXWaves, kErrArg, "exchange partial count");
  const long blocks = (long)npairs * nchunk;
  AFQ_CHECK(blocks < (1L << 31), kErrShape, "grid too large");
  constexpr int NSP = 8 * NB;
  const size_t lds = sizeof(float) * kXWaves * 2 * NSP * (2 * NSP + 8);
  const bool vec = (N % 4) == 0 && (reinterpret_cast<uintptr_t>(G) % 16) == 0 &&
                   (reinterpret_cast<uintptr_t>(L) % 16) == 0;
  auto kern = vec ? k_exx_fused<NB, true> : k_exx_fused<NB, false>;
  if (lds > 64 * 1024)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(64 * kXWaves), lds, s, W, ns, N, M, G, L,
                     vpw, npairs, nchunk, part, nparts, Kl);
  AFQ_LAUNCH_CHECK();
}

}  // namespace

int exx_fused_vpw() { return 4; }

int exx_fused_parts(int M) { return ceil_div(M, 2 * kXWaves * exx_fused_vpw()) * kXWaves; }

// The kernel addresses L with 32-bit byte offsets.
bool exx_fused_supported(int ns, int N, int M) {
  return ns >= 1 && ns <= 40 && N >= 1 && M >= 1 &&
         (long)M * ns * N * (long)sizeof(float) < (1L << 32);
}

void exchange_fused(hipStream_t s, int W, int ns, int N, int M, const c32* G, const float* L,
                    z64* part, int nparts, z64* Kl) {
  AFQ_CHECK(exx_fused_supported(ns, N, M), kErrShape,
            "fused exchange supports 1 <= n <= 40 and L under 4 GiB");
  // walkers in runs (of even length) whose G stays under 4 GiB for the 32-bit offsets
  const long per_w = (long)ns * N * (long)sizeof(c32);
  const int wrun = (int)std::min<long>(W, ((1L << 32) / per_w) & ~1L);
  AFQ_CHECK(wrun >= 1, kErrShape, "fused exchange: one walker's G exceeds 2 GiB");
  for (int w = 0; w < W; w += wrun) {
    const int wc = std::min(wrun, W - w);
    const c32* Gw = G + (long)w * ns * N;
    z64* pw = part + (long)w * nparts;
    z64* kw = Kl ? Kl + (long)w * M : nullptr;
    switch ((ns + 7) / 8) {
      case 1: launch_exx<1>(s, wc, ns, N, M, Gw, L, pw, nparts, kw); break;
      case 2: launch_exx<2>(s, wc, ns, N, M, Gw, L, pw, nparts, kw); break;
      case 3: launch_exx<3>(s, wc, ns, N, M, Gw, L, pw, nparts, kw); break;
      case 4: launch_exx<4>(s, wc, ns, N, M, Gw, L, pw, nparts, kw); break;
      default: launch_exx<5>(s, wc, ns, N, M, Gw, L, pw, nparts, kw); break;
    }
  }
}

}  // namespace afq
