// Internal launcher declarations shared by the kernel translation units and the C-ABI layer.
#pragma once

#include "afq_common.hpp"

namespace afq {

// ---- dense_small.hip -------------------------------------------------------------------------
// Factors kept between calls on the same walkers: the matrices T [nb][n][n] they belong to (NaN
// until set), their inverses [nb][n][n] and raw log dets [nb].
struct LuKeep {
  z64* T;
  z64* Tinv;
  z64* ldet;
};
// n <= 48.  keep, !reuse: A is keep->T; the inverse goes to keep->Tinv and the raw log det to keep->ldet.
// keep, reuse: matrices equal bit for bit to keep->T take the kept log det (nothing recomputed,
// nothing written but ovlp); the others are copied into keep->T and factorized into the kept
// buffers.  Either way ovlp is updated as without keep and Ainv must be keep->Tinv.
void lu_inverse_logdet(hipStream_t s, const z64* A, long a_bs, int nb, int n, z64* Ainv, long i_bs,
                       z64* ovlp, double ov_scale, z64 ov_add, bool want_inverse,
                       const LuKeep* keep = nullptr, bool reuse = false);
void chol_upper_inverse(hipStream_t s, const z64* S, int nb, int n, z64* Rinv, z64* ldet);

// ---- fields.hip ------------------------------------------------------------------------------
struct FieldParams {
  double bound;          // vbias_bound
  int project_vbias;     // use_cp_constraint || project_force_bias
  int free_projection;
  uint64_t seed;         // Philox key (used when u == nullptr)
  uint32_t stream;
  uint32_t step;
  const uint32_t* step_ptr;  // when set, the step index is read from device memory
  // bound-trigger counters (may be null): [0] += W M (walker, field) pairs, [1] += pairs with
  // |vbias| > bound (AFQMCBasePropagator.cpp:697-706)
  unsigned long long* stats = nullptr;
};
// X holds vbias on entry (complex [W][M]), fields on exit; HW, mf [W]. u == nullptr draws the
// uniforms in-kernel from Philox4x32-10 keyed by (seed, stream, step, walker, field). X32, when
// given, receives a complex64 copy of the fields.
void construct_x(hipStream_t s, const double* u, const z64* vMF, int W, int M,
                 const FieldParams& p, z64* X, z64* HW, z64* mf, c32* X32);
void field_uniforms(hipStream_t s, uint64_t seed, uint32_t stream, uint32_t step, int W, int M,
                    double* u);

struct UpdateParams {
  double dt, E0, lower_cut, upper_cut, weight_floor, weight_frac, target_pop;
  int free_projection, use_cp;
  int local_energy;  // 1: E_loc = sum of the energies E [W][3] (E0 included), midpoint rule
  // bound-trigger counters (may be null): [3..5] local-energy clamp total / upper / lower
  // (bound_local_energy, WalkerSetUpdate.hpp:46-58), [6..8] weight clamp total / upper / lower
  // (bound_walker_weights, :199-218)
  unsigned long long* stats = nullptr;
};
// Eshift is read from device memory (so a captured step needs no host value).  When wfac is
// given, the step's path-restoration factor (w / |w|) / scale of every walker (0 for a dropped
// walker; WalkerSetUpdate.hpp:136-146, 184-187) is stored in the ring wfac [W][hl] at slot
// *step_ptr % hl.
void walker_update(hipStream_t s, int W, const UpdateParams& p, const double* Eshift,
                   const z64* new_ovlp, const z64* mf, const z64* HW, z64* weight, z64* phase,
                   z64* eloc, z64* ovlp, z64* theta, z64* wfac = nullptr, int hl = 0,
                   const uint32_t* step_ptr = nullptr, const z64* E = nullptr);
// ring[w][*step_ptr % nring][m] = X[w][m] (WalkerSetBase::storeFields)
void store_fields(hipStream_t s, int W, int M, const z64* X, z64* ring, int nring,
                  const uint32_t* step_ptr);
// X[w][m] = ring[w][slot][m]
void gather_fields(hipStream_t s, int W, int M, const z64* ring, int nring, int slot, z64* X);
// weights[w] = weight[w] * prod_{k < n} wfac[w][(pos - 1 - k) mod hl]  (BackPropEstimator.hpp
// path restoration); n = 0 copies the weights
void restore_weights(hipStream_t s, int W, const z64* weight, const z64* wfac, int hl, uint32_t pos,
                     int n, z64* out);
// per-walker rows that move with their walker in population control (back-propagation state,
// the energy components): n[r] elements at p[r] + w * bs[r] (bs 0: n[r])
struct WalkerRows {
  static constexpr int kMax = 10;
  z64* p[kMax] = {};
  long n[kMax] = {};
  long bs[kMax] = {};
};
// e1[w bs], exx[w bs], ej[w bs] = E[w][0..2] (the walker's E1_, EXX_, EJ_ columns,
// WalkerSetUpdate.hpp:180-184)
void store_energy_components(hipStream_t s, int W, const z64* E, z64* e1, z64* exx, z64* ej,
                             long bs);
void increment_counter(hipStream_t s, uint32_t* c);
// comb population control of one rank (population_control.hpp:216-257); r_dev = injected
// uniform or nullptr for Philox(seed, stream, *step_ptr).  cum [W] doubles, src [W] ints and
// status [1] int are device scratch.  The Slater blocks of walker w are rowA (rowB) elements at
// PhiA + w bsA (PhiB + w bsB).
void population_control(hipStream_t s, int W, const double* r_dev, uint64_t seed, uint32_t stream,
                        const uint32_t* step_ptr, long rowA, long rowB, long bsA, long bsB,
                        z64* PhiA, z64* PhiB, z64* weight, z64* phase, z64* eloc, z64* ovlp,
                        z64* theta, double* cum, int* src, int* status,
                        const WalkerRows& extra = WalkerRows{});
// growth estimator log(sum w e^{dt E} / sum w)/dt -> out[0]; Eshift[0] += f (out - Eshift) when
// relax > 0.
void average_pseudo_energy(hipStream_t s, int W, const z64* weight, const z64* eloc, double dt,
                           double relax, double* out, double* Eshift);

// ---- estimators.hip (rdm_measure, which needs the context's scratch, is declared in ops.hpp) ---
// Reference coefficients of a multi-reference measurement (MixedEstimator.hpp:61-103,
// BackPropEstimator.hpp:62-120): x[d][w] = cc[d] exp(scale (lo[d][w] + conj(ad[w][d]))) (ad
// may be null), coeff[d][w] = x[d][w] / sum_d' x[d'][w].  lo [R][W] log overlaps, cc [R]
// conj(ci); scale 2 doubles the alpha block of CLOSED walkers where the estimator does.
void reference_coefficients(hipStream_t s, int W, int R, const z64* lo, const z64* cc,
                            const z64* ad, z64* coeff, double scale = 1.0);

// ---- elementwise.hip -------------------------------------------------------------------------
void scale_add_cplx(hipStream_t s, long n, const z64* x, double ar, double ai, z64* y);  // y = a x
void cast_z64_c32(hipStream_t s, long n, const z64* x, c32* y);
void cast_c32_z64(hipStream_t s, long n, const c32* x, z64* y);
void cast_f64_f32(hipStream_t s, long n, const double* x, float* y);
void fill_z64(hipStream_t s, long n, z64 v, z64* y);
void fill_blocks_z64(hipStream_t s, long B, long n, long stride, z64 v, z64* y);
// out[b][i][j] = in[b][j][i] for B square N x N blocks (out != in)
void transpose_blocks(hipStream_t s, long B, int N, const z64* in, z64* out);
void zero_imag(hipStream_t s, long n, z64* y);
void truncate_small(hipStream_t s, long n, double cut, z64* y);
void hermitize_scale(hipStream_t s, int N, double scale, z64* A);
void add_one_body(hipStream_t s, int N, double dt, const double* h, const double* vexx,
                  const z64* shift, z64* H1);
void identity(hipStream_t s, int N, z64* A);
void axpy_z64(hipStream_t s, long n, double ar, double ai, const z64* x, z64* y);  // y += a x
void permute_akm_mak(hipStream_t s, int A, int K, int M, const z64* in, z64* out);
void conj_transpose(hipStream_t s, int R, int C, const z64* in, z64* out);
void broadcast_z64(hipStream_t s, long W, long n, const z64* src, z64* dst);  // dst[w][:] = src
// dst[r][:] = src[idx[r]][:] for R rows of C elements (idx on the device)
void gather_rows(hipStream_t s, long R, int C, const int* idx, const z64* src, z64* dst);
// K strided columns of W walkers (element w of column k at p[k] + w * stride) copied to or from
// packed [K][W]
struct ColumnSet {
  static constexpr int kMax = 8;
  z64* p[kMax] = {};
};
void copy_columns(hipStream_t s, int K, int W, long stride, const ColumnSet& cols, z64* packed,
                  bool to_packed);
// out[b][j][i][:] = in[b][i][j][:] with in [B][D0][D1][K] and out [B][D1][D0][K]
void swap_mid(hipStream_t s, long B, int D0, int D1, int K, const z64* in, z64* out);
// true when every Cholesky vector L[.][.][m] of an [N][N][M] tensor is exactly symmetric
bool symmetric_cholesky(hipStream_t s, int N, int M, const double* L);
void real_part(hipStream_t s, long n, const z64* x, double* y);
void real_part_f32(hipStream_t s, long n, const z64* x, float* y);
// out[w*stride] += c * sum_k G[w][k] H[k] + add
void dot_rows(hipStream_t s, int W, long K, const z64* G, const z64* H, double cr, double ci,
              double add, z64* out, long out_stride);
void exchange_reduce(hipStream_t s, int W, int n, int MC, int m0, int M, const z64* T,
                     z64* exx_part, int nchunk, int chunk_id, z64* Kl);
// Cholesky vectors per workgroup of exchange_blocks for n electrons
int exchange_group(int n);
// exchange / Coulomb partial sums of a chunk of mcount Cholesky vectors starting at m0, from
// T [W][n][MC][n] (complex TT): exx_part[w][g0 + g] for the chunk's groups g of mg vectors
// (ng_total partials per walker; both spin blocks of a walker use the same mg, the group size
// of the larger block, so their partials line up), Kl[w][m0 + m] += tr T^m (Kl may be null)
template <typename TT>
void exchange_blocks(hipStream_t s, int W, int n, int mg, int MC, int m0, int mcount, int M,
                     const cplx<TT>* T, z64* exx_part, int ng_total, int g0, z64* Kl);
void energy_finish(hipStream_t s, int W, int M, int nchunk, const z64* exx_part, const z64* Kl,
                   double cx, double cj, int addEXX, int addEJ, z64* E);
// Fused FP32 exchange of a real half-rotated tensor (exchange.hip): T^m = G_w L^m^T formed and
// reduced on chip.  G [W][ns][N] complex64, L [M][ns][N] real; part [W][nparts] exchange
// partials (nparts = exx_fused_parts(M)), Kl[w][m] += tr T^m (Kl may be null).
bool exx_fused_supported(int ns, int N, int M);
int exx_fused_vpw();
int exx_fused_parts(int M);
void exchange_fused(hipStream_t s, int W, int ns, int N, int M, const c32* G, const float* L,
                    z64* part, int nparts, z64* Kl);

}  // namespace afq
