// phmsd_oracle.cpp -- FP64 CPU restatement of SAFIRE's particle-hole multi-determinant (PHMSD)
// trial operations for COLLINEAR walkers.
//
// THIS IS TEST INFRASTRUCTURE, like afq_oracle.cpp: only the test suite, the smoke check and the
// benchmark's cpu_baseline leg load it; the product never does.
//
// Restated (host branches, paths relative to the SAFIRE tree):
//   det_ops::MixedDensityMatrixForWoodbury  density_matrix.hpp:651-717 (Log_OverlapForWoodbury
//                                           :487-532 is its first half)
//   calculate_overlaps                      phmsd_helpers.hpp:112-198
//   calculate_R                             phmsd_helpers.hpp:201-290
//   get_compact_ph_R_matrices               phmsd_helpers.hpp:312-390
//   PHMSD::MixedDensityMatrix (compact)     PHMSD.cpp:549-790 (COLLINEAR branch)
//   PHMSD::Log_Overlap                      PHMSD.cpp:804-900
//   PHMSD::energy_alg1                      PHMSD.cpp:376-544
//   ph_reference_energy / ph_ref_energy_impl Real3IndexFactorization.hpp:212-233, 598-696
//   ph_excited_energy                       Real3IndexFactorization.hpp:235-273
//   ph_excited_1body_energy                 detail/phmsd_impl.hpp:152-190
//   ph_excited_2body_energy_dense_cholesky  detail/phmsd_impl.hpp:33-150
//   PHMSD::vMF (trial density, COLLINEAR)   PHMSD.cpp:1329-1573
// The small excitation determinants and inverses, which the reference writes out as cofactor
// expansions for ranks <= 5 (D2x2..D5x5, I2x2, I3x3), are computed here by LU; the two agree to
// rounding.
//
// orc_msd_brute is independent of all of the above: it evaluates the same trial as an explicit
// sum over its determinants (sorted occupation strings, unsigned coefficients), each one a
// single-determinant overlap, Green's function and local energy. It pins the Woodbury/Wick
// algebra, the excitation ordering and the permutation signs.
//
// Layouts (row-major, complex interleaved):
//   orbs[s]   [nact_s][N]       trial orbitals of spin s (PsiT_MO rows, already conjugated)
//   Phi_a/b   [W][N][nup|ndn]   walkers
//   G         [W][nact_a + nact_b][N]  compact mixed density matrix, alpha rows then beta rows
//   haj       [nact_a + nact_b][N]     half-rotated h1 (rows = orbs)
//   Lnak      [M][nact_a][N] followed by [M][nact_b][N]
//   E         [W][3] = (E1 + E0, EXX, EJ)
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "orc_dense.hpp"

using namespace orc;

extern "C" {

// The PHMSD trial in the reference's ph_excitations layout (Excitations.hpp:203-530):
// per spin, unique configurations ordered by excitation rank (index 0 = reference); a
// configuration of rank r >= 1 is 2r integers at exc[s][exc_start[s][u]]: r hole positions
// (indices into the reference configuration) then r particle orbitals (active indices).
// Determinants couple an alpha and a beta configuration with coefficient ci (permutation sign
// included, as build_ph_struct stores it).
struct orc_ph {
  int N;
  int nel[2];
  int nact[2];
  const double* orbs[2];
  int nconf[2];
  const int* rank[2];
  const int* exc_start[2];
  const int* exc[2];
  const int* refc[2];
  int ndet;
  const int* det_a;
  const int* det_b;
  const double* ci;
};

}  // extern "C"

namespace {

struct Spin {
  int N, nel, nact, nconf;
  const cd* orbs;
  const int* rank;
  const int* exc_start;
  const int* exc;
  const int* refc;
  const int* e(int u) const { return exc + exc_start[u]; }
};

Spin spin(const orc_ph* P, int s) {
  return Spin{P->N, P->nel[s], P->nact[s], P->nconf[s], (const cd*)P->orbs[s], P->rank[s],
              P->exc_start[s], P->exc[s], P->refc[s]};
}

// orbs(i) of configuration u: refc with the holes replaced in place by the particles
// (ph_excitations::get_configuration, Excitations.hpp:440-481).
void configuration(const Spin& S, int u, int* orbs) {
  for (int i = 0; i < S.nel; ++i) orbs[i] = S.refc[i];
  int r = S.rank[u];
  if (r == 0) return;
  const int* e = S.e(u);
  for (int p = 0; p < r; ++p) orbs[e[p]] = e[p + r];
}

// MixedDensityMatrixForWoodbury (density_matrix.hpp:651-717), one walker:
//   TAB = orbs * Phi [nact][nel]; TNN = rows refc of TAB; log_ov += log det TNN;
//   TNN <- TNN^{-1} (zeros if singular); QQ0 = TAB * TNN^{-1}; GA = TNN^{-T} Phi^T [nel][N].
void woodbury(const Spin& S, const cd* Phi, cd* GA, cd* QQ0, cd& log_ov) {
  const int N = S.N, n = S.nel, na = S.nact;
  std::vector<cd> TAB((size_t)na * n), TNN((size_t)n * n);
  zgemm('N', 'N', na, n, N, 1.0, S.orbs, N, Phi, n, 0.0, TAB.data(), n);
  for (int i = 0; i < n; ++i)
    std::copy(TAB.begin() + (size_t)S.refc[i] * n, TAB.begin() + (size_t)(S.refc[i] + 1) * n,
              TNN.begin() + (size_t)i * n);
  std::vector<int> piv(n);
  int info = lu_factor(TNN.data(), n, piv.data());
  logdet_from_getrf(TNN.data(), piv.data(), n, log_ov);
  inverse_from_lu(TNN.data(), piv.data(), n, info);
  zgemm('N', 'N', na, n, n, 1.0, TAB.data(), n, TNN.data(), n, 0.0, QQ0, n);
  if (GA) zgemm('T', 'T', n, N, n, 1.0, TNN.data(), n, Phi, n, 0.0, GA, N);
}

// determinant and inverse of the r x r excitation block Q[p][q] = T[e[p+r]][e[q]]
// (calculate_overlaps / calculate_R); inverse left in Q when invert (zeros if singular).
cd excitation_block(const cd* T, int nel, const int* e, int r, cd* Q, bool invert) {
  for (int p = 0; p < r; ++p)
    for (int q = 0; q < r; ++q) Q[p * r + q] = T[(size_t)e[p + r] * nel + e[q]];
  std::vector<int> piv(r);
  std::vector<cd> lu(Q, Q + r * r);
  int info = lu_factor(lu.data(), r, piv.data());
  cd det = 1.0;
  for (int i = 0; i < r; ++i) det *= (piv[i] == i + 1) ? lu[i * r + i] : -lu[i * r + i];
  if (invert) {
    inverse_from_lu(lu.data(), piv.data(), r, info);
    std::copy(lu.begin(), lu.end(), Q);
  }
  return det;
}

// calculate_overlaps (phmsd_helpers.hpp:112-198): ov[0] = 1, ov[u] = det of u's block of QQ0.
void overlaps(const Spin& S, const cd* QQ0, cd* ov) {
  ov[0] = 1.0;
  std::vector<cd> Q(64);
  for (int u = 1; u < S.nconf; ++u) {
    int r = S.rank[u];
    Q.resize((size_t)r * r);
    ov[u] = excitation_block(QQ0, S.nel, S.e(u), r, Q.data(), false);
  }
}

// get_compact_ph_R_matrices host branch (phmsd_helpers.hpp:312-390) for configuration u:
// R[p][a] (r x nact), zero when the block is singular.
void compact_R(const Spin& S, const cd* T, int u, cd* R) {
  const int n = S.nel, na = S.nact, r = S.rank[u];
  const int* e = S.e(u);
  std::vector<int> orbs(n);
  configuration(S, u, orbs.data());
  std::fill(R, R + (size_t)r * na, cd(0.0));
  std::vector<cd> Q((size_t)r * r);
  cd ov_a = excitation_block(T, n, e, r, Q.data(), true);
  if (std::abs(ov_a) == 0.0) return;
  for (int p = 0; p < r; ++p)
    for (int q = 0; q < r; ++q) {
      cd Ipq = Q[p * r + q];
      const cd* Tq = T + (size_t)e[q + r] * n;
      for (int i = 0; i < n; ++i) R[(size_t)p * na + orbs[i]] -= Ipq * Tq[i];
      R[(size_t)p * na + orbs[e[q]]] += Ipq;
    }
}

// calculate_R host branch (phmsd_helpers.hpp:201-258): R[i][a] (nel x nact).
void calculate_R(const Spin& S, const cd* T, const cd* wgt, cd* R) {
  const int n = S.nel, na = S.nact;
  std::fill(R, R + (size_t)n * na, cd(0.0));
  for (int i = 0; i < n; ++i) R[(size_t)i * na + S.refc[i]] += wgt[0];
  std::vector<int> orbs(n);
  for (int u = 1; u < S.nconf; ++u) {
    int r = S.rank[u];
    const int* e = S.e(u);
    configuration(S, u, orbs.data());
    std::vector<cd> Q((size_t)r * r);
    cd ov_a = excitation_block(T, n, e, r, Q.data(), true);
    cd w = wgt[u];
    if (std::abs(ov_a) == 0.0) continue;
    for (int i = 0; i < n; ++i) R[(size_t)i * na + orbs[i]] += w;
    for (int p = 0; p < r; ++p)
      for (int q = 0; q < r; ++q) {
        cd Ipq = Q[p * r + q];
        const cd* Tq = T + (size_t)e[q + r] * n;
        for (int i = 0; i < n; ++i) R[(size_t)e[p] * na + orbs[i]] -= w * Ipq * Tq[i];
        R[(size_t)e[p] * na + orbs[e[q]]] += w * Ipq;
      }
  }
}

// spin couplings (Wavefunction.cpp:425-451): C[0] rows alpha configurations, C[1] rows beta;
// values conj(ci). Kept as per-row lists in determinant order.
struct Couplings {
  std::vector<std::vector<std::pair<int, cd>>> row[2];
  explicit Couplings(const orc_ph* P) {
    row[0].resize(P->nconf[0]);
    row[1].resize(P->nconf[1]);
    const cd* ci = (const cd*)P->ci;
    for (int d = 0; d < P->ndet; ++d) {
      row[0][P->det_a[d]].push_back({P->det_b[d], std::conj(ci[d])});
      row[1][P->det_b[d]].push_back({P->det_a[d], std::conj(ci[d])});
    }
  }
  // out[u] = sum_v C[s][u][v] x[v]  (csrmm<'N'>)
  void apply(int s, const cd* x, int nrows, cd* out) const {
    for (int u = 0; u < nrows; ++u) {
      cd acc = 0.0;
      for (auto& kv : row[s][u]) acc += kv.second * x[kv.first];
      out[u] = acc;
    }
  }
};

// ph_ref_energy_impl host branch (Real3IndexFactorization.hpp:598-696) for one walker and spin:
// Swia [nel][nact], Twina [nel][M][nact]; E[0] += sum_i Swia[i][i]; E[1] exchange; Kl += diag.
void ph_ref_energy(int nel, int nact, int N, int M, const cd* GA, const cd* haj, const cd* L,
                   cd* E, cd* Kl, cd* Swia, cd* Twina) {
  zgemm('N', 'T', nel, nact, N, 1.0, GA, N, haj, N, 0.0, Swia, nact);
  for (int i = 0; i < nel; ++i) E[0] += Swia[(size_t)i * nact + i];
  // Twina[i][n][a] = sum_k GA[i][k] L[n][a][k]
  zgemm('N', 'T', nel, M * nact, N, 1.0, GA, N, L, N, 0.0, Twina, M * nact);
  auto T = [&](int i, int n, int a) { return Twina[((size_t)i * M + n) * nact + a]; };
  for (int ia = 0; ia < nel; ++ia) {
    cd s = 0.0;
    for (int n = 0; n < M; ++n) s += T(ia, n, ia) * T(ia, n, ia);
    E[1] += -0.5 * s;
    for (int ib = ia + 1; ib < nel; ++ib) {
      cd t = 0.0;
      for (int n = 0; n < M; ++n) t += T(ia, n, ib) * T(ib, n, ia);
      E[1] += -1.0 * t;
    }
  }
  for (int ia = 0; ia < nel; ++ia)
    for (int n = 0; n < M; ++n) Kl[n] += T(ia, n, ia);
}

}  // namespace

extern "C" {

// ph_excited_1body_energy host branch (detail/phmsd_impl.hpp:152-190).
// S [nwalk][nelec][nact], R [nwalk][ndet][nex][nact], wgt [ndet][nwalk], iexcit [ndet][2 nex].
void orc_ph_excited_1body(const int* iexcit, const int* refc, const double* S_, const double* R_,
                          const double* wgt_, int nwalk, int ndet, int nex, int nact, int nelec,
                          double* E_) {
  const cd* S = (const cd*)S_;
  const cd* R = (const cd*)R_;
  const cd* wgt = (const cd*)wgt_;
  cd* E = (cd*)E_;
  for (int iw = 0; iw < nwalk; ++iw)
    for (int d = 0; d < ndet; ++d) {
      cd e_ = 0.0;
      const int* ex = iexcit + (size_t)d * 2 * nex;
      for (int p = 0; p < nex; ++p) {
        int ip = ex[p];
        const cd* Si = S + ((size_t)iw * nelec + ip) * nact;
        const cd* Rp = R + (((size_t)iw * ndet + d) * nex + p) * nact;
        cd s = 0.0;
        for (int a = 0; a < nact; ++a) s += Si[a] * Rp[a];
        e_ += s + Si[ex[nex + p]] - Si[refc[ip]];
      }
      E[iw] += wgt[(size_t)d * nwalk + iw] * e_;
    }
}

// ph_excited_2body_energy_dense_cholesky host branch (detail/phmsd_impl.hpp:33-150).
// Twina [nwalk][nelec][nchol][nact], R [nwalk][ndet][nex][nact], wgt [ndet][nwalk],
// EX, EJ [nwalk] accumulated, KE [ndet][nwalk][nchol] written.
void orc_ph_excited_2body(const int* iexcit, const int* refc, const double* Tw_, const double* R_,
                          const double* wgt_, int nwalk, int ndet, int nex, int nact, int nelec,
                          int nchol, double* EX_, double* EJ_, double* KE_) {
  const cd* Tw = (const cd*)Tw_;
  const cd* Rall = (const cd*)R_;
  const cd* wgt = (const cd*)wgt_;
  cd* EX = (cd*)EX_;
  cd* EJ = (cd*)EJ_;
  cd* KE = (cd*)KE_;
  std::vector<int> occps(nelec);
  std::vector<cd> Fn(nchol), Fn1(nchol), Fa(nact), Kn(nchol);
  for (int iw = 0; iw < nwalk; ++iw) {
    auto T = [&](int i, int n, int a) -> cd {
      return Tw[(((size_t)iw * nelec + i) * nchol + n) * nact + a];
    };
    for (int idet = 0; idet < ndet; ++idet) {
      const int* holes = iexcit + (size_t)idet * 2 * nex;
      const int* parts = holes + nex;
      const cd* Rx = Rall + ((size_t)iw * ndet + idet) * nex * nact;
      std::fill(Kn.begin(), Kn.end(), cd(0.0));
      cd eX = 0.0;
      for (int i = 0; i < nelec; ++i) {
        occps[i] = refc[i];
        for (int e1 = 0; e1 < nex; ++e1)
          if (holes[e1] == i) { occps[i] = parts[e1]; break; }
      }
      for (int n = 0; n < nchol; ++n) {
        cd f = 0.0;
        for (int i = 0; i < nelec; ++i) f += T(i, n, refc[i]);
        Fn[n] = f;
      }
      cd eJ0 = 0.0;
      for (int n = 0; n < nchol; ++n) eJ0 += Fn[n] * Fn[n];
      for (int e1 = 0; e1 < nex; ++e1) {
        int ip = holes[e1];
        const cd* R1 = Rx + (size_t)e1 * nact;
        for (int n = 0; n < nchol; ++n) {
          cd f = 0.0;
          for (int a = 0; a < nact; ++a) f += T(ip, n, a) * R1[a];
          Fn[n] = f;
        }
        for (int n = 0; n < nchol; ++n) { eX += Fn[n] * Fn[n]; Kn[n] += Fn[n]; }
        for (int e2 = e1 + 1; e2 < nex; ++e2) {
          int iq = holes[e2];
          const cd* R2 = Rx + (size_t)e2 * nact;
          cd s = 0.0;
          for (int n = 0; n < nchol; ++n) {
            cd f = 0.0, g = 0.0;
            for (int a = 0; a < nact; ++a) {
              f += T(iq, n, a) * R1[a];
              g += T(ip, n, a) * R2[a];
            }
            s += f * g;
          }
          eX += 2.0 * s;
        }
        for (int j = 0; j < nelec; ++j) {
          int Oj = occps[j];
          for (int a = 0; a < nact; ++a) {
            cd f = 0.0;
            for (int n = 0; n < nchol; ++n) f += T(j, n, a) * T(ip, n, Oj);
            Fa[a] = f;
          }
          cd s = 0.0;
          for (int a = 0; a < nact; ++a) s += Fa[a] * R1[a];
          eX += 2.0 * s;
        }
      }
      for (int i = 0; i < nelec; ++i) {
        int Oi = occps[i], ri = refc[i];
        if (Oi != ri) {
          cd s = 0.0;
          for (int n = 0; n < nchol; ++n) s += T(i, n, Oi) * T(i, n, Oi) - T(i, n, ri) * T(i, n, ri);
          eX += s;
        }
        for (int j = i + 1; j < nelec; ++j) {
          int Oj = occps[j], rj = refc[j];
          if (Oi != ri || Oj != rj) {
            cd s = 0.0;
            for (int n = 0; n < nchol; ++n)
              s += T(i, n, Oj) * T(j, n, Oi) - T(i, n, rj) * T(j, n, ri);
            eX += 2.0 * s;
          }
        }
      }
      for (int i = 0; i < nelec; ++i)
        for (int n = 0; n < nchol; ++n) Kn[n] += T(i, n, occps[i]);
      cd w = wgt[(size_t)idet * nwalk + iw];
      EX[iw] -= 0.5 * w * eX;
      cd kk = 0.0;
      for (int n = 0; n < nchol; ++n) kk += Kn[n] * Kn[n];
      EJ[iw] += 0.5 * w * (kk - eJ0);
      std::copy(Kn.begin(), Kn.end(), KE + ((size_t)idet * nwalk + iw) * nchol);
    }
  }
}

// PHMSD::MixedDensityMatrix, compact, COLLINEAR (PHMSD.cpp:549-790): G [W][nact_a+nact_b][N] and
// Ov [W] = log(sum_d conj(ci_d) ov_a ov_b) + log_ov of the reference.
void orc_phmsd_mdm(const orc_ph* P, const double* Phi_a, const double* Phi_b, int W, double* G_,
                   double* Ov_) {
  const Spin Sa = spin(P, 0), Sb = spin(P, 1);
  const int N = P->N, nact = Sa.nact + Sb.nact;
  Couplings C(P);
  cd* G = (cd*)G_;
  cd* Ov = (cd*)Ov_;
#pragma omp parallel for schedule(dynamic)
  for (int w = 0; w < W; ++w) {
    const cd* Pa = (const cd*)Phi_a + (size_t)w * N * Sa.nel;
    const cd* Pb = (const cd*)Phi_b + (size_t)w * N * Sb.nel;
    std::vector<cd> GA((size_t)Sa.nel * N), GB((size_t)Sb.nel * N);
    std::vector<cd> Qa((size_t)Sa.nact * Sa.nel), Qb((size_t)Sb.nact * Sb.nel);
    std::vector<cd> ova(Sa.nconf), ovb(Sb.nconf), wa(Sa.nconf), wb(Sb.nconf);
    cd log_ov = 0.0;
    woodbury(Sa, Pa, GA.data(), Qa.data(), log_ov);
    overlaps(Sa, Qa.data(), ova.data());
    woodbury(Sb, Pb, GB.data(), Qb.data(), log_ov);
    overlaps(Sb, Qb.data(), ovb.data());
    C.apply(0, ovb.data(), Sa.nconf, wa.data());
    cd ov = 0.0;
    for (int u = 0; u < Sa.nconf; ++u) ov += ova[u] * wa[u];
    for (int u = 0; u < Sa.nconf; ++u) wa[u] *= ova[u];
    std::vector<cd> Ra((size_t)Sa.nel * Sa.nact), Rb((size_t)Sb.nel * Sb.nact);
    calculate_R(Sa, Qa.data(), wa.data(), Ra.data());
    C.apply(1, ova.data(), Sb.nconf, wb.data());
    for (int v = 0; v < Sb.nconf; ++v) wb[v] *= ovb[v];
    calculate_R(Sb, Qb.data(), wb.data(), Rb.data());
    cd* Gw = G + (size_t)w * nact * N;
    zgemm('T', 'N', Sa.nact, N, Sa.nel, 1.0, Ra.data(), Sa.nact, GA.data(), N, 0.0, Gw, N);
    zgemm('T', 'N', Sb.nact, N, Sb.nel, 1.0, Rb.data(), Sb.nact, GB.data(), N, 0.0,
          Gw + (size_t)Sa.nact * N, N);
    cd rinv = 1.0 / ov;
    for (size_t e = 0; e < (size_t)nact * N; ++e) Gw[e] *= rinv;
    Ov[w] = std::log(ov) + log_ov;
  }
}

// PHMSD::Log_Overlap, COLLINEAR (PHMSD.cpp:804-900).
void orc_phmsd_log_overlap(const orc_ph* P, const double* Phi_a, const double* Phi_b, int W,
                           double* Ov_) {
  const Spin Sa = spin(P, 0), Sb = spin(P, 1);
  const int N = P->N;
  Couplings C(P);
  cd* Ov = (cd*)Ov_;
#pragma omp parallel for schedule(dynamic)
  for (int w = 0; w < W; ++w) {
    const cd* Pa = (const cd*)Phi_a + (size_t)w * N * Sa.nel;
    const cd* Pb = (const cd*)Phi_b + (size_t)w * N * Sb.nel;
    std::vector<cd> Qa((size_t)Sa.nact * Sa.nel), Qb((size_t)Sb.nact * Sb.nel);
    std::vector<cd> ova(Sa.nconf), ovb(Sb.nconf), wa(Sa.nconf);
    cd log_ov = 0.0;
    woodbury(Sa, Pa, nullptr, Qa.data(), log_ov);
    overlaps(Sa, Qa.data(), ova.data());
    woodbury(Sb, Pb, nullptr, Qb.data(), log_ov);
    overlaps(Sb, Qb.data(), ovb.data());
    C.apply(0, ovb.data(), Sa.nconf, wa.data());
    cd ov = 0.0;
    for (int u = 0; u < Sa.nconf; ++u) ov += ova[u] * wa[u];
    Ov[w] = std::log(ov) + log_ov;
  }
}

// PHMSD::energy_alg1, COLLINEAR (PHMSD.cpp:376-544), walker by walker.
void orc_phmsd_energy(const orc_ph* P, const double* haj_, const double* Lnak_, int M, double E0,
                      const double* Phi_a, const double* Phi_b, int W, double* E_, double* Ov_) {
  const Spin Sa = spin(P, 0), Sb = spin(P, 1);
  const int N = P->N;
  Couplings C(P);
  const cd* haj = (const cd*)haj_;
  const cd* La = (const cd*)Lnak_;
  const cd* Lb = La + (size_t)M * Sa.nact * N;
  cd* E = (cd*)E_;
  cd* Ov = (cd*)Ov_;
  const int na = Sa.nconf, nb = Sb.nconf;
#pragma omp parallel for schedule(dynamic)
  for (int w = 0; w < W; ++w) {
    const cd* Pa = (const cd*)Phi_a + (size_t)w * N * Sa.nel;
    const cd* Pb = (const cd*)Phi_b + (size_t)w * N * Sb.nel;
    std::vector<cd> GA((size_t)Sa.nel * N), GB((size_t)Sb.nel * N);
    std::vector<cd> Qa((size_t)Sa.nact * Sa.nel), Qb((size_t)Sb.nact * Sb.nel);
    std::vector<cd> ova(na), ovb(nb);
    cd log_ov = 0.0;
    // 1. overlaps
    woodbury(Sa, Pa, GA.data(), Qa.data(), log_ov);
    overlaps(Sa, Qa.data(), ova.data());
    woodbury(Sb, Pb, GB.data(), Qb.data(), log_ov);
    overlaps(Sb, Qb.data(), ovb.data());
    cd Ew[3] = {0.0, 0.0, 0.0};
    std::vector<cd> KEr((size_t)std::max(na, nb) * M, cd(0.0));
    cd ov = 0.0;
    {  // spin up
      std::vector<cd> wgt(na);
      C.apply(0, ovb.data(), na, wgt.data());
      for (int u = 0; u < na; ++u) ov += ova[u] * wgt[u];
      for (int u = 0; u < na; ++u) wgt[u] *= ova[u];
      const int nel = Sa.nel, nact = Sa.nact;
      std::vector<cd> Swia((size_t)nel * nact), Tw((size_t)nel * M * nact);
      cd el[3] = {E0, 0.0, 0.0};
      std::fill(KEr.begin(), KEr.begin() + M, cd(0.0));
      ph_ref_energy(nel, nact, N, M, GA.data(), haj, La, el, KEr.data(), Swia.data(), Tw.data());
      {
        cd s = 0.0;
        for (int n = 0; n < M; ++n) s += KEr[n] * KEr[n];
        el[2] += 0.5 * s;
      }
      for (int c = 0; c < 3; ++c) Ew[c] += el[c] * ov;
      // first step: excited configurations shell by shell
      cd ex[3] = {0.0, 0.0, 0.0};
      for (int u = 1; u < na; ++u) {
        int r = Sa.rank[u];
        std::vector<cd> R((size_t)r * nact);
        compact_R(Sa, Qa.data(), u, R.data());
        orc_ph_excited_1body(Sa.e(u), Sa.refc, (const double*)Swia.data(), (const double*)R.data(),
                             (const double*)&wgt[u], 1, 1, r, nact, nel, (double*)&ex[0]);
        orc_ph_excited_2body(Sa.e(u), Sa.refc, (const double*)Tw.data(), (const double*)R.data(),
                             (const double*)&wgt[u], 1, 1, r, nact, nel, M, (double*)&ex[1],
                             (double*)&ex[2], (double*)(KEr.data() + (size_t)u * M));
      }
      for (int c = 0; c < 3; ++c) Ew[c] += ex[c];
      // KEr[u] *= ov_up[u]; KEr[v] = sum_u C[1][v][u] KEr[u]; KEr[v] *= ov_dn[v]
      for (int u = 0; u < na; ++u)
        for (int n = 0; n < M; ++n) KEr[(size_t)u * M + n] *= ova[u];
      std::vector<cd> Tdn(KEr.begin(), KEr.begin() + (size_t)na * M);
      for (int v = 0; v < nb; ++v) {
        for (int n = 0; n < M; ++n) {
          cd acc = 0.0;
          for (auto& kv : C.row[1][v]) acc += kv.second * Tdn[(size_t)kv.first * M + n];
          KEr[(size_t)v * M + n] = acc;
        }
        for (int n = 0; n < M; ++n) KEr[(size_t)v * M + n] *= ovb[v];
      }
    }
    {  // spin down
      std::vector<cd> wgt(nb);
      C.apply(1, ova.data(), nb, wgt.data());
      for (int v = 0; v < nb; ++v) wgt[v] *= ovb[v];
      const int nel = Sb.nel, nact = Sb.nact;
      std::vector<cd> Swia((size_t)nel * nact), Tw((size_t)nel * M * nact), KEl(M, cd(0.0));
      cd el[3] = {0.0, 0.0, 0.0};
      ph_ref_energy(nel, nact, N, M, GB.data(), haj + (size_t)Sa.nact * N, Lb, el, KEl.data(),
                    Swia.data(), Tw.data());
      {
        cd s = 0.0;
        for (int n = 0; n < M; ++n) s += KEl[n] * KEl[n];
        el[2] += 0.5 * s;
      }
      for (int c = 0; c < 3; ++c) Ew[c] += el[c] * ov;
      {
        cd t = 0.0;
        for (int n = 0; n < M; ++n) t += KEr[n] * KEl[n];
        Ew[2] += t;
      }
      cd ex[3] = {0.0, 0.0, 0.0};
      std::vector<cd> KEv(M);
      for (int v = 1; v < nb; ++v) {
        int r = Sb.rank[v];
        std::vector<cd> R((size_t)r * nact);
        compact_R(Sb, Qb.data(), v, R.data());
        orc_ph_excited_1body(Sb.e(v), Sb.refc, (const double*)Swia.data(), (const double*)R.data(),
                             (const double*)&wgt[v], 1, 1, r, nact, nel, (double*)&ex[0]);
        orc_ph_excited_2body(Sb.e(v), Sb.refc, (const double*)Tw.data(), (const double*)R.data(),
                             (const double*)&wgt[v], 1, 1, r, nact, nel, M, (double*)&ex[1],
                             (double*)&ex[2], (double*)KEv.data());
        cd t = 0.0;
        for (int n = 0; n < M; ++n) t += KEr[(size_t)v * M + n] * KEv[n];
        ex[2] += t;
      }
      for (int c = 0; c < 3; ++c) Ew[c] += ex[c];
    }
    for (int c = 0; c < 3; ++c) E[(size_t)w * 3 + c] = Ew[c] / ov;
    Ov[w] = std::log(ov) + log_ov;
  }
}

// The trial's own density for PHMSD::vMF, COLLINEAR host path (PHMSD.cpp:1329-1573): per spin,
// the diagonal sum over configurations (each configuration's compact density, herm = false
// overlaps Orbs^H Orbs) weighted by |ci|^2 and the configuration overlaps, plus the
// single-excitation couplings between configurations that share the other spin's
// configuration; divided by the trial norm. Gfull [nact_a + nact_b][N].
void orc_phmsd_gmf(const orc_ph* P, double* Gfull_) {
  const int N = P->N;
  Couplings C(P);
  cd* Gfull = (cd*)Gfull_;
  std::vector<cd> ovl[2];
  const cd* ci = (const cd*)P->ci;
  for (int s = 0; s < 2; ++s) {
    Spin S = spin(P, s);
    ovl[s].assign(S.nconf, cd(0.0));
    std::vector<int> confg(S.nel);
    std::vector<cd> A((size_t)N * S.nel), T((size_t)S.nel * S.nel);
    std::vector<int> piv(S.nel);
    for (int u = 0; u < S.nconf; ++u) {
      configuration(S, u, confg.data());
      // A(:, i) = Orbs(:, confg(i)) with Orbs = orbs^H; T = A^H A
      for (int k = 0; k < N; ++k)
        for (int i = 0; i < S.nel; ++i) A[(size_t)k * S.nel + i] = std::conj(S.orbs[(size_t)confg[i] * N + k]);
      for (int i = 0; i < S.nel; ++i)
        for (int j = 0; j < S.nel; ++j) {
          cd t = 0.0;
          for (int k = 0; k < N; ++k) t += std::conj(A[(size_t)k * S.nel + i]) * A[(size_t)k * S.nel + j];
          T[(size_t)i * S.nel + j] = t;
        }
      lu_factor(T.data(), S.nel, piv.data());
      logdet_from_getrf(T.data(), piv.data(), S.nel, ovl[s][u]);
    }
    for (int u = 1; u < S.nconf; ++u) ovl[s][u] = std::exp(ovl[s][u] - ovl[s][0]);
    ovl[s][0] = 1.0;
  }
  cd ov = 0.0;
  for (int d = 0; d < P->ndet; ++d) ov += std::norm(ci[d]) * ovl[0][P->det_a[d]] * ovl[1][P->det_b[d]];
  int row0 = 0;
  for (int s = 0; s < 2; ++s) {
    Spin S = spin(P, s);
    const int other = 1 - s, n = S.nel;
    cd* Gs = Gfull + (size_t)row0 * N;
    std::fill(Gs, Gs + (size_t)S.nact * N, cd(0.0));
    std::vector<int> confg(n), confgB(n), piv(n);
    std::vector<cd> A((size_t)N * n), T((size_t)n * n), G2((size_t)n * N);
    for (int u = 0; u < S.nconf; ++u) {  // diagonal
      configuration(S, u, confg.data());
      for (int k = 0; k < N; ++k)
        for (int i = 0; i < n; ++i) A[(size_t)k * n + i] = std::conj(S.orbs[(size_t)confg[i] * N + k]);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          cd t = 0.0;
          for (int k = 0; k < N; ++k) t += std::conj(A[(size_t)k * n + i]) * A[(size_t)k * n + j];
          T[(size_t)i * n + j] = t;
        }
      int info = lu_factor(T.data(), n, piv.data());
      inverse_from_lu(T.data(), piv.data(), n, info);
      // G2 = T^{-T} A^T: G2[i][k] = sum_j Tinv[j][i] A[k][j]
      zgemm('T', 'T', n, N, n, 1.0, T.data(), n, A.data(), n, 0.0, G2.data(), N);
      cd wgt = 0.0;
      for (auto& kv : C.row[s][u]) wgt += ovl[other][kv.first] * std::norm(kv.second);
      wgt *= ovl[s][u];
      for (int i = 0; i < n; ++i)
        for (int k = 0; k < N; ++k) Gs[(size_t)confg[i] * N + k] += wgt * G2[(size_t)i * N + k];
    }
    for (int v = 0; v < P->nconf[other]; ++v) {  // off-diagonal single excitations
      auto& row = C.row[other][v];
      for (size_t n1 = 0; n1 < row.size(); ++n1) {
        configuration(S, row[n1].first, confg.data());
        std::sort(confg.begin(), confg.end());
        for (size_t n2 = n1 + 1; n2 < row.size(); ++n2) {
          configuration(S, row[n2].first, confgB.data());
          std::sort(confgB.begin(), confgB.end());
          int np = 0, hole = -1, part = -1;
          for (int i = 0; i < n; ++i)
            if (!std::binary_search(confgB.begin(), confgB.end(), confg[i])) { ++np; hole = confg[i]; }
          for (int i = 0; i < n; ++i)
            if (!std::binary_search(confg.begin(), confg.end(), confgB[i])) part = confgB[i];
          if (np != 1) continue;
          cd wgt = row[n1].second * std::conj(row[n2].second);
          for (int k = 0; k < N; ++k) {
            Gs[(size_t)part * N + k] += wgt * S.orbs[(size_t)hole * N + k];
            Gs[(size_t)hole * N + k] += std::conj(wgt) * S.orbs[(size_t)part * N + k];
          }
        }
      }
    }
    row0 += S.nact;
  }
  cd rinv = 1.0 / ov;
  for (size_t e = 0; e < (size_t)row0 * N; ++e) Gfull[e] *= rinv;
}

// Explicit multi-determinant sum (the metamorphic pin): determinant d occupies the sorted
// active orbitals occa[d] / occb[d] with coefficient ci[d]. G, E and Ov as orc_phmsd_mdm /
// orc_phmsd_energy; the imaginary part of Ov is only defined modulo 2 pi.
void orc_msd_brute(int N, int nup, int ndn, int nact_a, int nact_b, const double* orbs_a,
                   const double* orbs_b, int ndet, const int* occa, const int* occb,
                   const double* ci_, const double* haj_, const double* Lnak_, int M, double E0,
                   const double* Phi_a, const double* Phi_b, int W, double* G_, double* E_,
                   double* Ov_) {
  const cd* ci = (const cd*)ci_;
  const cd* haj = (const cd*)haj_;
  const cd* L[2] = {(const cd*)Lnak_, (const cd*)Lnak_ + (size_t)M * nact_a * N};
  const cd* orbs[2] = {(const cd*)orbs_a, (const cd*)orbs_b};
  const int nel[2] = {nup, ndn}, nact[2] = {nact_a, nact_b}, row0[2] = {0, nact_a};
  const int* occ[2] = {occa, occb};
  const int nrow = nact_a + nact_b;
  cd* G = (cd*)G_;
  cd* E = (cd*)E_;
  cd* Ov = (cd*)Ov_;
#pragma omp parallel for schedule(dynamic)
  for (int w = 0; w < W; ++w) {
    const cd* Phi[2] = {(const cd*)Phi_a + (size_t)w * N * nup, (const cd*)Phi_b + (size_t)w * N * ndn};
    std::vector<cd> Gacc((size_t)nrow * N, cd(0.0));
    cd Eacc[3] = {0.0, 0.0, 0.0}, wsum = 0.0, lref = 0.0;
    for (int d = 0; d < ndet; ++d) {
      cd ld = 0.0, e1 = 0.0, exx = 0.0;
      std::vector<std::vector<cd>> Gd(2);
      for (int s = 0; s < 2; ++s) {
        const int n = nel[s];
        std::vector<cd> A((size_t)n * N), T((size_t)n * n);
        for (int i = 0; i < n; ++i)
          std::copy(orbs[s] + (size_t)occ[s][(size_t)d * n + i] * N,
                    orbs[s] + (size_t)(occ[s][(size_t)d * n + i] + 1) * N, A.begin() + (size_t)i * N);
        log_overlap_one(A.data(), Phi[s], N, n, T.data(), ld, true);
        Gd[s].resize((size_t)n * N);
        zgemm('T', 'T', n, N, n, 1.0, T.data(), n, Phi[s], n, 0.0, Gd[s].data(), N);
      }
      if (d == 0) lref = ld;
      cd wd = std::conj(ci[d]) * std::exp(ld - lref);
      std::vector<cd> Kl(M, cd(0.0));
      for (int s = 0; s < 2; ++s) {
        const int n = nel[s];
        const int* o = occ[s] + (size_t)d * n;
        for (int i = 0; i < n; ++i)
          for (int k = 0; k < N; ++k) {
            e1 += Gd[s][(size_t)i * N + k] * haj[(size_t)(row0[s] + o[i]) * N + k];
            Gacc[(size_t)(row0[s] + o[i]) * N + k] += wd * Gd[s][(size_t)i * N + k];
          }
        std::vector<cd> T((size_t)n * n);
        for (int m = 0; m < M; ++m) {
          // T[b][a] = sum_k G[b][k] L[m][occ(a)][k]
          for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a) {
              const cd* l = L[s] + ((size_t)m * nact[s] + o[a]) * N;
              cd t = 0.0;
              for (int k = 0; k < N; ++k) t += Gd[s][(size_t)b * N + k] * l[k];
              T[(size_t)b * n + a] = t;
            }
          cd sx = 0.0, tr = 0.0;
          for (int a = 0; a < n; ++a) {
            tr += T[(size_t)a * n + a];
            for (int b = 0; b < n; ++b) sx += T[(size_t)a * n + b] * T[(size_t)b * n + a];
          }
          exx += -0.5 * sx;
          Kl[m] += tr;
        }
      }
      cd ej = 0.0;
      for (int m = 0; m < M; ++m) ej += 0.5 * Kl[m] * Kl[m];
      Eacc[0] += wd * e1;
      Eacc[1] += wd * exx;
      Eacc[2] += wd * ej;
      wsum += wd;
    }
    for (size_t e = 0; e < (size_t)nrow * N; ++e) G[(size_t)w * nrow * N + e] = Gacc[e] / wsum;
    E[(size_t)w * 3 + 0] = Eacc[0] / wsum + E0;
    E[(size_t)w * 3 + 1] = Eacc[1] / wsum;
    E[(size_t)w * 3 + 2] = Eacc[2] / wsum;
    Ov[w] = std::log(wsum) + lref;
  }
}

}  // extern "C"
