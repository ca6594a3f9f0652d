// orc_dense.hpp -- dense helpers shared by the oracle sources (defined in afq_oracle.cpp).
//
// THIS IS TEST INFRASTRUCTURE (see afq_oracle.cpp). Row-major complex matrices.
#pragma once

#include <complex>

typedef std::complex<double> cd;

namespace orc {

// C[m][n] = alpha * opA(A) opB(B) + beta C, op = 'N' or 'T' (no conjugation).
void zgemm(char ta, char tb, int m, int n, int k, cd alpha, const cd* A, int lda, const cd* B,
           int ldb, cd beta, cd* C, int ldc);
// LU with partial pivoting (zgetf2 semantics), 1-based pivots; returns info.
int lu_factor(cd* a, int n, int* piv);
// res += sum log(+-U_ii + DBL_MIN), imaginary part wrapped into [-pi, pi].
void logdet_from_getrf(const cd* a, const int* piv, int n, cd& res);
// getri_or_zero: inverse from the LU factors, zeros when info != 0.
void inverse_from_lu(cd* a, const int* piv, int n, int info);
// T = A * B (A [n][N], B [N][n]), ovlp += log det T, optionally T <- T^{-1}.
void log_overlap_one(const cd* A, const cd* B, int N, int n, cd* T, cd& ovlp, bool invert);

}  // namespace orc
