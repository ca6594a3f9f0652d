"""The density-matrix observables (OneRDM, DiagonalTwoRDM) of the oracle, pinned by brute force in
Fock space, and the product's fold of the diagonal two-body blocks into the reference's layout.

For Slater determinants |Psi> (the trial, PsiT = Psi^H) and |Phi> (a walker) the mixed
expectation <Psi| c+_i c_j |Phi> / <Psi|Phi> and <Psi| c+_is c+_js' c_js' c_is |Phi> / <Psi|Phi> are
computed from explicit Jordan-Wigner operators on the 2^(2N)-dimensional Fock space; the oracle
builds the same numbers from the mixed density matrices through Wick's theorem
(diagonal_twordm.hpp:55-78).  A dropped exchange term, a transposed G or a wrong spin pairing
fails the comparison.
"""
import numpy as np
import pytest

import oracle as orc
from safire_amd import afq


def jw_annihilators(K):
    """c_p for p < K as dense 2^K x 2^K matrices (mode p = bit p of the basis index)."""
    dim = 1 << K
    ops = []
    for p in range(K):
        c = np.zeros((dim, dim))
        for s in range(dim):
            if s >> p & 1:
                sign = (-1) ** bin(s & ((1 << p) - 1)).count("1")
                c[s ^ (1 << p), s] = sign
        ops.append(c)
    return ops


def slater_state(cdag, orbs_by_spin, N):
    """prod_k (sum_i orb[i,k] c+_{i,spin}) |0>; spin-orbital (i, s) is mode i + s N."""
    v = np.zeros(cdag[0].shape[0], np.complex128)
    v[0] = 1.0
    for s, orbs in enumerate(orbs_by_spin):
        for k in range(orbs.shape[1]):
            b = sum(orbs[i, k] * cdag[i + s * N] for i in range(N))
            v = b @ v
    return v


@pytest.mark.parametrize("closed", [True, False])
def test_oracle_rdm_vs_fock_space(closed):
    rng = np.random.default_rng(11 if closed else 12)
    N, nup, ndn = 4, 2, 2 if closed else 1
    K = 2 * N
    c = jw_annihilators(K)
    cdag = [x.T.copy() for x in c]
    W = 3
    PsiT = [np.linalg.qr(rng.normal(size=(N, nup)) + 1j * rng.normal(size=(N, nup)))[0].conj().T]
    PsiT.append(PsiT[0] if closed else
                np.linalg.qr(rng.normal(size=(N, ndn)) + 1j * rng.normal(size=(N, ndn)))[0].conj().T)
    weights = rng.normal(size=W) + 1j * rng.normal(size=W)
    coeff = rng.normal(size=W) + 1j * rng.normal(size=W)
    nsp = 1 if closed else 2
    G = np.empty((W, nsp, N, N), np.complex128)
    one_bf = np.zeros((nsp, N, N), np.complex128)
    two_bf = np.zeros((2 if closed else 3, N, N), np.complex128)
    psi = slater_state(cdag, [PsiT[0].conj().T, PsiT[1].conj().T], N)
    for w in range(W):
        Pu = rng.normal(size=(N, nup)) + 1j * rng.normal(size=(N, nup))
        Pd = Pu if closed else rng.normal(size=(N, ndn)) + 1j * rng.normal(size=(N, ndn))
        phi = slater_state(cdag, [Pu, Pd], N)
        norm = np.vdot(psi, phi)

        def ev(op):
            return np.vdot(psi, op @ phi) / norm

        G[w, 0] = orc.mixed_density_matrix(PsiT[0], Pu[None], compact=False)[0][0]
        if not closed:
            G[w, 1] = orc.mixed_density_matrix(PsiT[1], Pd[None], compact=False)[0][0]
        cw = weights[w] * coeff[w]
        for s in range(nsp):
            for i in range(N):
                for j in range(N):
                    one_bf[s, i, j] += cw * ev(cdag[i + s * N] @ c[j + s * N])
        # (aaaa), (aabb), (bbbb): <c+_is c+_js' c_js' c_is>
        pairs = [(0, 0), (0, 1)] + ([] if closed else [(1, 1)])
        for b, (s1, s2) in enumerate(pairs):
            for i in range(N):
                for j in range(N):
                    a, bb = i + s1 * N, j + s2 * N
                    two_bf[b, i, j] += cw * ev(cdag[a] @ cdag[bb] @ c[bb] @ c[a])
    one_bf /= weights.sum()
    two_bf /= weights.sum()
    one, two = orc.rdm_measure(G, weights, closed, coeff)
    assert np.abs(one - one_bf).max() < 1e-12
    assert np.abs(two - two_bf).max() < 1e-12
    # G[i][j] is <c+_i c_j>: the trace counts the electrons of each spin
    assert abs(np.trace(G[0, 0]) - nup) < 1e-12


@pytest.mark.parametrize("closed", [True, False])
def test_fold_matches_reference_layout(closed):
    rng = np.random.default_rng(3)
    N = 7
    full = rng.normal(size=(2 if closed else 3, N, N)) + 1j * rng.normal(size=(2 if closed else 3, N, N))
    ref = orc.fold_diag2rdm(full, closed)
    size = N * (2 * N - 1) - (N * (N - 1) // 2 if closed else 0)   # diagonal_twordm.hpp:95-101
    assert ref.shape == (size,)
    got = afq.fold_diag2rdm(full, afq.AFQ_CLOSED if closed else afq.AFQ_COLLINEAR)
    assert np.array_equal(got, ref)
