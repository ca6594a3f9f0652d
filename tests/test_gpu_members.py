"""HamiltonianOperations members beyond vbias / vHS / energy, through the C ABI.

The per-spin energy (Real3IndexFactorization.hpp:176-210): for COLLINEAR walkers the alpha and
beta calls, combined with the opposite-spin Coulomb term sum_m EJn_a EJn_b, give the full
energy of both blocks (the oracle's, 1e-10 relative), and EJn equals tr(L^m G_s) (the force
bias of the block over sqrt(dt)); for CLOSED walkers the alpha call is the full energy.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd.synthetic import make_problem, random_walkers  # noqa: E402

TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import afq
    return afq, torch


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_energy_spin_collinear(gpu, precision):
    afq, torch = gpu
    N, nup, ndn, M, W = 37, 7, 5, 60, 9
    say(f"case: per-spin energy COLLINEAR N={N} n={nup}+{ndn} M={M} W={W} {precision}")
    p = make_problem(N, nup, M, seed=44)
    PsiT = np.concatenate([p.PsiT, p.PsiT[:ndn]])
    ctx = afq.Context(0)
    prec = afq.AFQ_FP64 if precision == "fp64" else afq.AFQ_MIXED
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, PsiT, nup, ndn,
                                    walker_type=afq.AFQ_COLLINEAR, E0=1.5, precision=prec)
    Pa = random_walkers(p, W, seed=1)
    Pb = np.ascontiguousarray(random_walkers(p, W, seed=2)[:, :, :ndn])
    Ga, _ = orc.mixed_density_matrix(p.PsiT, Pa)
    Gb, _ = orc.mixed_density_matrix(PsiT[nup:], Pb)
    Gfull = np.concatenate([Ga, Gb], axis=1)
    E = afq.empty((W, 3), torch.complex128)
    ham.energy(E, afq.device_tensor(Gfull), precision=prec)
    Es, K = [], []
    for spin, G in ((0, Ga), (1, Gb)):
        e = afq.empty((W, 3), torch.complex128)
        k = afq.empty((W, M), torch.complex128)
        ham.energy_spin(spin, e, afq.device_tensor(G), k, precision=prec)
        Es.append(afq.host(e))
        K.append(afq.host(k))
    tot = Es[0] + Es[1]
    tot[:, 2] += np.sum(K[0] * K[1], axis=1)
    tol = TOL if precision == "fp64" else 1e-5
    assert rel(tot, afq.host(E)) < tol
    L, haj, _ = ham.tensors()
    La = L[: M * nup * N].reshape(M, nup, N)
    Lb = L[M * nup * N:].reshape(M, ndn, N)
    if precision == "fp64":
        Er = orc.energy(haj, L, Gfull, nup, ndn, closed=False, E0=1.5)
        assert rel(afq.host(E), Er) < TOL
        for k, G, Ls in ((K[0], Ga, La), (K[1], Gb, Lb)):
            assert rel(k, orc.vbias(Ls, G, 1.0, c=1.0)) < TOL
        # the beta block carries no E0: E1_b = sum G_b haj_b
        e1b = np.einsum("wak,ak->w", Gb, haj[nup:])
        assert rel(Es[1][:, 0], e1b) < TOL
    # without EJ: no Coulomb vectors needed
    e = afq.empty((W, 3), torch.complex128)
    ham.energy_spin(0, e, afq.device_tensor(Ga), None, addEJ=False, precision=prec)
    assert np.array_equal(afq.host(e)[:, :2], Es[0][:, :2]) and np.all(afq.host(e)[:, 2] == 0)


def test_energy_spin_closed_is_the_full_energy(gpu):
    afq, torch = gpu
    p = make_problem(24, 5, 96, seed=1001)
    W = 11
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, p.nup, E0=0.25)
    G, _ = orc.mixed_density_matrix(p.PsiT, random_walkers(p, W, seed=3))
    dG = afq.device_tensor(G)
    E = afq.empty((W, 3), torch.complex128)
    ham.energy(E, dG)
    e = afq.empty((W, 3), torch.complex128)
    k = afq.empty((W, p.M), torch.complex128)
    ham.energy_spin(0, e, dG, k)
    # CLOSED: E[2] = 1/2 c^2 sum K^2 with c = 2, the per-spin call returns 1/2 sum K^2
    eh, Eh = afq.host(e), afq.host(E)
    assert rel(eh[:, :2], Eh[:, :2]) < TOL
    assert rel(4.0 * eh[:, 2], Eh[:, 2]) < TOL
    with pytest.raises(afq.AfqError):
        ham.energy_spin(1, e, dG, k)
