"""The register-tiled complex FP64 product (zgemm_rt.hpp) behind the density matrix and overlap.

Through the C ABI: MixedDensityMatrix (compact and full) and Log_Overlap against the FP64 oracle
per walker (relative 1e-12 of each walker's largest entry; log overlaps to 1e-11) and against
the LDS-staged GEMM engine it replaces (AFQ_ZRT=0), at shapes whose electron counts fill one,
two or three 16-row tiles or fall off them, with ragged k steps (N % 4 != 0, n % 4 != 0), a
column count past one 32-column block, a walker stride wider than the Slater block, and a
single electron.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import afq
    return afq, afq.Context(0), torch


def _with_zrt(on, fn):
    old = os.environ.get("AFQ_ZRT")
    os.environ["AFQ_ZRT"] = "1" if on else "0"
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["AFQ_ZRT"]
        else:
            os.environ["AFQ_ZRT"] = old


def _walkers(rng, W, N, n, PsiT, noise=0.3):
    base = PsiT.conj().T
    pert = rng.normal(size=(W, N, n)) + 1j * rng.normal(size=(W, N, n))
    return np.ascontiguousarray(base[None] + noise * pert / np.sqrt(N))


def _trial(rng, N, n):
    q, _ = np.linalg.qr(rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
    return np.ascontiguousarray(q[:, :n].conj().T)


@pytest.mark.parametrize("N,n,W", [(200, 40, 64), (24, 5, 7), (33, 17, 9), (64, 48, 5),
                                   (50, 1, 3), (37, 30, 11), (120, 16, 4)])
def test_zgemm_rt_density_and_overlap(gpu, N, n, W):
    afq, ctx, torch = gpu
    say(f"case: register-tiled density matrix N={N} n={n} W={W}")
    rng = np.random.default_rng(N * 100 + n)
    PsiT = _trial(rng, N, n)
    Phi = _walkers(rng, W, N, n, PsiT)
    dP = afq.device_tensor(PsiT)
    Gr, ovr = orc.mixed_density_matrix(PsiT, Phi, compact=True)
    Gfr, _ = orc.mixed_density_matrix(PsiT, Phi, compact=False)
    lor = orc.log_overlap(PsiT, Phi)

    def run():
        dPhi = afq.device_tensor(Phi)
        G = afq.empty((W, n, N), torch.complex128)
        ov = afq.zeros(W, torch.complex128)
        afq.det_ops.MixedDensityMatrix(ctx, dP, dPhi, G, ov, compact=True)
        Gf = afq.empty((W, N, N), torch.complex128)
        ovf = afq.zeros(W, torch.complex128)
        afq.det_ops.MixedDensityMatrix(ctx, dP, dPhi, Gf, ovf, compact=False)
        lo = afq.zeros(W, torch.complex128)
        afq.det_ops.Log_Overlap(ctx, dP, dPhi, lo)
        return afq.host(G), afq.host(ov), afq.host(Gf), afq.host(lo)
    G, ov, Gf, lo = _with_zrt(True, run)
    Ge, ove, Gfe, loe = _with_zrt(False, run)
    for w in range(W):
        sc = np.abs(Gr[w]).max()
        assert np.abs(G[w] - Gr[w]).max() < 1e-12 * sc, w
        assert np.abs(G[w] - Ge[w]).max() < 1e-12 * sc, w
        scf = np.abs(Gfr[w]).max()
        assert np.abs(Gf[w] - Gfr[w]).max() < 1e-12 * scf, w
    assert np.abs(ov - ovr).max() < 1e-11 and np.abs(ov - ove).max() < 1e-11
    assert np.abs(lo - lor).max() < 1e-11 and np.abs(lo - loe).max() < 1e-11
    say(f"  G vs oracle {np.abs(G - Gr).max() / np.abs(Gr).max():.2e}, "
        f"log overlap {np.abs(lo - lor).max():.2e}")


def test_zgemm_rt_strided_walkers(gpu):
    """Walkers inside a wider buffer (block stride past N n): the stride reaches the kernel."""
    afq, ctx, torch = gpu
    N, n, W, pad = 40, 12, 6, 37
    rng = np.random.default_rng(3)
    PsiT = _trial(rng, N, n)
    Phi = _walkers(rng, W, N, n, PsiT)
    buf = np.zeros((W, N * n + pad), np.complex128)
    buf[:, :N * n] = Phi.reshape(W, N * n)
    dbuf = afq.device_tensor(buf)
    view = dbuf[:, :N * n].view(W, N, n)
    G = afq.empty((W, n, N), torch.complex128)
    ov = afq.zeros(W, torch.complex128)
    afq.det_ops.MixedDensityMatrix(ctx, afq.device_tensor(PsiT), view, G, ov, compact=True)
    Gr, ovr = orc.mixed_density_matrix(PsiT, Phi, compact=True)
    assert np.abs(afq.host(G) - Gr).max() < 1e-12 * np.abs(Gr).max()
    assert np.abs(afq.host(ov) - ovr).max() < 1e-11


def test_zgemm_rt_blockings_same_bits(gpu):
    """Every blocking the tuner may pick accumulates each output tile in the same order: forcing
    each candidate of the density matrix's two products gives the same bits."""
    afq, ctx, torch = gpu
    N, n, W = 72, 40, 5
    rng = np.random.default_rng(11)
    PsiT = _trial(rng, N, n)
    Phi = _walkers(rng, W, N, n, PsiT)
    dP, dPhi = afq.device_tensor(PsiT), afq.device_tensor(Phi)
    out = []
    for i in range(5):
        afq.tune_force("mdm_T_rt:f64", i)
        afq.tune_force("mdm_G_rt:f64", i)
        try:
            G = afq.empty((W, n, N), torch.complex128)
            ov = afq.zeros(W, torch.complex128)
            afq.det_ops.MixedDensityMatrix(ctx, dP, dPhi, G, ov, compact=True)
            out.append((afq.host(G), afq.host(ov)))
        finally:
            afq.tune_force(None)
    for G, ov in out[1:]:
        assert np.array_equal(G, out[0][0]) and np.array_equal(ov, out[0][1])
    Gr, _ = orc.mixed_density_matrix(PsiT, Phi, compact=True)
    assert np.abs(out[0][0] - Gr).max() < 1e-12 * np.abs(Gr).max()
