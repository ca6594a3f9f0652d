"""The register-tiled FP32 complex x real product (crgemm.hpp) behind the mixed-precision vHS.

Checks, through the C ABI: against the FP64 oracle's vHS within the FP32 bound (the inputs
rounded to FP32 and an FP32 sum over M terms: relative error below 1e-5 of the largest entry,
with the largest per-walker error stated); against the LDS-staged GEMM engine it replaces
(AFQ_CRGEMM=0) within the same bound; bit-for-bit repeatable; exactly symmetric V; at shapes
whose walkers, pairs and Cholesky vectors fall off the 64 x 112 x 16 tiling (clamped rows,
invalid pairs and a ragged k group, taken first with the reversed order).
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd.synthetic import make_problem  # noqa: E402

BOUND = 1e-5


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


def _vhs(afq, torch, ham, X, N, crg):
    W = X.shape[0]
    old = os.environ.get("AFQ_CRGEMM")
    os.environ["AFQ_CRGEMM"] = "1" if crg else "0"
    try:
        V = afq.empty((W, N, N), torch.complex64)
        ham.vHS(afq.device_tensor(X.astype(np.complex64), np.complex64), V, 0.01,
                precision=afq.AFQ_MIXED)
        return afq.host(V)
    finally:
        if old is None:
            del os.environ["AFQ_CRGEMM"]
        else:
            os.environ["AFQ_CRGEMM"] = old


# M % 4 != 0 (the last case) is outside the kernel's alignment: both runs take the GEMM engine
@pytest.mark.parametrize("N,M,W", [(200, 1600, 96), (21, 40, 37), (64, 332, 130), (8, 20, 5),
                                   (10, 33, 7)])
def test_crgemm_vhs(gpu, N, M, W):
    afq, ctx, torch = gpu
    say(f"case: register-tiled vHS N={N} M={M} W={W}")
    p = make_problem(N, 4, M, 13)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, 4, precision=afq.AFQ_MIXED)
    rng = np.random.default_rng(5)
    X = rng.normal(size=(W, M)) + 1j * rng.normal(size=(W, M))
    Vc = _vhs(afq, torch, ham, X, N, True)
    Vc2 = _vhs(afq, torch, ham, X, N, True)
    assert np.array_equal(Vc, Vc2)
    assert np.array_equal(Vc, np.transpose(Vc, (0, 2, 1)))
    Vg = _vhs(afq, torch, ham, X, N, False)
    # the oracle on the FP32-rounded inputs isolates the product's own rounding
    X32 = X.astype(np.complex64).astype(np.complex128)
    L32 = p.Likn.astype(np.float32).astype(np.float64)
    Vr = orc.vhs(L32, X32, 0.01)
    scale = np.abs(Vr).max()
    e_or = np.abs(Vc - Vr).max() / scale
    e_gm = np.abs(Vc - Vg).max() / scale
    e_w = max(np.abs(Vc[w] - Vr[w]).max() / np.abs(Vr[w]).max() for w in range(W))
    say(f"  vs oracle {e_or:.2e} (worst walker {e_w:.2e}), vs the LDS-staged GEMM {e_gm:.2e}")
    assert e_or < BOUND and e_gm < BOUND and e_w < 10 * BOUND


def _step(afq, torch, ham, Pa, Pb, u, crg):
    old = os.environ.get("AFQ_CRGEMM")
    os.environ["AFQ_CRGEMM"] = "1" if crg else "0"
    try:
        W = Pa.shape[0]
        prop = afq.Propagator(ham, W, dt=0.005, seed=3)
        ws = afq.WalkerSet(Pa, Pb)
        E = afq.empty((W, 3), torch.complex128)
        ov = afq.empty(W, torch.complex128)
        prop.Energy(ws, E, ov, afq.AFQ_FP64)
        ws.eloc.copy_(E.sum(1))
        ws.ovlp.copy_(ov)
        Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).mean().real)]))
        prop.Propagate(ws, Esh, u=afq.device_tensor(u))
        out = [afq.host(ws.Phi), None if Pb is None else afq.host(ws.Phi_b),
               afq.host(ws.weight), afq.host(ws.ovlp)]
        prop.close()
        return out
    finally:
        if old is None:
            del os.environ["AFQ_CRGEMM"]
        else:
            os.environ["AFQ_CRGEMM"] = old


@pytest.mark.parametrize("closed", [True, False])
def test_crgemm_mixed_step_matches_gemm_engine(gpu, closed):
    """A whole mixed-precision step (vbias and vHS through the register-tiled kernel, vbias
    split over K) against the same step on the LDS-staged engine: the walkers, weights and
    overlaps agree within the FP32 bound."""
    afq, ctx, torch = gpu
    from safire_amd.synthetic import random_walkers
    N, nup, M, W = 96, 12, 512, 200
    say(f"case: mixed step, register-tiled vs GEMM engine, closed={closed} N={N} M={M} W={W}")
    p = make_problem(N, nup, M, 21)
    if closed:
        ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, nup, precision=afq.AFQ_MIXED)
        Pa, Pb = random_walkers(p, W, seed=4, noise=0.2), None
    else:
        ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, np.concatenate([p.PsiT, p.PsiT[:9]]),
                                        nup, 9, walker_type=afq.AFQ_COLLINEAR,
                                        precision=afq.AFQ_MIXED)
        import dataclasses
        Pa = random_walkers(p, W, seed=4, noise=0.2)
        Pb = random_walkers(dataclasses.replace(p, nup=9, PsiT=p.PsiT[:9]), W, seed=5, noise=0.2)
    u = orc.field_uniforms(8, 1, 0, W, M)
    a = _step(afq, torch, ham, Pa, Pb, u, True)
    b = _step(afq, torch, ham, Pa, Pb, u, False)
    for name, x, y in zip(("Phi", "Phi_b", "weight", "ovlp"), a, b):
        if x is None:
            continue
        e = np.abs(x - y).max() / np.abs(y).max()
        say(f"  {name} {e:.2e}")
        assert e < 1e-5, name


def _with_crg(crg, fn):
    old = os.environ.get("AFQ_CRGEMM")
    os.environ["AFQ_CRGEMM"] = "1" if crg else "0"
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["AFQ_CRGEMM"]
        else:
            os.environ["AFQ_CRGEMM"] = old


@pytest.mark.parametrize("N,M,W", [(200, 1600, 64), (21, 40, 37), (64, 332, 130), (8, 20, 5)])
def test_crgemm_fp64_vhs_and_vbias(gpu, N, M, W):
    """FP64 vHS and vbias through the register-tiled kernel (vbias split over K) against the
    oracle at 1e-12 and against the LDS-staged engine."""
    afq, ctx, torch = gpu
    say(f"case: register-tiled FP64 vHS + vbias N={N} M={M} W={W}")
    nup = min(4, N)
    p = make_problem(N, nup, M, 17)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, nup)
    rng = np.random.default_rng(9)
    X = rng.normal(size=(W, M)) + 1j * rng.normal(size=(W, M))
    G = rng.normal(size=(W, nup, N)) + 1j * rng.normal(size=(W, nup, N))

    def run():
        V = afq.empty((W, N, N), torch.complex128)
        ham.vHS(afq.device_tensor(X), V, 0.01)
        v = afq.empty((W, M), torch.complex128)
        ham.vbias(afq.device_tensor(G), v, 0.01)
        return afq.host(V), afq.host(v)
    Vc, vc = _with_crg(True, run)
    Vg, vg = _with_crg(False, run)
    Vr = orc.vhs(p.Likn, X, 0.01)
    Lnak, _, _ = ham.tensors()
    vr = orc.vbias(Lnak.reshape(M, nup, N), G, 0.01, c=2.0)
    eV = np.abs(Vc - Vr).max() / np.abs(Vr).max()
    ev = np.abs(vc - vr).max() / np.abs(vr).max()
    gV = np.abs(Vc - Vg).max() / np.abs(Vr).max()
    gv = np.abs(vc - vg).max() / np.abs(vr).max()
    say(f"  vHS vs oracle {eV:.2e} vs engine {gV:.2e}; vbias vs oracle {ev:.2e} vs engine {gv:.2e}")
    assert max(eV, ev, gV, gv) < 1e-12
    assert np.array_equal(Vc, np.transpose(Vc, (0, 2, 1)))
