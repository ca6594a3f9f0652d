"""Kept overlap factors: the step's density matrix takes T^-1 and log det T from the previous
step's overlap for every walker whose T = PsiT Phi comes out bit for bit the same.

The kept values are the factorization of the same bits, so a run must be bitwise identical to
one that recomputes them (AFQ_LU_KEEP=0): checked over graph-replayed and eager steps with QR
and comb population control in between (both change walkers, so their T no longer match and
the factors are recomputed), for a closed and a collinear walker set, in FP64 and mixed
precision, and for a walker set swapped for another between steps.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd.synthetic import make_problem  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import afq
    return afq, torch


def _run(afq, torch, p, W, closed, prec, keep, graph, seed=77, steps=11):
    old = os.environ.get("AFQ_LU_KEEP")
    os.environ["AFQ_LU_KEEP"] = "1" if keep else "0"
    try:
        c = afq.Context(0)
        if closed:
            ham = afq.HamiltonianOperations(c, p.h1, p.Likn, p.PsiT, p.nup, precision=prec)
            ws = afq.WalkerSet(p.walkers(W))
        else:
            nd = p.nup - 2
            ham = afq.HamiltonianOperations(c, p.h1, p.Likn, np.concatenate([p.PsiT, p.PsiT[:nd]]),
                                            p.nup, nd, walker_type=afq.AFQ_COLLINEAR,
                                            precision=prec)
            Pa = p.walkers(W)
            ws = afq.WalkerSet(Pa, np.ascontiguousarray(Pa[:, :, :nd]))
        prop = afq.Propagator(ham, W, dt=0.005, seed=seed)
        E = afq.empty((W, 3), torch.complex128)
        ov = afq.empty(W, torch.complex128)
        prop.Energy(ws, E, ov)
        ws.eloc.copy_(E.sum(1))
        ws.ovlp.copy_(ov)
        Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).real.mean())]))
        e_out = afq.empty(1, torch.float64)
        relax = 1 - np.exp(-0.1)
        for step in range(steps):
            ortho, pc = step % 5 == 0, step % 4 == 3
            if graph:
                prop.step_graph(ws, Esh, ortho=ortho, popctl=pc, relax=relax, energy_out=e_out)
            else:
                prop.Propagate(ws, Esh)
                if ortho:
                    prop.Orthogonalize(ws)
                if pc:
                    ctr = orc.philox4x32_10([0, 0, step, 1 + 1000], [seed & 0xffffffff, seed >> 32])
                    bits = (int(ctr[0]) << 32) | int(ctr[1])
                    prop.population_control(ws, ((bits >> 11) + 0.5) * 2.0 ** -53)
                afq.average_pseudo_energy(c, ws.weight, ws.eloc, 0.005, e_out, relax, Esh)
        torch.cuda.synchronize()
        out = [afq.host(ws.Phi), afq.host(ws.weight), afq.host(ws.ovlp), afq.host(Esh)]
        if not closed:
            out.append(afq.host(ws.Phi_b))
        prop.close()
        return out
    finally:
        if old is None:
            del os.environ["AFQ_LU_KEEP"]
        else:
            os.environ["AFQ_LU_KEEP"] = old


@pytest.mark.parametrize("closed", [True, False], ids=["closed", "collinear"])
@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_kept_factors_bitwise(gpu, closed, graph, mixed):
    afq, torch = gpu
    prec = afq.AFQ_MIXED if mixed else afq.AFQ_FP64
    say(f"case: kept overlap factors closed={closed} graph={graph} mixed={mixed} N=48 n=12 W=40")
    p = make_problem(48, 12, 160, 5, lscale=0.1)
    a = _run(afq, torch, p, 40, closed, prec, True, graph)
    b = _run(afq, torch, p, 40, closed, prec, False, graph)
    for x, y in zip(a, b):
        assert np.all(np.isfinite(x))
        assert np.array_equal(x, y)


def test_kept_factors_other_walker_set(gpu):
    """A second walker set stepped by the same propagator between steps of the first: every T
    differs from the kept ones, so nothing stale is taken."""
    afq, torch = gpu
    p = make_problem(40, 8, 120, 9, lscale=0.1)
    W = 24
    res = {}
    for keep in (True, False):
        old = os.environ.get("AFQ_LU_KEEP")
        os.environ["AFQ_LU_KEEP"] = "1" if keep else "0"
        try:
            c = afq.Context(0)
            ham = afq.HamiltonianOperations(c, p.h1, p.Likn, p.PsiT, p.nup)
            prop = afq.Propagator(ham, W, dt=0.005, seed=5)
            from safire_amd.synthetic import random_walkers
            sets = [afq.WalkerSet(random_walkers(p, W, seed=s, noise=0.2)) for s in (1, 2)]
            Esh = afq.device_tensor(np.array([0.0]))
            for ws in sets:
                E = afq.empty((W, 3), torch.complex128)
                ov = afq.empty(W, torch.complex128)
                prop.Energy(ws, E, ov)
                ws.eloc.copy_(E.sum(1))
                ws.ovlp.copy_(ov)
            for step in range(6):
                prop.Propagate(sets[step % 2], Esh)
            torch.cuda.synchronize()
            res[keep] = [afq.host(ws.Phi) for ws in sets] + [afq.host(ws.weight) for ws in sets]
            prop.close()
        finally:
            if old is None:
                del os.environ["AFQ_LU_KEEP"]
            else:
                os.environ["AFQ_LU_KEEP"] = old
    for x, y in zip(res[True], res[False]):
        assert np.all(np.isfinite(x))
        assert np.array_equal(x, y)
