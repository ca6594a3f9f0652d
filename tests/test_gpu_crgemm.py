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
