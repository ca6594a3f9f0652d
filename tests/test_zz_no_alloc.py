"""The allocation-free density-matrix entry point against the allocating one, through the C ABI.

afq_mixed_density_matrix_noalloc takes its workspace from a fixed per-GPU array instead of the
context's scratch buffers and times no tile; its results must equal afq_mixed_density_matrix's
bit for bit (the same kernels in the same order).  Two calls in a row, the second smaller, show
that the shared workspace carries nothing from one call to the next; both the compact and the
full density matrix are checked, and a call above the workspace limit fails cleanly.  The
allocating entry point is itself checked against the oracle on the first input.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd.synthetic import make_problem, random_walkers  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import afq
    return afq, torch


def _run(afq, torch, ctx, PsiT, Phi, compact, noalloc):
    n, N = PsiT.shape
    W = Phi.shape[0]
    G = afq.empty((W, n, N) if compact else (W, N, N), torch.complex128)
    G.fill_(complex(7.0, -3.0))
    ov = afq.device_tensor(np.full(W, 0.25 - 0.5j))
    f = afq.det_ops.MixedDensityMatrixNoAlloc if noalloc else afq.det_ops.MixedDensityMatrix
    f(ctx, afq.device_tensor(PsiT), afq.device_tensor(Phi), G, ov, compact=compact)
    torch.cuda.synchronize()
    return afq.host(G), afq.host(ov)


@pytest.mark.parametrize("compact", [True, False])
def test_noalloc_equals_allocating_entry_point(env, compact):
    afq, torch = env
    ctx = afq.Context(0)
    # (N, n, W, seed), the second call smaller; the full density matrix's workspace,
    # W n (n + N) complex, fits the 32 MiB limit up to W = 218 at N = 200, n = 40
    cases = [(200, 40, 300 if compact else 200, 1003), (108, 24, 77, 1002)]
    for N, n, W, seed in cases:
        say(f"case: noalloc density matrix compact={compact} N={N} n={n} W={W}")
        p = make_problem(N, n, 8, seed)
        Phi = random_walkers(p, W, seed=seed + 5, noise=0.3)
        G0, ov0 = _run(afq, torch, ctx, p.PsiT, Phi, compact, noalloc=False)
        G1, ov1 = _run(afq, torch, ctx, p.PsiT, Phi, compact, noalloc=True)
        assert np.array_equal(G0, G1)
        assert np.array_equal(ov0, ov1)
        if N == 200:   # the allocating path against the oracle (1e-10 per walker)
            Gr, ovr = orc.mixed_density_matrix(p.PsiT, Phi[:8], compact=compact,
                                               ovlp=np.full(8, 0.25 - 0.5j))
            d = np.max(np.abs(G0[:8] - Gr).reshape(8, -1), axis=1)
            s = np.max(np.abs(Gr).reshape(8, -1), axis=1)
            assert np.max(d / s) < 1e-10
            assert np.max(np.abs(ov0[:8] - ovr) / np.abs(ovr)) < 1e-10


def test_noalloc_rejects_workspace_overflow(env):
    afq, torch = env
    from safire_amd.afq import AfqError
    ctx = afq.Context(0)
    N, n = 64, 64
    W = afq.NOALLOC_SCRATCH_BYTES // (16 * n * n) + 1
    say(f"case: noalloc overflow W={W} n={n}")
    PsiT = afq.device_tensor(np.eye(n, N, dtype=np.complex128))
    Phi = afq.empty((W, N, n), torch.complex128)
    G = afq.empty((W, n, N), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    with pytest.raises(AfqError, match="workspace"):
        afq.det_ops.MixedDensityMatrixNoAlloc(ctx, PsiT, Phi, G, ov, compact=True)
