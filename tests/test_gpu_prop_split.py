"""The FP64 fused propagation with the stripes past a multiple of four split by k over four
wavefronts (one split stripe: N = 200, 72; two: N = 90), through the C ABI, against the FP64
oracle's propagation per walker at 1e-12 of the walker's largest entry, with and without the split
(AFQ_PROP_SPLIT=0), and with a ragged last k group (N % 16 != 0) and a walker count that fills no
tile.
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


@pytest.mark.parametrize("N,n,W", [(200, 40, 6), (90, 20, 9), (72, 33, 5)])
def test_fp64_split_propagation(gpu, N, n, W):
    afq, ctx, torch = gpu
    say(f"case: FP64 propagation, split tail, N={N} n={n} W={W}")
    rng = np.random.default_rng(N + n)
    h = rng.normal(size=(N, N))
    h = 0.05 * (h + h.T)
    w, U = np.linalg.eigh(h)
    P1 = (U * np.exp(-w)) @ U.T + 0j
    V = rng.normal(size=(W, N, N)) + 1j * rng.normal(size=(W, N, N))
    V = 0.02 * (V + np.transpose(V, (0, 2, 1)))
    Phi = rng.normal(size=(W, N, n)) + 1j * rng.normal(size=(W, N, n))
    ref = orc.propagate(P1, V, Phi, order=6)
    out = {}
    for split in ("1", "0"):
        old = os.environ.get("AFQ_PROP_SPLIT")
        os.environ["AFQ_PROP_SPLIT"] = split
        try:
            dPhi = afq.device_tensor(Phi)
            afq.det_ops.Propagate(ctx, afq.device_tensor(P1), afq.device_tensor(V), dPhi, order=6,
                                  precision=afq.AFQ_FP64)
            out[split] = afq.host(dPhi)
        finally:
            if old is None:
                del os.environ["AFQ_PROP_SPLIT"]
            else:
                os.environ["AFQ_PROP_SPLIT"] = old
    for split, got in out.items():
        for k in range(W):
            e = np.abs(got[k] - ref[k]).max() / np.abs(ref[k]).max()
            assert e < 1e-12, (split, k, e)
    say(f"  split vs oracle {np.abs(out['1'] - ref).max() / np.abs(ref).max():.2e}, "
        f"unsplit {np.abs(out['0'] - ref).max() / np.abs(ref).max():.2e}")
