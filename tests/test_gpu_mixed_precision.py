"""Mixed precision against the FP64 path at the benchmark shape (N=200, n=40+40, M=1600).

The per-walker local-energy deviation after one step, with each FP32 stage alone (the force
bias vbias, the vHS contraction, the Taylor propagation) and all of them together
(bench.stage_errors), and the floor: the FP64 step's walker matrices rounded to complex64 once
(u/2 per element, u = 2^-24 = 6.0e-8).  Any FP32 computation of the propagated walkers rounds
them at least that much, so the floor bounds every FP32 propagation from below.

Bounds asserted:
* at the generator's default Cholesky amplitude 0.5 (local energies of ~5e4 Ha) the floor alone
  exceeds 0.1 mHa, so that bar is out of reach for any FP32 propagation there;
* at both amplitudes every FP32 stage, and the mixed step, stays within 100 times the floor
  (maximum over walkers of |dE| / |E|): an FP32 contraction over K terms carries ~sqrt(K) u
  relative error, sqrt(1600) = 40 for vHS, i.e. ~80 times the floor's u / 2 (measured: 20-60
  times at amplitude 0.5, 5-35 times at 0.1);
* at Cholesky amplitude 0.1 (local energies of ~600 Ha; the benchmark's workload) the mixed
  step meets 0.1 mHa per walker (measured 9.3e-5 Ha at most over 128 walkers).
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def budget():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import bench
    from safire_amd.synthetic import make_problem
    out = {}
    for amp in (0.5, 0.1):
        say(f"case: stage error budget N=200 n=40 M=1600 amplitude {amp}, 128 walkers")
        p = make_problem(200, 40, 1600, 1003, lscale=amp)
        out[amp] = bench.stage_errors(p, W_e=128)
        say(f"  {out[amp]}")
    return out


def test_floor_exceeds_the_bar_at_the_default_amplitude(budget):
    b = budget[0.5]
    assert b["floor"]["max_Ha"] > 1e-4
    assert b["floor"]["max_Ha"] <= b["mixed"]["max_Ha"]


@pytest.mark.parametrize("amp", [0.5, 0.1])
@pytest.mark.parametrize("stage", ["vbias", "vhs", "taylor", "mixed"])
def test_stage_errors_within_bound(budget, stage, amp):
    b = budget[amp]
    assert b[stage]["max_rel"] <= 100.0 * b["floor"]["max_rel"], (b[stage], b["floor"])


def test_mixed_meets_the_bar_at_physical_amplitude(budget):
    b = budget[0.1]
    assert b["mixed"]["max_Ha"] <= 1e-4, b["mixed"]
