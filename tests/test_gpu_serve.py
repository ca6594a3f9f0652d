"""Serving loop: jobs of different sizes admitted while others run give the same result alone and
in the mix (fixed per-job seeds, fixed-order reductions: bit-identical energies expected; the
stated tolerance is 1e-12 relative)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_alone_equals_mix():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import serve
    from safire_amd.synthetic import make_problem
    specs = [serve.JobSpec(i, 0.002 * i, N, N // 4, 5 * N, W, 500 + i, steps=60, equilibration=20,
                           measure_interval=2)
             for i, (N, W) in enumerate([(32, 64), (48, 96), (64, 72), (48, 128)])]
    problems = {s.id: make_problem(s.N, s.nup, s.M, s.seed, lscale=s.lscale) for s in specs}
    say("case: serving loop, 4 jobs N=32..64, alone vs mixed")
    alone = {}
    for s in specs:
        r = serve.Server(max_active=1).run([s], problems)
        alone[s.id] = r[0]
    mixed = {r.id: r for r in serve.Server(max_active=4, chunk=3).run(specs, problems)}
    for s in specs:
        a, m = alone[s.id], mixed[s.id]
        say(f"job {s.id}: alone {a.energy:.10f} +- {a.error:.4f}  mixed {m.energy:.10f}")
        assert np.isfinite(a.energy) and not a.collapsed
        assert abs(a.energy - m.energy) <= 1e-12 * abs(a.energy)
    summ = serve.summarize(list(mixed.values()))
    assert summ["jobs"] == 4 and summ["aggregate_walker_steps_per_s"] > 0
