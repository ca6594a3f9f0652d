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


def test_largest_serving_job_alone_equals_mix():
    """The serving workload's largest job (N=160, 40+40 electrons, M=800, 400 mixed steps, 256
    walkers, Cholesky amplitude 0.1 as every serving job) alone and among three other jobs of
    the workload's sizes arriving while it runs: the same energy bit for bit."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import serve
    from safire_amd.synthetic import make_problem
    sizes = [(160, 256), (32, 512), (96, 128), (128, 64)]
    specs = [serve.JobSpec(i, 0.05 * i, N, N // 4, 5 * N, W, 900 + i) for i, (N, W) in
             enumerate(sizes)]
    assert specs[0].steps == 400 and specs[0].lscale == 0.1
    problems = {s.id: make_problem(s.N, s.nup, s.M, s.seed, lscale=s.lscale) for s in specs}
    say("case: serving, N=160 M=800 400 steps alone vs with 3 jobs (N=32, 96, 128)")
    alone = serve.Server(max_active=1).run([specs[0]], problems)[0]
    mixed = {r.id: r for r in serve.Server(max_active=4, chunk=7).run(specs, problems)}
    m = mixed[0]
    say(f"N=160: alone {alone.energy:.10f} +- {alone.error:.4f}  mixed {m.energy:.10f}, "
        f"time to energy alone {alone.time_to_energy:.2f} s, in the mix {m.time_to_energy:.2f} s")
    assert np.isfinite(alone.energy) and not alone.collapsed
    assert alone.energy == m.energy
    for r in mixed.values():
        assert np.isfinite(r.energy) and not r.collapsed, r


def test_collapse_at_the_benchmark_amplitude_is_reported():
    """At the benchmark recipe's Cholesky amplitude 0.5 the N=160 job's fields rotate each
    walker's overlap phase by O(1) per step (||V||_2 ~ 3), the phaseless constraint zeroes about
    half of the weights every step, and the population dies between two population controls:
    the loop reports the job as collapsed with no energy instead of a number."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import serve
    from safire_amd.synthetic import make_problem
    spec = serve.JobSpec(0, 0.0, 160, 40, 800, 256, 900, lscale=0.5)
    say("case: serving, N=160 M=800 400 steps at amplitude 0.5")
    r = serve.Server(max_active=1).run([spec], {0: make_problem(160, 40, 800, 900, lscale=0.5)})[0]
    say(f"  collapsed {r.collapsed}, energy {r.energy}")
    assert r.collapsed and not np.isfinite(r.energy)
