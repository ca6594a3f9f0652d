"""Continuous-batching serving loop for independent AFQMC jobs on one GPU.

Jobs (each its own hamiltonian, trial, walker count and seed) arrive over time.  A job is
admitted as soon as a slot is free and its HBM footprint fits the budget; it then gets its own
stream and library context, its tensors and walkers stay resident on the device, and it is
advanced in chunks of graph-replayed steps.  The host never waits for one job to finish before
issuing another job's work: each job's chunk completion is polled with an event, so the kernels
of all admitted jobs overlap on the device.  A job's result is its mixed-precision energy
estimate (mean and blocking error of the weighted local energies measured every
`measure_interval` steps after equilibration); time-to-energy is measured from its arrival.

Every job draws its fields and comb uniforms from Philox keyed by its own seed and all
reductions have a fixed order, so a job's result does not depend on what else is running.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import afq, driver
from .synthetic import make_problem


@dataclasses.dataclass
class JobSpec:
    id: int
    arrival: float          # seconds after the loop starts
    N: int
    nup: int
    M: int
    W: int
    seed: int
    steps: int = 400
    equilibration: int = 100
    measure_interval: int = 10
    ortho_interval: int = 5
    popctl_interval: int = 10
    dt: float = 0.005
    lscale: float = 0.1


@dataclasses.dataclass
class JobResult:
    id: int
    N: int
    W: int
    arrival: float
    admitted: float
    finished: float
    energy: float
    error: float
    walker_steps: int
    collapsed: bool

    @property
    def time_to_energy(self):
        return self.finished - self.arrival


def poisson_jobs(n_jobs=200, rate=4.0, seed=2024, steps=400, lscale=0.1):
    """The serving workload: Poisson arrivals, N uniform in {32, 48, ..., 160}, N/4 electrons per
    spin, M = 5N, 64-512 walkers, per-job seeds."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_jobs))
    t -= t[0]
    Ns = rng.choice(np.arange(32, 161, 16), size=n_jobs)
    Ws = rng.integers(64, 513, size=n_jobs)
    return [JobSpec(i, float(t[i]), int(Ns[i]), int(Ns[i]) // 4, 5 * int(Ns[i]), int(Ws[i]),
                    10_000 + i, steps=steps, lscale=lscale) for i in range(n_jobs)]


def footprint(spec: JobSpec) -> int:
    """Upper estimate of a job's device memory (bytes): resident tensors, walkers, step and
    energy scratch of a mixed-precision context."""
    N, n, M, W = spec.N, spec.nup, spec.M, spec.W
    ham = N * N * M * 12 + M * n * N * (16 + 12 + 8) + n * N * M * 16 + N * N * 64
    walkers = W * N * n * 16 + W * 5 * 16
    step = (W * n * N * 16 + W * M * 24 + W * N * N * 8 + 3 * W * N * n * 8
            + 4 * W * n * n * 16 + W * N * n * 16)
    per_m = W * n * n * 16
    mc = max(1, min(M, (1 << 30) // max(per_m, 1)))
    energy = per_m * mc + W * M * 16 + W * n * N * 24 + W * 64
    return int(1.15 * (ham + walkers + step + energy)) + (64 << 20)


class _Active:
    def __init__(self, spec: JobSpec, problem, precision, now):
        import torch
        self.spec = spec
        self.admitted = now
        self.stream = torch.cuda.Stream()
        with torch.cuda.stream(self.stream):
            self.ctx = afq.Context(0, stream=self.stream)
            self.ham = afq.HamiltonianOperations(self.ctx, problem.h1, problem.Likn, problem.PsiT,
                                                 problem.nup, precision=precision)
            self.prop = afq.Propagator(self.ham, spec.W, dt=spec.dt, seed=spec.seed)
            self.ws = afq.WalkerSet(problem.walkers(spec.W))
            self.Esh = driver.initialize(self.prop, self.ws)
            self.e_out = afq.empty(1, torch.float64)
            self.E = afq.empty((spec.W, 3), torch.complex128)
            self.ov = afq.empty(spec.W, torch.complex128)
        self.relax = 1.0 - np.exp(-10.0 / spec.equilibration)
        self.step = 0
        self.samples = []      # device scalars: weighted-average energy per measurement
        self.event = None

    def busy(self):
        return self.event is not None and not self.event.query()

    def advance(self, chunk):
        """Issue up to `chunk` steps (and the energy measurements due) on the job's stream."""
        import torch
        s = self.spec
        with torch.cuda.stream(self.stream):
            for _ in range(min(chunk, s.steps - self.step)):
                st = self.step
                self.prop.step_graph(self.ws, self.Esh, ortho=(st % s.ortho_interval == 0),
                                     popctl=(st % s.popctl_interval == 0),
                                     relax=(self.relax if st < s.equilibration else 0.0),
                                     energy_out=self.e_out)
                if st >= s.equilibration and st % s.measure_interval == 0:
                    self.prop.Energy(self.ws, self.E, self.ov, afq.AFQ_MIXED)
                    w = self.ws.weight
                    self.samples.append(((w[:, None] * self.E).sum() / w.sum()).reshape(1))
                self.step += 1
            self.event = torch.cuda.Event()
            self.event.record(self.stream)

    def done(self):
        return self.step >= self.spec.steps and not self.busy()

    def result(self, now):
        import torch
        with torch.cuda.stream(self.stream):
            e = torch.cat(self.samples).cpu().numpy() if self.samples else np.zeros(0, complex)
        collapsed = self.prop.collapsed()
        er = np.real(e)
        mean = float(er.mean()) if er.size else float("nan")
        err = driver.blocking_error(er) if er.size >= 8 else float("nan")
        s = self.spec
        return JobResult(s.id, s.N, s.W, s.arrival, self.admitted, now, mean, err,
                         s.W * s.steps, collapsed)

    def close(self):
        self.prop.close()
        self.ham.close()
        self.ctx.close()


class Server:
    def __init__(self, hbm_budget: int | None = None, max_active: int = 16, chunk: int = 10,
                 precision=afq.AFQ_MIXED):
        import torch
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.6 * free)
        self.budget = hbm_budget
        self.max_active = max_active
        self.chunk = chunk
        self.precision = precision

    def run(self, specs, problems=None, log=None):
        """Serves `specs` (arrival times relative to the start of the loop).  problems: optional
        dict id -> synthetic Problem (the request payloads); generated up front otherwise."""
        import torch
        if problems is None:
            problems = {s.id: make_problem(s.N, s.nup, s.M, s.seed, lscale=s.lscale) for s in specs}
        pending = sorted(specs, key=lambda s: s.arrival)
        active: list[_Active] = []
        results: list[JobResult] = []
        used = 0
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            # admission: arrived jobs, in arrival order, while a slot and the HBM budget allow
            while (pending and pending[0].arrival <= now and len(active) < self.max_active
                   and used + footprint(pending[0]) <= self.budget):
                spec = pending.pop(0)
                used += footprint(spec)
                active.append(_Active(spec, problems[spec.id], self.precision,
                                      time.perf_counter() - t0))
                if log:
                    log(f"t={now:7.2f}s admit job {spec.id} N={spec.N} W={spec.W} "
                        f"({len(active)} active, {used / 2**30:.1f} GiB)")
            progressed = False
            for job in list(active):
                if job.busy():
                    continue
                if job.done():
                    now = time.perf_counter() - t0
                    results.append(job.result(now))
                    used -= footprint(job.spec)
                    job.close()
                    active.remove(job)
                    progressed = True
                    if log:
                        r = results[-1]
                        log(f"t={now:7.2f}s job {r.id} done: E={r.energy:.5f} +- {r.error:.5f} "
                            f"tte={r.time_to_energy:.2f}s")
                    continue
                job.advance(self.chunk)
                progressed = True
            if not progressed:
                time.sleep(0.0005)
        torch.cuda.synchronize()
        return results


def summarize(results):
    tte = np.array([r.time_to_energy for r in results])
    first = min(r.arrival for r in results)
    last = max(r.finished for r in results)
    ws = sum(r.walker_steps for r in results)
    return {"jobs": len(results), "p50_time_to_energy_s": float(np.percentile(tte, 50)),
            "p99_time_to_energy_s": float(np.percentile(tte, 99)),
            "aggregate_walker_steps_per_s": ws / (last - first), "makespan_s": last - first,
            "collapsed_jobs": int(sum(r.collapsed for r in results))}
