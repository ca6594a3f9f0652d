"""Walker-step throughput on one MI355X.

Workload (the metric's configuration): N=200 orbitals, 40+40 electrons (closed shell), M=1600
Cholesky vectors, 1024 walkers, dt=0.005, 6-term Taylor propagator, QR every 5 steps, synthetic
hamiltonian (safire_amd.synthetic, seed 1003) with Cholesky amplitude 0.1.  A step is one hybrid
propagation of every walker (G, vbias, fields, vHS, P1 exp(iV) P1, overlap, weights) plus the
growth estimator and Eshift relaxation, with the QR every fifth step, replayed from a HIP graph.
`value` is the mixed precision path (FP32 MFMA contractions, FP64 G/LU/QR/weights); the FP64
path is timed beside it.  At the generator's default amplitude 0.5 the walkers' weights all
reach zero within about four steps (the FP64 oracle shows the same) and the matrices overflow,
so the timed steps would run on dead walkers and the mixed-vs-FP64 condition cannot be met;
those timings are reported beside the headline (`recipe_amplitude_0.5`).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-cpu] [--only MODE]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, NUP, M, W = 200, 40, 1600, 1024
DT, ORDER, ORTHO = 0.005, 6, 5
SEED = 1003
AMPLITUDE = 0.1          # Cholesky amplitude of the headline workload
RECIPE_AMPLITUDE = 0.5   # the generator's default
PEAK = {"fp32": 157.3e12, "fp64": 78.6e12}
HBM_PEAK = 8.0e12
# HBM bytes of one FP32 vHS launch (k_crgemm over the symmetric pairs) at this workload:
# FETCH_SIZE x 2 + WRITE_SIZE x 1 in KiB (the factors that make a 1 GiB copy read and write
# 1 GiB), rocprofv3 --pmc of `bench.py --only step`
# (profiles/step_N200_n40_M1600_W1024_pmc.md).  Algorithmic: X 13.1 MB + L 128.6 MB + V 327.7 MB.
VHS_TRAFFIC_BYTES = (226327 * 2 + 562803) * 1024
VHS_ALGORITHMIC_BYTES = W * M * 8 + (N * (N + 1) // 2) * M * 4 + W * N * N * 8


def flops_per_walker_step(N, n, M, K=ORDER, k_ortho=ORTHO, performed=True):
    """FLOP of one hybrid step of one walker.  performed=True counts the products the kernels
    issue for a real trial, real Cholesky vectors and a real one-body propagator: vbias and the
    two P1 products are complex x real (4 FLOP per MAC), the Taylor terms complex x complex
    (8), vHS the N(N+1)/2 distinct entries of the symmetric V.  performed=False is the count
    with every MAC complex (8 FLOP) and vHS over all N^2 entries (4 N^2 M), the formula the
    scope table states for the reference's operations."""
    g = 16 * n * n * N + 16 * n ** 3
    ov = 8 * n * n * N + (8 / 3) * n ** 3
    qr = 16 * n * n * N / k_ortho
    if performed:
        vb = 4 * M * n * N
        prop = 8 * K * N * N * n + 4 * 2 * N * N * n
        vhs = vhs_flops_per_walker(N, M)
    else:
        vb = 8 * M * n * N
        prop = 8 * (K + 2) * N * N * n
        vhs = 4.0 * N * N * M
    return g + vb + prop + ov + qr + vhs


def vhs_flops_per_walker(N, M):
    """vHS of one walker: every L^m is symmetric, so V is symmetric and its N(N+1)/2 distinct
    entries each take M complex x real MACs (4 FLOP each)."""
    return 2.0 * N * (N + 1) * M


def phmsd_flops_per_walker_step(N, M, spins, ndet, real_L=True, K=ORDER, k_ortho=ORTHO):
    """FLOP of one hybrid step of one COLLINEAR walker with the particle-hole trial, counted
    from the drawn excitation list.  spins: per spin (n electrons, nact active orbitals, ranks of
    the unique excited configurations).  Per spin, the density matrix (PHMSD::MixedDensityMatrix)
    costs the Woodbury set-up TAB = orbs Phi (nact n N complex MACs), the LU and inverse of its
    n x n reference block (n^3), QQ0 (nvir n^2) and GA (n^2 N); per excitation of rank r the
    LU of its r x r block, that block's inverse (r^2 (r - 1)) and its compact R rows (r^2 nact);
    the R assembly (r nact per excitation) and G = R^T GA (nact n N).  The step's closing
    Log_Overlap repeats the Woodbury set-up and the excitation LUs.  The spin-coupled weights
    take one complex MAC per determinant and spin, in both calls.  vbias contracts the nact x N
    G with the half-rotated Cholesky vectors (complex x real with a real trial: 4 FLOP per MAC),
    vHS, the Taylor propagation and the QR share (every k_ortho steps) are the single
    determinant's.  Complex MACs count 8 FLOP."""
    def lu(r):
        return sum((r - 1 - k) ** 2 for k in range(r))
    f = vhs_flops_per_walker(N, M) + 2 * 2 * 8 * ndet
    for n, na, ranks in spins:
        ranks = [int(r) for r in ranks if r > 0]
        nvir = na - n
        wood = 8 * (na * n * N + n ** 3 + nvir * n * n)
        mdm = (wood + 8 * n * n * N + 8 * sum(lu(r) + r * r * (r - 1) + r * r * na + r * na
                                              for r in ranks) + 8 * na * n * N)
        ovl = wood + 8 * sum(lu(r) for r in ranks)
        vb = (4 if real_L else 8) * M * na * N
        prop = 8 * K * N * N * n + 4 * 2 * N * N * n
        qr = 16 * n * n * N / k_ortho
        f += mdm + ovl + vb + prop + qr
    return f


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def gpu_setup(problem, precision, W):
    import torch
    from safire_amd import afq
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, problem.h1, problem.Likn, problem.PsiT, problem.nup,
                                    precision=precision)
    prop = afq.Propagator(ham, W, dt=DT, order=ORDER, seed=4242)
    ws = afq.WalkerSet(problem.walkers(W))
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    prop.Energy(ws, E, ov)
    ws.eloc.copy_(E.sum(1))
    ws.ovlp.copy_(ov)
    Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).real.mean())]))
    eout = afq.empty(1, torch.float64)
    return ctx, ham, prop, ws, Esh, eout


def time_steps(prop, ws, Esh, eout, warmup, steps):
    import torch
    relax = 1.0 - np.exp(-10.0 / 100.0)
    for i in range(warmup):
        prop.step_graph(ws, Esh, ortho=(i % ORTHO == 0), relax=relax, energy_out=eout)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(warmup, warmup + steps):
        prop.step_graph(ws, Esh, ortho=(i % ORTHO == 0), relax=relax, energy_out=eout)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def time_vhs_in_step(prop, ws, Esh, eout, warmup=3, reps=10):
    """Average duration of the vHS launch as the step runs it: HIP events recorded around it
    inside the step graph, read after each of `reps` replayed steps (after the timed region)."""
    relax = 1.0 - np.exp(-10.0 / 100.0)
    prop.vhs_timing(True)
    try:
        for i in range(warmup):
            prop.step_graph(ws, Esh, ortho=(i % ORTHO == 0), relax=relax, energy_out=eout)
        tot = 0.0
        for i in range(reps):
            prop.step_graph(ws, Esh, ortho=(i % ORTHO == 0), relax=relax, energy_out=eout)
            tot += prop.vhs_ms()
    finally:
        prop.vhs_timing(False)
    return tot / reps / 1e3


def time_vhs_in_eager_step(prop, ws, Esh, warmup=3, reps=10):
    """Average duration of the vHS launch inside eagerly launched steps: HIP events recorded on
    the library's stream right before and after the launch (no graph event nodes)."""
    prop.vhs_timing(True)
    try:
        for _ in range(warmup):
            prop.Propagate(ws, Esh)
        tot = 0.0
        for _ in range(reps):
            prop.Propagate(ws, Esh)
            tot += prop.vhs_ms()
    finally:
        prop.vhs_timing(False)
    return tot / reps / 1e3


def time_vhs(ham, precision, W, reps=10):
    """Average duration of one vHS GEMM launch (the dominant kernel), HIP events on the stream
    the library launches on (the shared torch stream)."""
    import torch
    from safire_amd import afq
    cdt = torch.complex64 if precision == afq.AFQ_MIXED else torch.complex128
    X = (torch.randn(W, M, dtype=torch.float64, device="cuda") +
         1j * torch.randn(W, M, dtype=torch.float64, device="cuda")).to(cdt).contiguous()
    V = torch.empty((W, N, N), dtype=cdt, device="cuda")
    ham.vHS(X, V, DT, precision=precision)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(reps):
        ham.vHS(X, V, DT, precision=precision)
    ev1.record()
    ev1.synchronize()
    return ev0.elapsed_time(ev1) / 1e3 / reps


def time_energy(prop, ws, precision, reps=3):
    """Seconds per local-energy evaluation of the whole walker set (Energy: G, E1, exchange and
    Coulomb terms), timed around synchronized calls after one warm-up."""
    import torch
    from safire_amd import afq
    Wn = ws.W
    E = afq.empty((Wn, 3), torch.complex128)
    ov = afq.empty(Wn, torch.complex128)
    prop.Energy(ws, E, ov, precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        prop.Energy(ws, E, ov, precision)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def bh_mixed_error(W_e=256, steps=10):
    """Mixed vs FP64 on BH read from the reference's inputs (19 orbitals, 78 Cholesky vectors,
    RHF trial): the same walkers and field draws propagated `steps` steps in each precision, then
    the per-walker local energies (evaluated in FP64) compared, in Hartree."""
    import torch
    from safire_amd import afq, h5inputs
    g = os.path.join(ROOT, "tests", "golden", "bh")
    s = h5inputs.load_system(os.path.join(g, "afqmc_H_rhf_closed.h5"),
                             os.path.join(g, "afqmc_rhf_nomsd.h5"))
    res = {}
    for prec in (afq.AFQ_FP64, afq.AFQ_MIXED):
        ctx = afq.Context(0)
        ham = afq.HamiltonianOperations(ctx, s.h1, s.Likn, s.PsiT, s.nup, E0=s.E0, precision=prec)
        prop = afq.Propagator(ham, W_e, dt=0.01, seed=42)
        ws = afq.WalkerSet(s.walkers(W_e)[0])
        E = afq.empty((W_e, 3), torch.complex128)
        ov = afq.empty(W_e, torch.complex128)
        prop.Energy(ws, E, ov)
        ws.eloc.copy_(E.sum(1))
        ws.ovlp.copy_(ov)
        Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).real.mean())]))
        for k in range(steps):
            prop.set_step(k)
            prop.Propagate(ws, Esh)
        prop.Energy(ws, E, ov)
        res[prec] = afq.host(E.sum(1))
        prop.close()
    d = np.abs(res[afq.AFQ_MIXED] - res[afq.AFQ_FP64])
    return {"system": "BH (19 orbitals, 78 Cholesky vectors, 3+3 electrons, RHF trial) from the "
                      "reference's test inputs", "walkers": W_e, "steps": steps,
            "max_Ha": float(d.max()), "median_Ha": float(np.median(d)),
            "mean_energy_diff_Ha": float(abs(res[afq.AFQ_MIXED].mean() - res[afq.AFQ_FP64].mean()))}


def estimator_bench(reps=3):
    """Cost of one back-propagated and one mixed measurement (OneRDM + DiagonalTwoRDM) on BH read
    from the reference's inputs at its run parameters (1536 walkers, dt 0.01, 100-step
    back-propagation with QR every 10 steps): the RHF trial (one reference) and the CASCI
    PHMSD trial (44 references); and the time-evolved estimator over the same 100 steps (the
    operator propagation once, the dressed measurement)."""
    import torch
    from safire_amd import afq, h5inputs
    from safire_amd.phmsd import PHMSD
    g = os.path.join(ROOT, "tests", "golden", "bh")
    W, L = 1536, 100
    res = {}
    for name, wfn in (("rhf", "afqmc_rhf_nomsd.h5"), ("casci_phmsd", "afqmc_casci_rhf_phmsd.h5")):
        s = h5inputs.load_system(os.path.join(g, "afqmc_H_rhf_closed.h5"), os.path.join(g, wfn),
                                 walker_type=None)
        ctx = afq.Context(0)
        trial = None
        if s.ph is None:
            ham = afq.HamiltonianOperations(ctx, s.h1, s.Likn, s.PsiT, s.nup, s.ndown,
                                            walker_type=s.walker_type, E0=s.E0)
            ws = afq.WalkerSet(*s.walkers(W))
        else:
            ham = afq.HamiltonianOperations(ctx, s.h1, s.Likn, s.PsiT, s.nact, s.nact,
                                            walker_type=afq.AFQ_COLLINEAR, E0=s.E0)
            trial = PHMSD(ham, s.ph)
            up = np.ascontiguousarray(s.PsiT[:s.nup].conj().T)
            Pa = np.broadcast_to(up, (W,) + up.shape).copy()
            ws = afq.WalkerSet(Pa, Pa.copy())
        prop = afq.Propagator(ham, W, dt=0.01, seed=42, trial=trial)
        prop.enable_backprop(L)
        E = afq.empty((W, 3), torch.complex128)
        ov = afq.empty(W, torch.complex128)
        prop.Energy(ws, E, ov)
        ws.eloc.copy_(E.sum(1))
        ws.ovlp.copy_(ov)
        Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).real.mean())]))
        prop.bp_anchor(ws)
        prop.te_anchor()
        t0 = time.perf_counter()
        for k in range(L):
            prop.step_graph(ws, Esh, ortho=(k % 10 == 0), popctl=(k % 10 == 0))
        torch.cuda.synchronize()
        t_step = (time.perf_counter() - t0) / L
        # time-evolved operators over the same window: one PropagateOperators call over the
        # 100 stored fields (timed on a second call, after the first has tuned its products),
        # then one dressed measurement
        prop.te_propagate(ws, L)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        prop.te_propagate(ws, L)
        torch.cuda.synchronize()
        t_tep = time.perf_counter() - t0
        prop.te_measure(ws, L)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            prop.te_measure(ws, L)
        torch.cuda.synchronize()
        t_tem = (time.perf_counter() - t0) / reps
        prop.bp_measure(ws, L, 10, 1)
        prop.mixed_measure(ws)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            prop.bp_measure(ws, L, 10, 1)
        torch.cuda.synchronize()
        t_bp = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            prop.mixed_measure(ws)
        torch.cuda.synchronize()
        t_mx = (time.perf_counter() - t0) / reps
        res[name] = {"references": prop.number_of_references(), "walkers": W,
                     "step_ms": t_step * 1e3, "backprop_100_ms": t_bp * 1e3,
                     "mixed_ms": t_mx * 1e3,
                     "time_evolved_propagate_100_ms": t_tep * 1e3,
                     "time_evolved_measure_ms": t_tem * 1e3,
                     "backprop_share_at_one_per_100_steps": t_bp / (t_bp + 100 * t_step)}
        prop.close()
        if trial is not None:
            trial.close()
    return res


def energy_error(problem, W_e=64):
    """Per-walker local energy of the mixed path vs the FP64 path after one step from the same
    state and the same field draws (energies evaluated in FP64)."""
    import torch
    from safire_amd import afq
    res = {}
    for prec in (afq.AFQ_FP64, afq.AFQ_MIXED):
        ctx, ham, prop, ws, Esh, eout = gpu_setup(problem, prec, W_e)
        prop.set_step(0)
        prop.Propagate(ws, Esh)
        E = afq.empty((W_e, 3), torch.complex128)
        ov = afq.empty(W_e, torch.complex128)
        prop.Energy(ws, E, ov)
        res[prec] = afq.host(E.sum(1))
    d = np.abs(res[afq.AFQ_MIXED] - res[afq.AFQ_FP64])
    return float(d.max()), float(np.median(d)), float((d / np.abs(res[afq.AFQ_FP64])).max()), \
        float(np.median(np.abs(res[afq.AFQ_FP64])))


def config_energy_errors(W_e=64):
    """Per-walker local-energy deviation of the mixed step from the FP64 step (one step from the
    same state and field draws, energies in FP64) for each configuration's system: the CPU case
    (N=24), the FP64 GPU case (N=108), the benchmark shape at the recipe's Cholesky amplitude 0.5
    and at 0.1, and the smallest and largest serving jobs (amplitude 0.1)."""
    cases = {"N=24 n=5 M=96": (24, 5, 96, 1001, 0.5),
             "N=108 n=24 M=600": (108, 24, 600, 1002, 0.5),
             "N=200 n=40 M=1600": (N, NUP, M, SEED, 0.5),
             "N=200 n=40 M=1600 amplitude 0.1": (N, NUP, M, SEED, 0.1),
             "serving job N=32 n=8 M=160": (32, 8, 160, 10_000, 0.1),
             "serving job N=160 n=40 M=800": (160, 40, 800, 10_001, 0.1)}
    from safire_amd.synthetic import make_problem
    out = {}
    for name, (n_, nup, m_, seed, amp) in cases.items():
        emax, emed, erel, emag = energy_error(make_problem(n_, nup, m_, seed, lscale=amp), W_e)
        out[name] = {"amplitude": amp, "max_Ha": emax, "median_Ha": emed, "max_rel": erel,
                     "median_abs_E_Ha": emag}
    return out


def stage_errors(problem, W_e=128, seed=29, steps=1):
    """Per-stage error budget of the mixed-precision step against the FP64 step from the same
    walkers and the same field uniforms.  Each variant runs `steps` steps with one stage in FP32
    (the force-bias contraction vbias, the vHS contraction, or the Taylor propagation) and the
    rest in FP64; 'mixed' runs all three in FP32 (the step's kernels); 'floor' is the FP64 step
    with the propagated walker matrices rounded to complex64 once, the smallest perturbation any
    FP32 computation of them can make.  Local energies are evaluated in FP64 from FP64 G.
    Returns per-variant max / median |dE| (Ha) and max |dE|/|E| over the walkers."""
    import torch
    from safire_amd import afq
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, problem.h1, problem.Likn, problem.PsiT, problem.nup,
                                    precision=afq.AFQ_MIXED)
    prop = afq.Propagator(ham, W_e, dt=DT, order=ORDER)
    P1 = prop.P1()
    vmf = prop.vMF()
    Nn, n, Mm = problem.N, problem.nup, problem.M
    dP = afq.device_tensor(problem.PsiT)
    Phi0 = afq.device_tensor(problem.walkers(W_e))
    c64, c128 = torch.complex64, torch.complex128

    def energy(Phi):
        G = afq.empty((W_e, n, Nn), c128)
        ov = afq.zeros(W_e, c128)
        afq.det_ops.MixedDensityMatrix(ctx, dP, Phi, G, ov, compact=True)
        E = afq.empty((W_e, 3), c128)
        ham.energy(E, G, precision=afq.AFQ_FP64)
        return afq.host(E.sum(1))

    def run(fp32_vbias, fp32_vhs, fp32_taylor, round_out=False):
        Phi = Phi0.clone()
        for k in range(steps):
            G = afq.empty((W_e, n, Nn), c128)
            ov = afq.zeros(W_e, c128)
            afq.det_ops.MixedDensityMatrix(ctx, dP, Phi, G, ov, compact=True)
            X = afq.empty((W_e, Mm), c128)
            ham.vbias(G, X, DT, precision=afq.AFQ_MIXED if fp32_vbias else afq.AFQ_FP64)
            u = afq.empty((W_e, Mm), torch.float64)
            afq.field_uniforms(ctx, seed, 1, k, u)
            HW, mf = afq.empty(W_e, c128), afq.empty(W_e, c128)
            afq.construct_x(ctx, u, vmf, X, HW, mf)
            if fp32_vhs:
                V32 = afq.empty((W_e, Nn, Nn), c64)
                ham.vHS(X.to(c64), V32, DT, precision=afq.AFQ_MIXED)
                V = V32.to(c128)
            else:
                V = afq.empty((W_e, Nn, Nn), c128)
                ham.vHS(X, V, DT, precision=afq.AFQ_FP64)
            if fp32_taylor:
                afq.det_ops.Propagate(ctx, P1, V.to(c64).contiguous(), Phi, order=ORDER,
                                      precision=afq.AFQ_MIXED)
            else:
                afq.det_ops.Propagate(ctx, P1, V, Phi, order=ORDER, precision=afq.AFQ_FP64)
        if round_out:
            Phi = Phi.to(c64).to(c128)
        return energy(Phi)

    ref = run(False, False, False)
    out = {"walkers": W_e, "steps": steps, "median_abs_E_Ha": float(np.median(np.abs(ref)))}
    for name, flags in (("floor", (False, False, False, True)), ("vbias", (True, False, False)),
                        ("vhs", (False, True, False)), ("taylor", (False, False, True)),
                        ("mixed", (True, True, True))):
        d = np.abs(run(*flags) - ref)
        out[name] = {"max_Ha": float(d.max()), "median_Ha": float(np.median(d)),
                     "max_rel": float((d / np.abs(ref)).max())}
    prop.close()
    ham.close()
    return out


def cpu_baseline(problem, budget_s=12.0, warmup=3):
    """The FP64 oracle (CPU restatement of the reference step, OpenMP over walkers, its own
    loops with no BLAS) on a bounded sample: a few walkers of the same workload, `warmup`
    untimed steps, then steps timed one by one for about budget_s seconds; the value is the
    median of the per-step rates (QR steps included in their step's time)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    lib = orc.lib()
    threads = lib.orc_num_threads()
    osys = orc.OracleSystem(problem, dt=DT)
    Ws = max(threads, 8)
    st, Eshift = osys.init_walkers(problem.walkers(Ws))
    rates, t_tot, step = [], 0.0, 0
    while (t_tot < budget_s or len(rates) < 3) and step < 1000:
        u = orc.field_uniforms(7, 1, step, Ws, problem.M)
        t0 = time.perf_counter()
        osys.step(st, u, Eshift)
        if step % ORTHO == 0:
            osys.orthogonalize(st)
        dt = time.perf_counter() - t0
        if step >= warmup:
            t_tot += dt
            rates.append(Ws / dt)
        step += 1
    return {"value": float(np.median(rates)), "unit": "walker-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{Ws} walkers of the same N=200/M=1600 workload, FP64 oracle (its own "
                      f"loops, no BLAS, OpenMP over walkers on {threads} threads): median of "
                      f"{len(rates)} timed steps ({t_tot:.1f} s) after {warmup} untimed steps"}


def walker_state(ws):
    """Finiteness and survival of a walker set: the fraction of walkers whose matrices are
    finite and whose weights are finite and non-zero."""
    import torch
    from safire_amd import afq
    phi_ok = torch.isfinite(torch.view_as_real(ws.Phi)).flatten(1).all(1)
    w = ws.weight
    w_ok = torch.isfinite(torch.view_as_real(w)).all(-1) & (w.abs() > 0)
    return {"finite_matrices": float(phi_ok.double().mean()),
            "alive_finite_weights": float((phi_ok & w_ok).double().mean())}


def recipe_workload(warmup, steps, amplitude=RECIPE_AMPLITUDE):
    """The benchmark shapes at the generator's default Cholesky amplitude 0.5: mixed and FP64
    step times and the state of the walkers after the timed steps.  The propagator's norm grows
    the walker matrices by 1e5-1e9 per step and every walker's weight reaches zero within about
    four steps (the FP64 oracle shows the same), so these steps run on dead, overflowing walkers;
    the one-step mixed-vs-FP64 energy error from the initial walkers is in
    `energy_error_mixed_vs_fp64_per_config`."""
    from safire_amd.synthetic import make_problem
    import torch
    pr = make_problem(N, NUP, M, SEED, lscale=amplitude)
    res = {"amplitude": amplitude}
    from safire_amd import afq
    for prec, name in ((afq.AFQ_MIXED, "mixed"), (afq.AFQ_FP64, "fp64")):
        ctx, ham, prop, ws, Esh, eout = gpu_setup(pr, prec, W)
        sec = time_steps(prop, ws, Esh, eout, warmup, steps)
        res[name] = {"ms_per_step": sec * 1e3, "value": W / sec, "unit": "walker-steps/s",
                     "walkers_after_timed_steps": walker_state(ws)}
        del ctx, ham, prop, ws
        torch.cuda.empty_cache()
    return res


def phmsd_bench(warmup, steps, energy_reps=5):
    """The multi-determinant workload: N=64, 10+10 electrons, M=320, 2000 particle-hole
    determinants (ranks <= 4, 15 active orbitals per spin), 256 COLLINEAR walkers in FP64.
    Walkers start from the reference determinant with a small seeded perturbation (exactly at
    the reference every excitation overlap vanishes and the Wick energy terms cancel
    catastrophically). Returns step and energy timings."""
    import torch
    from safire_amd import afq
    from safire_amd.phmsd import PHMSD, build_ph_struct
    from safire_amd.synthetic import make_phmsd_problem
    Wp = 256
    P = make_phmsd_problem()
    ph = build_ph_struct(P.occa, P.occb, P.ci)
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, P.h1, P.Likn, P.PsiT, P.nact, P.nact,
                                    walker_type=afq.AFQ_COLLINEAR)
    trial = PHMSD(ham, ph)
    prop = afq.Propagator(ham, Wp, dt=DT, order=ORDER, seed=4343, trial=trial)
    ws = afq.WalkerSet(*P.walkers(Wp, seed=11, noise=0.05))
    E = afq.empty((Wp, 3), torch.complex128)
    ov = afq.empty(Wp, torch.complex128)
    prop.Energy(ws, E, ov)
    ws.eloc.copy_(E.sum(1))
    ws.ovlp.copy_(ov)
    Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).real.mean())]))
    eout = afq.empty(1, torch.float64)
    sec = time_steps(prop, ws, Esh, eout, warmup, steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(energy_reps):
        prop.Energy(ws, E, ov)
    torch.cuda.synchronize()
    te = (time.perf_counter() - t0) / energy_reps
    Emean = float(afq.host(E.sum(1)).real.mean())
    fstep = phmsd_flops_per_walker_step(
        P.N, P.M, [(P.nup, P.nact, ph.rank[0]), (P.ndn, P.nact, ph.rank[1])], ph.ndet,
        real_L=not np.iscomplexobj(P.orbs) or not np.any(np.imag(P.orbs)))
    return {"workload": "N=64, 10+10 electrons, M=320, 2000 determinants (ranks <= 4, "
                        f"{ph.nconf[0]}+{ph.nconf[1]} unique configurations), 256 walkers, FP64",
            "value": Wp / sec, "unit": "walker-steps/s", "ms_per_step": sec * 1e3,
            "energy_ms": te * 1e3, "mean_local_energy": Emean,
            # FLOP of the step from the drawn excitation list, over the FP64 MFMA peak
            "step_flop_per_walker": fstep,
            "step_roofline_frac": fstep * Wp / sec / PEAK["fp64"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--only", choices=["vhs", "vhs64", "copy", "step", "step64", "phmsd", "vendor",
                                       "estimators", "energy", "serve", "recipe"],
                    default=None,
                    help="profiling modes: launch only the named kernel(s)")
    ap.add_argument("--serve-jobs", type=int, default=200,
                    help="jobs of the serving run (Poisson arrivals at 4/s; 200 = the full "
                         "serving workload; 0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(walkers shard across GPUs; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    from safire_amd import afq
    from safire_amd.synthetic import make_problem
    log(f"generating workload N={N} n={NUP} M={M} W={W}")
    problem = make_problem(N, NUP, M, SEED, lscale=AMPLITUDE)

    if args.only in ("vhs", "vhs64"):
        prec = afq.AFQ_MIXED if args.only == "vhs" else afq.AFQ_FP64
        ctx = afq.Context(0)
        ham = afq.HamiltonianOperations(ctx, problem.h1, problem.Likn, problem.PsiT, NUP,
                                        precision=prec)
        t = time_vhs(ham, prec, W, reps=args.steps)
        log(f"vHS launch {t * 1e3:.3f} ms")
        return
    if args.only == "energy":
        for prec, name in ((afq.AFQ_MIXED, "mixed"), (afq.AFQ_FP64, "fp64")):
            ctx, ham, prop, ws, Esh, eout = gpu_setup(problem, prec, W)
            te = time_energy(prop, ws, prec, reps=args.steps)
            log(f"[{name}] energy {te * 1e3:.2f} ms, exchange "
                f"{4.0 * NUP * NUP * N * M * W / te / 1e12:.1f} TFLOP/s")
            del ctx, ham, prop, ws
            torch.cuda.empty_cache()
        return
    if args.only == "estimators":
        log("back-propagation and mixed estimators on BH")
        log(json.dumps(estimator_bench()))
        return
    if args.only == "phmsd":
        log(f"PHMSD workload: {args.steps} steps after {args.warmup} warmup")
        log(json.dumps(phmsd_bench(args.warmup, args.steps)))
        return
    if args.only == "vendor":
        # practical ceiling: the vendor GEMM (torch.matmul -> hipBLASLt / rocBLAS) on real
        # operands of the vHS and vbias shapes and on a large cube
        for dt in (torch.float32, torch.float64):
            for (m, k, n) in ((2048, M, N * (N + 1) // 2), (W, NUP * N, 2 * M), (8192, 8192, 8192)):
                a = torch.randn(m, k, dtype=dt, device="cuda")
                b = torch.randn(k, n, dtype=dt, device="cuda")
                a @ b
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(5):
                    a @ b
                torch.cuda.synchronize()
                t = (time.perf_counter() - t0) / 5
                log(f"{str(dt)[6:]} {m}x{k}x{n}: {t * 1e3:.3f} ms {2 * m * k * n / t / 1e12:.1f} TFLOP/s")
                del a, b
        return
    if args.only == "copy":
        # calibration kernel of known traffic for the FETCH/WRITE counters: copy 1 GiB
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 1 GiB x %d" % args.steps)
        return

    out = {}
    for prec, name in ((afq.AFQ_MIXED, "mixed"), (afq.AFQ_FP64, "fp64")):
        if args.only in ("serve", "recipe"):
            break
        if (name == "fp64" and args.skip_fp64) or (name == "mixed" and args.only == "step64"):
            continue
        log(f"[{name}] setup")
        ctx, ham, prop, ws, Esh, eout = gpu_setup(problem, prec, W)
        log(f"[{name}] timing {args.steps} steps after {args.warmup} warmup")
        sec = time_steps(prop, ws, Esh, eout, args.warmup, args.steps)
        state = walker_state(ws)
        tvs = time_vhs_in_step(prop, ws, Esh, eout)
        tve = time_vhs_in_eager_step(prop, ws, Esh)
        tv = time_vhs(ham, prec, W)
        te = time_energy(prop, ws, prec) if args.only is None else None
        out[name] = dict(sec=sec, vhs=tvs, vhs_eager=tve, vhs_b2b=tv, energy=te, state=state)
        log(f"[{name}] walkers after the timed steps: {state}")
        log(f"[{name}] {sec * 1e3:.2f} ms/step, {W / sec:.4g} walker-steps/s, vHS {tvs * 1e3:.3f} ms "
            f"in the step graph, {tve * 1e3:.3f} ms in eager steps ({tv * 1e3:.3f} ms back to back)"
            + (f", energy {te * 1e3:.2f} ms" if te else ""))
        del ctx, ham, prop, ws
        torch.cuda.empty_cache()
        if args.only == "step":
            break

    shard = None
    if args.only is None:
        # one GPU's share of an 8-GPU node: the same workload with W / 8 walkers per GPU
        Ws = W // 8
        log(f"[mixed] {Ws} walkers (one GPU's share of {W} over 8 GPUs)")
        ctx, ham, prop, ws, Esh, eout = gpu_setup(problem, afq.AFQ_MIXED, Ws)
        ss = time_steps(prop, ws, Esh, eout, args.warmup, args.steps)
        shard = {"walkers": Ws, "ms_per_step": ss * 1e3, "value": Ws / ss,
                 "unit": "walker-steps/s", "node_of_8_if_linear": 8 * Ws / ss}
        log(f"[mixed] W={Ws}: {ss * 1e3:.3f} ms/step, {Ws / ss:.4g} walker-steps/s")
        del ctx, ham, prop, ws
        torch.cuda.empty_cache()

    recipe = None
    if args.only in (None, "recipe"):
        log(f"recipe amplitude {RECIPE_AMPLITUDE}: timing")
        recipe = recipe_workload(args.warmup, args.steps)
        log(f"recipe amplitude {RECIPE_AMPLITUDE}: {recipe}")
        if args.only == "recipe":
            return

    if args.only != "serve":
        log("autotuned tiles:\n" + afq.tuning_report())
    if args.only == "step64":
        return
    emax = emed = erel = emag = None
    bh_err = None
    cfg_err = budget = None
    if args.only is None:
        log("energy error mixed vs fp64 (one step, 64 walkers)")
        emax, emed, erel, emag = energy_error(problem)
        log("energy error mixed vs fp64 on BH (10 steps, 256 walkers)")
        bh_err = bh_mixed_error()
        log(f"bh: {bh_err}")
        log("energy error mixed vs fp64 per configuration (one step, 64 walkers)")
        cfg_err = config_energy_errors()
        log(f"per configuration: {cfg_err}")
        log("stage error budget at the benchmark shape (one step, 128 walkers)")
        budget = {}
        for amp in (0.5, 0.1):
            b = stage_errors(make_problem(N, NUP, M, SEED, lscale=amp), W_e=128)
            budget[f"amplitude_{amp}"] = {k: b[k] for k in ("floor", "vbias", "vhs", "taylor",
                                                            "mixed")}
            budget[f"amplitude_{amp}"]["median_abs_E_Ha"] = b["median_abs_E_Ha"]
        log(f"budget: {budget}")

    phmsd = None
    if args.only is None:
        log("PHMSD workload (2000 determinants, 256 walkers, FP64)")
        phmsd = phmsd_bench(args.warmup, min(args.steps, 30))
        log(f"phmsd: {phmsd}")

    serve_summary = None
    if args.only in (None, "serve") and args.serve_jobs > 0:
        from safire_amd import serve
        log(f"serving run: {args.serve_jobs} jobs, Poisson arrivals at 4 jobs/s, 400 mixed steps each")
        specs = serve.poisson_jobs(args.serve_jobs, rate=4.0)
        problems = {}
        for i, sp in enumerate(specs):
            problems[sp.id] = make_problem(sp.N, sp.nup, sp.M, sp.seed, lscale=sp.lscale)
            if (i + 1) % 50 == 0:
                log(f"  generated {i + 1} job inputs")
        res = serve.Server(max_active=16).run(
            specs, problems, log=log if args.serve_jobs > 50 else None)
        serve_summary = serve.summarize(res)
        serve_summary.update({"arrival_rate_per_s": 4.0, "steps_per_job": 400,
                              "workload": "N in {32..160 step 16}, n=N/4, M=5N, W in [64,512], "
                                          "Cholesky amplitude 0.1, mixed precision"})
        log(f"serve: {serve_summary}")
        if args.only == "serve":
            print(json.dumps(serve_summary))
            return

    cpu = None
    if not args.skip_cpu and args.only is None:
        log("cpu baseline (FP64 oracle)")
        cpu = cpu_baseline(problem)

    sec = out["mixed"]["sec"]
    fstep = flops_per_walker_step(N, NUP, M)
    fstep_all = flops_per_walker_step(N, NUP, M, performed=False)
    vhs_flops = vhs_flops_per_walker(N, M) * W
    # achieved: the vHS launch's duration inside ten eagerly launched steps, HIP events recorded
    # on the library's stream right around it, which a kernel trace of the steps reproduces
    # launch for launch; beside it the average of ten back-to-back launches of the kernel alone
    # and the launch inside the step graph (event nodes in the graph add their own cost)
    tv = out["mixed"]["vhs_eager"]
    achieved = vhs_flops / tv
    line = {
        "metric": "walker-steps/s @N=200,M=1600: mixed vs FP64 at <=0.1 mHa; serve p50/p99",
        "value": W / sec,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 FP32-MFMA contractions + c128 G/LU/QR/weights",
        "data": "synthetic (seeded hamiltonian, trial, fields)",
        "config": {"workload": "N=200 orbitals, 40+40 electrons, M=1600 Cholesky vectors, "
                               "1024 walkers, dt=0.005, taylor 6, QR every 5 steps",
                   "cholesky_amplitude": AMPLITUDE,
                   "precision": "mixed", "graph": "one HIP graph per step"},
        "roofline": {"bound": "mfma",
                     "kernel": "vHS: k_crgemm, register-tiled FP32 MFMA, X [W x M] . the "
                               "N(N+1)/2 symmetric pairs of L [M], V mirrored",
                     "achieved": achieved / 1e12, "peak": PEAK["fp32"] / 1e12,
                     "vhs_ms_in_eager_steps": tv * 1e3,
                     "vhs_ms_back_to_back": out["mixed"]["vhs_b2b"] * 1e3,
                     "vhs_ms_in_step_graph_events": out["mixed"]["vhs"] * 1e3,
                     "achieved_in_step_graph_events": vhs_flops / out["mixed"]["vhs"] / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK["fp32"],
                     "traffic": VHS_TRAFFIC_BYTES,
                     "traffic_over_algorithmic": VHS_TRAFFIC_BYTES / VHS_ALGORITHMIC_BYTES},
        # the whole step: the FLOP the kernels issue (flops_per_walker_step) per second over the
        # MFMA peak of the contraction precision; *_all_complex: the count with every MAC
        # complex and vHS over all N^2 entries
        "step_flop_per_walker": fstep,
        "step_roofline_frac": (fstep * W / sec) / PEAK["fp32"],
        "step_roofline_frac_all_complex_count": (fstep_all * W / sec) / PEAK["fp32"],
        "fp64": None if "fp64" not in out else {
            "value": W / out["fp64"]["sec"], "ms_per_step": out["fp64"]["sec"] * 1e3,
            "vhs_ms": out["fp64"]["vhs_eager"] * 1e3,
            "vhs_ms_back_to_back": out["fp64"]["vhs_b2b"] * 1e3,
            "step_roofline_frac": (fstep * W / out["fp64"]["sec"]) / PEAK["fp64"],
            "step_roofline_frac_all_complex_count": (fstep_all * W / out["fp64"]["sec"])
                                                    / PEAK["fp64"]},
        "mixed_over_fp64": None if "fp64" not in out else out["fp64"]["sec"] / sec,
        # one local-energy evaluation of all walkers (measured every measure_interval steps);
        # the exchange contraction dominates: 4 n^2 N M FLOP per walker with the real trial
        # walker energy evaluations/s, and the headline with one evaluation every 10 steps
        # (the mixed estimator's default measure interval) charged to the steps
        "energy": None if out["mixed"].get("energy") is None else {
            name: {"ms": out[name]["energy"] * 1e3,
                   "exchange_tflops": 4.0 * NUP * NUP * N * M * W / out[name]["energy"] / 1e12,
                   "walker_evaluations_per_s": W / out[name]["energy"],
                   "walker_steps_per_s_with_energy_every_10":
                       W / (out[name]["sec"] + out[name]["energy"] / 10)}
            for name in out if out[name].get("energy")},
        "shard_one_of_8": shard,
        # the state of the walkers after the timed steps, and the same shapes at the generator's
        # default Cholesky amplitude 0.5 (walkers dead within about four steps)
        "walkers_after_timed_steps": {k: out[k]["state"] for k in out},
        "recipe_amplitude_0.5": recipe,
        "energy_error_mixed_vs_fp64_Ha": None if emax is None else
            {"max": emax, "median": emed, "max_rel": erel, "median_abs_E": emag},
        "energy_error_mixed_vs_fp64_BH": bh_err,
        "energy_error_mixed_vs_fp64_per_config": cfg_err,
        "mixed_precision_error_budget": budget,
        "phmsd": phmsd,
        "serve": serve_summary,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
