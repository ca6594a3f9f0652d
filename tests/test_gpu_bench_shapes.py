"""Parity at the benchmark's own shapes, through the C ABI.

The benchmark workload is N=200 orbitals, 40+40 electrons (CLOSED), M=1600 Cholesky vectors,
1024 walkers (safire_amd.synthetic seed 1003, here at the generator's default Cholesky amplitude;
the benchmark times the same shapes, hence the same kernels and tiles, at amplitude 0.1).  The
autotuner picks tiles per shape bucket, so these tests run at exactly that shape:

* the FP64 step of all 1024 walkers (tuned tiles, as the benchmark runs it) against the oracle
  on eight sampled walkers, 1e-10 relative per walker;
* every tile candidate of every tuned product site of the mixed step, the FP64 step and both
  local-energy evaluations, pinned one at a time (afq_tune_force): each must leave the step's
  walkers, overlaps and weights (or the energies) bit-identical to the tuned run, which is the
  tuner's claim that a choice never changes a bit;
* the captured (graph) step with in-kernel Philox fields, 1024 walkers, against the oracle on
  the sampled walkers.
"""
import re
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd.synthetic import make_problem, random_walkers  # noqa: E402

N, NUP, M, W = 200, 40, 1600, 1024
SAMPLE = [0, 1, 137, 511, 512, 777, 1000, 1023]
TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def per_walker_rel(a, b):
    """max over walkers of max |a_w - b_w| / max |b_w| (leading axis = walkers)."""
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    d = np.max(np.abs(a - b), axis=1)
    s = np.max(np.abs(b), axis=1)
    return float(np.max(d / np.where(s > 0, s, 1.0)))


def weight_rel(got, st, w_old, ov_old, Esh, dt=0.005):
    """Per-walker weight error in the units its inputs fix.  The step multiplies the weight by
    max(0, cos dtheta) |e^{-dt (E_loc - E_shift)}| with dtheta = Im(ov_new - ov_old - mf), so an
    error d in the log overlaps (held to 1e-10 relative to their own size, i.e. d ~ 1e-10 |ov|)
    moves the weight by up to |w_old| |e^{...}| |sin dtheta| d: the unit is the unconstrained
    weight times max(1, |ov_new| + |ov_old|).  A relative check of the weight alone would
    measure the conditioning of the walker (a cosine near zero, log overlaps of size ~10^3), not
    the arithmetic."""
    unc = np.abs(w_old) * np.exp(-dt * (st["eloc"].real - Esh))
    scale = unc * np.maximum(1.0, np.abs(st["ovlp"]) + np.abs(ov_old))
    return float(np.max(np.abs(np.asarray(got) - st["weight"]) / np.where(scale > 0, scale, 1.0)))


def phase_rel(got, st, ov_old):
    """Per-walker phase error in units of the walker's log-overlap magnitude: the step turns the
    phase by Im(mf - HW - (ov_new - ov_old)), so the phase inherits the absolute error of the
    log overlaps, which are held to 1e-10 relative to their own size."""
    scale = np.maximum(1.0, np.maximum(np.abs(st["ovlp"]), np.abs(ov_old)))
    return float(np.max(np.abs(np.asarray(got) - st["phase"]) / scale))


@pytest.fixture(scope="module")
def bench():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import afq
    say(f"case: benchmark workload setup N={N} n={NUP} M={M} W={W}")
    p = make_problem(N, NUP, M, 1003)
    osys = orc.OracleSystem(p, dt=0.005)
    Phi = random_walkers(p, W, seed=31, noise=0.3)
    st_s, Esh = osys.init_walkers(Phi[SAMPLE])
    return afq, torch, p, osys, Phi, st_s, Esh


def _setup(afq, torch, p, precision, Phi, Esh, seed=4242):
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, p.nup, precision=precision)
    prop = afq.Propagator(ham, W, dt=0.005, seed=seed)
    ws = afq.WalkerSet(Phi)
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    prop.Energy(ws, E, ov)
    ws.eloc.copy_(E.sum(1))
    ws.ovlp.copy_(ov)
    return ctx, ham, prop, ws, afq.device_tensor(np.array([Esh]))


def test_fp64_step_at_bench_shape_vs_oracle(bench):
    afq, torch, p, osys, Phi, st_s, Esh = bench
    say("case: FP64 step, 1024 walkers, tuned tiles, vs oracle on 8 sampled walkers")
    ctx, ham, prop, ws, dE = _setup(afq, torch, p, afq.AFQ_FP64, Phi, Esh)
    st = {k: v.copy() for k, v in st_s.items()}
    # the initial energies and overlaps at the benchmark shape
    assert per_walker_rel(afq.host(ws.eloc)[SAMPLE], st["eloc"]) < TOL
    assert per_walker_rel(afq.host(ws.ovlp)[SAMPLE], st["ovlp"]) < TOL
    for step in range(2):
        u = orc.field_uniforms(17, 1, step, W, M)
        prop.Propagate(ws, dE, u=afq.device_tensor(u))
        w_old, ov_old = st["weight"].copy(), st["ovlp"].copy()
        osys.step(st, u[SAMPLE], Esh)
        if step == 0:
            prop.Orthogonalize(ws)
            osys.orthogonalize(st)
        e = per_walker_rel(afq.host(ws.Phi)[SAMPLE], st["Phi"])
        say(f"  step {step}: Phi per-walker rel {e:.2e}")
        assert e < TOL, step
        for k in ("ovlp", "eloc", "theta"):
            ek = per_walker_rel(afq.host(getattr(ws, k))[SAMPLE], st[k])
            assert ek < TOL, (step, k, ek)
        assert phase_rel(afq.host(ws.phase)[SAMPLE], st, ov_old) < TOL, step
        wg = afq.host(ws.weight)[SAMPLE]
        ew = weight_rel(wg, st, w_old, ov_old, Esh)
        say(f"  step {step}: weight error {ew:.2e} of the unconstrained weight; per walker "
            f"|dw| {np.abs(wg - st['weight'])}, w {np.abs(st['weight'])}, "
            f"cos dtheta {np.cos(st['theta'].real)}")
        assert ew < TOL, (step, ew)


def test_graph_step_at_bench_shape_vs_oracle(bench):
    """The captured step (in-kernel Philox fields keyed by the device step counter, QR in the
    first step) of all 1024 walkers; the oracle steps the sampled walkers with the same
    uniforms.  Eshift is not relaxed (relax = 0), so every walker's step is independent."""
    afq, torch, p, osys, Phi, st_s, Esh = bench
    say("case: graph step FP64, 1024 walkers, vs oracle on 8 sampled walkers")
    seed = 4242
    ctx, ham, prop, ws, dE = _setup(afq, torch, p, afq.AFQ_FP64, Phi, Esh, seed=seed)
    st = {k: v.copy() for k, v in st_s.items()}
    e_out = afq.empty(1, torch.float64)
    # two steps, as the eager test: at this shape a step amplifies rounding differences ~100x
    # (|P1|^2 e^|V| with |V|_2 ~ 3.4), so the oracle and the device, which sum in different
    # orders, part at 1e-10 relative in the third step (measured 4e-14, 2e-12, 2.6e-10)
    for step in range(2):
        ortho = step == 0
        prop.step_graph(ws, dE, ortho=ortho, relax=0.0, energy_out=e_out)
        u = orc.field_uniforms(seed, 1, step, W, M)
        w_old, ov_old = st["weight"].copy(), st["ovlp"].copy()
        osys.step(st, u[SAMPLE], Esh)
        if ortho:
            osys.orthogonalize(st)
        e = per_walker_rel(afq.host(ws.Phi)[SAMPLE], st["Phi"])
        say(f"  step {step}: Phi per-walker rel {e:.2e}")
        assert e < TOL, step
        for k in ("ovlp", "eloc"):
            assert per_walker_rel(afq.host(getattr(ws, k))[SAMPLE], st[k]) < TOL, (step, k)
        assert phase_rel(afq.host(ws.phase)[SAMPLE], st, ov_old) < TOL, step
        assert weight_rel(afq.host(ws.weight)[SAMPLE], st, w_old, ov_old, Esh) < TOL, step


def _sites(afq):
    """{site:precision -> number of candidates} seen by the tuner so far."""
    out = {}
    for line in afq.tuning_report().splitlines():
        m = re.match(r"^([a-z0-9_]+):(f32|f64):\S+ -> .* of (\d+)$", line.strip())
        if m:
            key = f"{m.group(1)}:{m.group(2)}"
            out[key] = max(out.get(key, 0), int(m.group(3)))
    return out


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_every_tile_candidate_is_bit_identical(bench, precision):
    """Pins each candidate of each product site the step and the energy touched at the
    benchmark shape and re-runs the same step (or energy) from the same state: walkers,
    overlaps, weights and energies must equal the tuned run's bits."""
    afq, torch, p, osys, Phi, st_s, Esh = bench
    prec = afq.AFQ_MIXED if precision == "mixed" else afq.AFQ_FP64
    say(f"case: every tile candidate, {precision} step + energy, 1024 walkers")
    ctx, ham, prop, ws, dE = _setup(afq, torch, p, prec, Phi, Esh)
    init = {k: getattr(ws, k).clone() for k in ("Phi", "weight", "phase", "eloc", "ovlp", "theta")}
    u = afq.device_tensor(orc.field_uniforms(23, 1, 0, W, M))
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)

    def run():
        for k, v in init.items():
            getattr(ws, k).copy_(v)
        prop.Propagate(ws, dE, u=u)
        prop.Orthogonalize(ws)
        prop.Energy(ws, E, ov, prec)
        torch.cuda.synchronize()
        return [afq.host(getattr(ws, k)) for k in ("Phi", "ovlp", "weight", "eloc")] + \
            [afq.host(E), afq.host(ov)]

    ref = run()          # tuned choices (first call tunes, as the benchmark's warm-up does)
    ref2 = run()
    for a, b in zip(ref, ref2):
        assert np.array_equal(a, b)
    sites = _sites(afq)
    prefix = "f32" if prec == afq.AFQ_MIXED else "f64"
    big = [s for s in sites if s.startswith(("vhs", "vbias", "exx", "prop_fused"))]
    say(f"  sites: {sites}")
    # vHS and vbias run the register-tiled kernel (crgemm.hpp), which has no tile candidates
    assert any(s.startswith("prop_fused:" + prefix) for s in big)
    assert any(s.startswith("exx_chunk") for s in big)
    runs = 0
    for site, ncand in sorted(sites.items()):
        for i in range(ncand):
            afq.tune_force(site, i)
            try:
                got = run()
            finally:
                afq.tune_force(site, -1)
            for name, a, b in zip(("Phi", "ovlp", "weight", "eloc", "E", "E_ovlp"), got, ref):
                assert np.array_equal(a, b), (site, i, name)
            runs += 1
    say(f"  {runs} pinned runs over {len(sites)} sites, all bit-identical")
    # the FP64 run is the oracle's: check the tuned result on the sampled walkers
    if prec == afq.AFQ_FP64:
        st = {k: v.copy() for k, v in st_s.items()}
        osys.step(st, afq.host(u)[SAMPLE], Esh)
        osys.orthogonalize(st)
        assert per_walker_rel(ref[0][SAMPLE], st["Phi"]) < TOL
        Gr, _ = orc.mixed_density_matrix(p.PsiT, st["Phi"])
        Er = orc.energy(osys.haj, osys.Lnak, Gr, p.nup)
        assert per_walker_rel(ref[4][SAMPLE], Er) < TOL


def test_forced_candidate_out_of_range_is_an_error(bench, monkeypatch):
    afq, torch, p, osys, Phi, st_s, Esh = bench
    # the tuned site is the LDS-staged GEMM engine's vHS; the register-tiled product is untuned
    monkeypatch.setenv("AFQ_CRGEMM", "0")
    from safire_amd.afq import AfqError
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, p.nup, precision=afq.AFQ_FP64)
    X = afq.device_tensor(np.ones((4, M), np.complex128))
    V = afq.empty((4, N, N), torch.complex128)
    afq.tune_force("vhs_sym:f64", 99)
    try:
        with pytest.raises(AfqError, match="candidates"):
            ham.vHS(X, V, 0.005, precision=afq.AFQ_FP64)
    finally:
        afq.tune_force(None)
    ham.vHS(X, V, 0.005, precision=afq.AFQ_FP64)
