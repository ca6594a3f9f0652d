"""Parity of the device PHMSD trial (COLLINEAR, particle-hole multi-determinant expansion)
with the FP64 oracle's restatement of the reference algorithm, through the C ABI.

Bar: 1e-10 relative (max |gpu - oracle| / max |oracle| per array) for the compact mixed
density matrix, the log overlaps (imaginary parts compared modulo 2 pi), the local energies,
vbias and vMF; whole propagation steps of a walker set against an oracle step built from the
oracle's pinned pieces; a captured graph step against the eager step.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd.phmsd import build_ph_struct  # noqa: E402
from safire_amd.synthetic import make_phmsd_problem  # noqa: E402

TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    """The larger of max |a - b| / max |b| and the per-walker error (leading axis), each
    walker's scale floored at 1e-6 max |b|."""
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    assert np.isfinite(a).all() and np.isfinite(b).all()
    d = np.max(np.abs(b))
    d = d if d > 0 else 1.0
    if a.ndim == 0:
        return float(np.abs(a - b) / d)
    dw = np.abs(a - b).reshape(len(a), -1).max(axis=1)
    sw = np.maximum(np.abs(b).reshape(len(b), -1).max(axis=1), 1e-6 * d)
    return float(max(dw.max() / d, np.max(dw / sw)))


def rel_log(a, b):
    """Relative difference of log overlaps with the imaginary parts taken modulo 2 pi."""
    a, b = np.asarray(a), np.asarray(b)
    d = a - b
    im = (d.imag + np.pi) % (2 * np.pi) - np.pi
    return float(np.max(np.abs(d.real + 1j * im)) / max(np.max(np.abs(b)), 1.0))


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import afq
    ctx = afq.Context(0)
    return afq, ctx, torch


def setup_case(afq, ctx, **kw):
    from safire_amd.phmsd import PHMSD
    P = make_phmsd_problem(**kw)
    ph = build_ph_struct(P.occa, P.occb, P.ci)
    ham = afq.HamiltonianOperations(ctx, P.h1, P.Likn, P.PsiT, P.nact, P.nact,
                                    walker_type=afq.AFQ_COLLINEAR)
    trial = PHMSD(ham, ph)
    L1, h1 = orc.half_rotate(P.orbs, P.Likn, P.h1)
    tens = dict(Lblk=np.concatenate([L1, L1]), hblk=np.concatenate([h1, h1]),
                Lstk=np.concatenate([L1, L1], axis=1))
    return P, ph, ham, trial, orc.PhTrial(ph, P.orbs), tens


SMALL = dict(N=24, nel=6, M=40, ndet=60, ncore=2, nactive=8, max_rank=4, seed=21)
RANK6 = dict(N=28, nel=8, M=32, ndet=80, ncore=1, nactive=14, max_rank=6, seed=22)
PH2000 = dict()   # make_phmsd_problem defaults: N=64, 10+10 electrons, M=320, 2000 determinants


@pytest.mark.parametrize("case,W", [("small", 16), ("rank6", 8), ("ph2000", 12)])
def test_phmsd_trial_ops(gpu, case, W):
    afq, ctx, torch = gpu
    kw = {"small": SMALL, "rank6": RANK6, "ph2000": PH2000}[case]
    say(f"case: PHMSD trial ops {case} W={W}")
    P, ph, ham, trial, ot, t = setup_case(afq, ctx, **kw)
    say("  nconf", ph.nconf, "ndet", ph.ndet, "max rank", ph.max_rank(0), ph.max_rank(1))
    Pa, Pb = P.walkers(W, seed=5, noise=0.3)
    ws = afq.WalkerSet(Pa, Pb)
    G = afq.empty((W, 2 * P.nact, P.N), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    trial.MixedDensityMatrix(ws, G, ov)
    Gr, ovr = ot.mixed_density_matrix(Pa, Pb)
    eG, eo = rel(afq.host(G), Gr), rel_log(afq.host(ov), ovr)
    say(f"  G {eG:.2e}  Ov {eo:.2e}")
    assert eG < TOL and eo < TOL
    ov2 = afq.empty(W, torch.complex128)
    trial.Log_Overlap(ws, ov2)
    assert rel_log(afq.host(ov2), ot.log_overlap(Pa, Pb)) < TOL
    E = afq.empty((W, 3), torch.complex128)
    ov3 = afq.empty(W, torch.complex128)
    trial.Energy(ws, E, ov3)
    Er, ovr3 = ot.energy(t["hblk"], t["Lblk"], Pa, Pb)
    eE = rel(afq.host(E), Er)
    say(f"  E {eE:.2e}  (E[0] gpu {afq.host(E)[0]}  oracle {Er[0]})")
    assert eE < TOL and rel_log(afq.host(ov3), ovr3) < TOL
    # PHMSD::vbias: compact multi-determinant G through the hamiltonian's vbias (c = 1)
    v = afq.empty((W, P.M), torch.complex128)
    trial.vbias(ws, v, 0.005)
    vr = orc.vbias(t["Lstk"], Gr, 0.005, c=1.0)
    assert rel(afq.host(v), vr) < TOL


def test_phmsd_vmf(gpu):
    afq, ctx, torch = gpu
    say("case: PHMSD vMF small + 2000 determinants")
    for kw in (SMALL, PH2000):
        P, ph, ham, trial, ot, t = setup_case(afq, ctx, **kw)
        v = afq.empty(P.M, torch.complex128)
        trial.vMF(v, 0.005)
        vr = orc.vbias(t["Lstk"], ot.gmf()[None], 0.005, c=1.0)[0]
        vr = vr.real + 0j
        e = rel(afq.host(v), vr)
        say(f"  N={P.N} vMF {e:.2e}")
        assert e < TOL


def oracle_step(P, ot, t, vMF, P1, state, u, Eshift, dt=0.005):
    """One hybrid step of COLLINEAR walkers over the PHMSD trial, from the oracle's pieces
    (AFQMCBasePropagator.cpp:289-392 with PHMSD::vbias and PHMSD::Log_Overlap)."""
    Pa, Pb = state["Pa"], state["Pb"]
    G, _ = ot.mixed_density_matrix(Pa, Pb)
    vb = orc.vbias(t["Lstk"], G, dt, c=1.0)
    X, HW, mf = orc.construct_x(u, vb, vMF, 50.0)
    V = orc.vhs(P.Likn, X, dt)
    state["Pa"] = orc.propagate(P1, V, Pa)
    state["Pb"] = orc.propagate(P1, V, Pb)
    nov = ot.log_overlap(state["Pa"], state["Pb"])
    orc.walker_update(state, dt, Eshift, 0.0, nov, mf, HW)


def test_phmsd_propagation_matches_oracle(gpu):
    afq, ctx, torch = gpu
    W, steps, dt = 16, 4, 0.005
    say(f"case: PHMSD propagation small W={W} steps={steps}")
    P, ph, ham, trial, ot, t = setup_case(afq, ctx, **SMALL)
    prop = afq.Propagator(ham, W, dt=dt, seed=77, trial=trial)
    # oracle setup: vMF from the trial density, P1 from the same generateP1 restatement
    vMF = orc.vbias(t["Lstk"], ot.gmf()[None], dt, c=1.0)[0].real + 0j
    assert rel(afq.host(prop.vMF()), vMF) < TOL
    P1 = orc.generate_p1(P.h1, orc.vexx(P.Likn), P.Likn, vMF, dt)
    assert rel(afq.host(prop.P1()), P1) < TOL
    Pa, Pb = P.walkers(W, seed=9, noise=0.2)
    ws = afq.WalkerSet(Pa, Pb)
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    prop.Energy(ws, E, ov)
    Er, ovr = ot.energy(t["hblk"], t["Lblk"], Pa, Pb)
    assert rel(afq.host(E), Er) < TOL
    ws.eloc.copy_(E.sum(1))
    ws.ovlp.copy_(ov)
    state = dict(Pa=Pa.copy(), Pb=Pb.copy(), weight=np.ones(W, complex), phase=np.ones(W, complex),
                 eloc=Er.sum(1), ovlp=ovr.copy(), theta=np.zeros(W, complex))
    Eshift = float(np.mean(Er.sum(1)).real)
    Esh = afq.device_tensor(np.array([Eshift]))
    for k in range(steps):
        u = orc.field_uniforms(77, 1, k, W, P.M)
        prop.Propagate(ws, Esh, u=afq.device_tensor(u))
        oracle_step(P, ot, t, vMF, P1, state, u, Eshift, dt)
        ea = rel(afq.host(ws.Phi), state["Pa"])
        eb = rel(afq.host(ws.Phi_b), state["Pb"])
        ew = rel(afq.host(ws.weight), state["weight"])
        eo = rel_log(afq.host(ws.ovlp), state["ovlp"])
        say(f"  step {k}: Phi_a {ea:.2e} Phi_b {eb:.2e} weight {ew:.2e} ovlp {eo:.2e}")
        assert max(ea, eb, ew, eo) < TOL
    prop.Orthogonalize(ws)
    Qa, _ = orc.orthogonalize(state["Pa"])
    Qb, _ = orc.orthogonalize(state["Pb"])
    assert rel(afq.host(ws.Phi), Qa) < TOL and rel(afq.host(ws.Phi_b), Qb) < TOL
    assert rel_log(afq.host(ws.ovlp), ot.log_overlap(Qa, Qb)) < TOL


def test_phmsd_graph_equals_eager(gpu):
    afq, ctx, torch = gpu
    W = 32
    say(f"case: PHMSD graph replay vs eager N=64 2000-determinant-shaped small W={W}")
    P, ph, ham, trial, ot, t = setup_case(afq, ctx, **SMALL)
    Pa, Pb = P.walkers(W, seed=3, noise=0.2)
    out = []
    for mode in ("eager", "graph"):
        prop = afq.Propagator(ham, W, dt=0.005, seed=5, trial=trial)
        ws = afq.WalkerSet(Pa, Pb)
        E = afq.empty((W, 3), torch.complex128)
        ov = afq.empty(W, torch.complex128)
        prop.Energy(ws, E, ov)
        ws.eloc.copy_(E.sum(1))
        ws.ovlp.copy_(ov)
        Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).mean().real)]))
        eo = afq.empty(1, torch.float64)
        for k in range(6):
            if mode == "eager":
                prop.Propagate(ws, Esh)
                if k % 3 == 0:
                    prop.Orthogonalize(ws)
            else:
                prop.step_graph(ws, Esh, ortho=(k % 3 == 0), energy_out=eo)
        out.append((afq.host(ws.Phi), afq.host(ws.Phi_b), afq.host(ws.weight)))
        prop.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
