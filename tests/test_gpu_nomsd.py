"""Parity of the device multi-determinant NOMSD trial (NOMSD.cpp with ndet > 1, CLOSED and
COLLINEAR walkers) with the FP64 oracle's restatement, through the C ABI.

Bar: 1e-10 relative (max |gpu - oracle| / max |oracle| per array, per walker) in FP64 for the
log overlaps (imaginary parts modulo 2 pi), the full mixed density matrix and the Ov it leaves,
vbias, the local energies and vMF; whole propagation steps against an oracle step built from
the oracle's pinned pieces; a captured graph step equal to the eager step.  Mixed precision
(FP32 tensors and contractions): 2e-5 relative for the energies and vbias.
"""
import dataclasses
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd.synthetic import make_problem, random_walkers  # noqa: E402

TOL = 1e-10
TOL32 = 2e-5


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


def _rotated(rng, PsiT, mix):
    """Orthonormal rows near PsiT: a complex perturbation, re-orthonormalised."""
    A = PsiT + mix * (rng.normal(size=PsiT.shape) + 1j * rng.normal(size=PsiT.shape))
    Q, _ = np.linalg.qr(A.conj().T)
    return np.ascontiguousarray(Q.conj().T)


@dataclasses.dataclass
class Case:
    p: object            # the CLOSED problem (alpha block)
    closed: bool
    nup: int
    ndn: int
    dets: list           # [(PsiT_a, PsiT_b or None)]
    ci: np.ndarray
    PsiT_b: np.ndarray   # beta rows of determinant 0 (COLLINEAR)

    def walkers(self, W, seed, noise=0.3):
        Pa = random_walkers(self.p, W, seed=seed, noise=noise)
        if self.closed:
            return Pa, None
        pb = dataclasses.replace(self.p, nup=self.ndn, PsiT=self.PsiT_b)
        return Pa, random_walkers(pb, W, seed=seed + 101, noise=noise)

    @property
    def c(self):
        return 2.0 if self.closed else 1.0


def make_case(closed, ndet=4, N=24, nup=4, ndn=3, M=48, seed=31, mix=0.25):
    p = make_problem(N, nup, M, seed)
    rng = np.random.default_rng(seed + 7)
    PsiT_b = np.ascontiguousarray(p.PsiT[:ndn])
    dets = [(p.PsiT, None if closed else PsiT_b)]
    for _ in range(ndet - 1):
        dets.append((_rotated(rng, p.PsiT, mix), None if closed else _rotated(rng, PsiT_b, mix)))
    ci = rng.normal(size=ndet) + 0.4j * rng.normal(size=ndet)
    ci[0] = 1.0
    return Case(p, closed, nup, 0 if closed else ndn, dets, ci, PsiT_b)


def build(afq, ctx, torch, cs, precision=None):
    from safire_amd.nomsd import NOMSD
    prec = afq.AFQ_FP64 if precision is None else precision
    if cs.closed:
        ham = afq.HamiltonianOperations(ctx, cs.p.h1, cs.p.Likn, cs.p.PsiT, cs.nup,
                                        precision=prec)
    else:
        ham = afq.HamiltonianOperations(ctx, cs.p.h1, cs.p.Likn,
                                        np.concatenate([cs.p.PsiT, cs.PsiT_b]), cs.nup, cs.ndn,
                                        walker_type=afq.AFQ_COLLINEAR, precision=prec)
    rows = [a if b is None else np.concatenate([a, b]) for a, b in cs.dets]
    PsiT = afq.device_tensor(np.ascontiguousarray(np.stack(rows)), np.complex128)
    return ham, NOMSD(ham, PsiT, cs.ci)


def oracle_tensors(cs):
    return orc.msd_half_rotated(cs.dets, cs.p.Likn, cs.p.h1)


@pytest.mark.parametrize("closed", [True, False])
def test_nomsd_trial_ops(gpu, closed):
    afq, ctx, torch = gpu
    W = 12
    cs = make_case(closed)
    say(f"case: NOMSD trial ops closed={closed} ndet={len(cs.ci)} W={W}")
    ham, trial = build(afq, ctx, torch, cs)
    Pa, Pb = cs.walkers(W, seed=5)
    ws = afq.WalkerSet(Pa, Pb)
    ovr = orc.msd_log_overlap(cs.dets, cs.ci, Pa, Pb)
    ov = afq.empty(W, torch.complex128)
    trial.Log_Overlap(ws, ov)
    eo = rel_log(afq.host(ov), ovr)
    say(f"  log overlap {eo:.2e}")
    assert eo < TOL
    # the full density matrix with a scale away from the overlap (the scale cancels in G)
    log_m = ovr + 0.7 - 0.3j
    G = afq.empty((W, (1 if closed else 2) * cs.p.N * cs.p.N), torch.complex128)
    Ov = afq.empty(W, torch.complex128)
    trial.MixedDensityMatrix(ws, G, Ov, log_m=afq.device_tensor(log_m))
    Gr, Ovr = orc.msd_mixed_density_matrix(cs.dets, cs.ci, Pa, Pb, log_m)
    eG, eOv = rel(afq.host(G), Gr), rel(afq.host(Ov), Ovr)
    say(f"  G {eG:.2e}  Ov {eOv:.2e}")
    assert eG < TOL and eOv < TOL
    # vbias scales by the walkers' stored log overlap
    ws.ovlp.copy_(afq.device_tensor(ovr))
    v = afq.empty((W, cs.p.M), torch.complex128)
    trial.vbias(ws, v, 0.005)
    vr = orc.vbias_full(cs.p.Likn, Gr, 0.005, cs.c)
    ev = rel(afq.host(v), vr)
    say(f"  vbias {ev:.2e}")
    assert ev < TOL
    E = afq.empty((W, 3), torch.complex128)
    ov3 = afq.empty(W, torch.complex128)
    trial.Energy(ws, E, ov3)
    Er, ovr3 = orc.msd_energy(cs.dets, cs.ci, oracle_tensors(cs), Pa, Pb)
    eE = rel(afq.host(E), Er)
    say(f"  E {eE:.2e}  (E[0] gpu {afq.host(E)[0]}  oracle {Er[0]})")
    assert eE < TOL and rel_log(afq.host(ov3), ovr3) < TOL
    trial.close()


@pytest.mark.parametrize("closed", [True, False])
def test_nomsd_vmf(gpu, closed):
    afq, ctx, torch = gpu
    cs = make_case(closed, ndet=3)
    say(f"case: NOMSD vMF closed={closed}")
    ham, trial = build(afq, ctx, torch, cs)
    v = afq.empty(cs.p.M, torch.complex128)
    trial.vMF(v, 0.005)
    vr = orc.msd_vmf(cs.dets, cs.ci, cs.p.Likn, 0.005)
    e = rel(afq.host(v), vr)
    say(f"  vMF {e:.2e}")
    assert e < TOL
    trial.close()


@pytest.mark.parametrize("closed", [True, False])
def test_nomsd_mixed_precision(gpu, closed):
    afq, ctx, torch = gpu
    W = 16
    cs = make_case(closed, N=40, nup=8, ndn=6, M=80, seed=41)
    say(f"case: NOMSD mixed precision closed={closed} W={W}")
    ham, trial = build(afq, ctx, torch, cs, precision=afq.AFQ_MIXED)
    Pa, Pb = cs.walkers(W, seed=8)
    ws = afq.WalkerSet(Pa, Pb)
    ovr = orc.msd_log_overlap(cs.dets, cs.ci, Pa, Pb)
    ws.ovlp.copy_(afq.device_tensor(ovr))
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    trial.Energy(ws, E, ov, precision=afq.AFQ_MIXED)
    Er, _ = orc.msd_energy(cs.dets, cs.ci, oracle_tensors(cs), Pa, Pb)
    eE = rel(afq.host(E), Er)
    v = afq.empty((W, cs.p.M), torch.complex128)
    trial.vbias(ws, v, 0.005, precision=afq.AFQ_MIXED)
    Gr, _ = orc.msd_mixed_density_matrix(cs.dets, cs.ci, Pa, Pb, ovr)
    ev = rel(afq.host(v), orc.vbias_full(cs.p.Likn, Gr, 0.005, cs.c))
    say(f"  mixed E {eE:.2e}  vbias {ev:.2e}")
    assert eE < TOL32 and ev < TOL32
    trial.close()


def oracle_step(cs, vMF, P1, state, u, Eshift, dt):
    """One hybrid step over the NOMSD trial from the oracle's pieces
    (AFQMCBasePropagator.cpp:289-392 with NOMSD::vbias, NOMSD.cpp:39-65, and
    NOMSD::Log_Overlap)."""
    Pa, Pb = state["Pa"], state.get("Pb")
    G, _ = orc.msd_mixed_density_matrix(cs.dets, cs.ci, Pa, Pb, state["ovlp"])
    vb = orc.vbias_full(cs.p.Likn, G, dt, cs.c)
    X, HW, mf = orc.construct_x(u, vb, vMF, 50.0)
    V = orc.vhs(cs.p.Likn, X, dt)
    state["Pa"] = orc.propagate(P1, V, Pa)
    if Pb is not None:
        state["Pb"] = orc.propagate(P1, V, Pb)
    nov = orc.msd_log_overlap(cs.dets, cs.ci, state["Pa"], state.get("Pb"))
    orc.walker_update(state, dt, Eshift, 0.0, nov, mf, HW)


@pytest.mark.parametrize("closed", [True, False])
def test_nomsd_propagation_matches_oracle(gpu, closed):
    afq, ctx, torch = gpu
    W, steps, dt = 12, 3, 0.005
    cs = make_case(closed, ndet=3)
    say(f"case: NOMSD propagation closed={closed} W={W} steps={steps}")
    ham, trial = build(afq, ctx, torch, cs)
    prop = afq.Propagator(ham, W, dt=dt, seed=77, trial=trial)
    vMF = orc.msd_vmf(cs.dets, cs.ci, cs.p.Likn, dt)
    assert rel(afq.host(prop.vMF()), vMF) < TOL
    P1 = orc.generate_p1(cs.p.h1, orc.vexx(cs.p.Likn), cs.p.Likn, vMF, dt)
    assert rel(afq.host(prop.P1()), P1) < TOL
    Pa, Pb = cs.walkers(W, seed=9, noise=0.2)
    ws = afq.WalkerSet(Pa, Pb)
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    prop.Energy(ws, E, ov)
    Er, ovr = orc.msd_energy(cs.dets, cs.ci, oracle_tensors(cs), Pa, Pb)
    assert rel(afq.host(E), Er) < TOL
    ws.eloc.copy_(E.sum(1))
    ws.ovlp.copy_(ov)
    state = dict(Pa=Pa.copy(), weight=np.ones(W, complex), phase=np.ones(W, complex),
                 eloc=Er.sum(1), ovlp=ovr.copy(), theta=np.zeros(W, complex))
    if not closed:
        state["Pb"] = Pb.copy()
    Eshift = float(np.mean(Er.sum(1)).real)
    Esh = afq.device_tensor(np.array([Eshift]))
    for k in range(steps):
        u = orc.field_uniforms(77, 1, k, W, cs.p.M)
        prop.Propagate(ws, Esh, u=afq.device_tensor(u))
        oracle_step(cs, vMF, P1, state, u, Eshift, dt)
        ea = rel(afq.host(ws.Phi), state["Pa"])
        eb = 0.0 if closed else rel(afq.host(ws.Phi_b), state["Pb"])
        ew = rel(afq.host(ws.weight), state["weight"])
        eo = rel_log(afq.host(ws.ovlp), state["ovlp"])
        say(f"  step {k}: Phi_a {ea:.2e} Phi_b {eb:.2e} weight {ew:.2e} ovlp {eo:.2e}")
        assert max(ea, eb, ew, eo) < TOL
    prop.Orthogonalize(ws)
    Qa, _ = orc.orthogonalize(state["Pa"])
    Qb = None if closed else orc.orthogonalize(state["Pb"])[0]
    assert rel(afq.host(ws.Phi), Qa) < TOL
    if not closed:
        assert rel(afq.host(ws.Phi_b), Qb) < TOL
    assert rel_log(afq.host(ws.ovlp), orc.msd_log_overlap(cs.dets, cs.ci, Qa, Qb)) < TOL
    prop.close()
    trial.close()


def test_nomsd_graph_equals_eager(gpu):
    afq, ctx, torch = gpu
    W = 16
    cs = make_case(False, ndet=3)
    say(f"case: NOMSD graph replay vs eager COLLINEAR W={W}")
    ham, trial = build(afq, ctx, torch, cs)
    Pa, Pb = cs.walkers(W, seed=3, noise=0.2)
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
    trial.close()


def test_nomsd_rejects_mismatched_first_determinant(gpu):
    afq, ctx, torch = gpu
    from safire_amd.nomsd import NOMSD
    cs = make_case(True, ndet=2)
    say("case: NOMSD determinant 0 differs from the hamiltonian's trial")
    ham, _ = build(afq, ctx, torch, cs)
    rows = np.stack([cs.dets[1][0], cs.dets[0][0]])
    with pytest.raises(afq.AfqError):
        NOMSD(ham, afq.device_tensor(rows, np.complex128), cs.ci)


@pytest.mark.parametrize("closed", [True, False])
def test_nomsd_measurements_vs_oracle(gpu, closed):
    """Every determinant is a reference of the mixed and back-propagated OneRDM /
    DiagonalTwoRDM (NOMSD::getReferences, NOMSD.cpp:103-129): the device measurements equal the
    oracle's sums over the references of density matrices weighted by conj(c_d) times the
    overlap (times e^{conj(log det R)} after back-propagation).  CLOSED walkers take the alpha
    overlap once in the mixed estimator and doubled, with log det R, in the back-propagated one
    (MixedEstimator.hpp:61-75, BackPropEstimator.hpp:62-81)."""
    afq, ctx, torch = gpu
    W, seed, nback, nbp, ortho, dt = 8, 41, 5, 4, 2, 0.005
    cs = make_case(closed, ndet=4)
    say(f"case: NOMSD measurements closed={closed} ndet=4 W={W}, window {nbp}, QR every {ortho}")
    ham, trial = build(afq, ctx, torch, cs)
    prop = afq.Propagator(ham, W, dt=dt, seed=seed, trial=trial)
    assert prop.number_of_references() == len(cs.ci)
    prop.enable_backprop(nback)
    ws = afq.WalkerSet(*cs.walkers(W, seed=3, noise=0.3))
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    prop.Energy(ws, E, ov)
    ws.eloc.copy_(E.sum(1))
    ws.ovlp.copy_(ov)
    Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).real.mean())]))
    prop.set_step(0)
    prop.Propagate(ws, Esh)
    prop.bp_anchor(ws)
    anchors = [afq.host(ws.Phi)] + ([] if closed else [afq.host(ws.Phi_b)])
    for _ in range(nbp):
        prop.Propagate(ws, Esh)
    step = 1 + nbp
    Ps = [afq.host(ws.Phi)] + ([] if closed else [afq.host(ws.Phi_b)])
    w = afq.host(ws.weight)
    refs = [[a] + ([] if closed else [b]) for a, b in cs.dets]

    def measure(Gs, los, weights, scale):
        lo = scale * np.array(los)
        x = np.conj(cs.ci)[:, None] * np.exp(lo - lo.real.max(0))
        cf = x / x.sum(0)
        one = two = 0
        for d, G in enumerate(Gs):
            o, t = orc.rdm_measure(G, weights, closed, cf[d])
            one, two = one + o, two + t
        return one, two

    # mixed estimator
    Gs, los = [], []
    for r in refs:
        blocks = [orc.mixed_density_matrix(r[sp], Ps[sp], compact=False) for sp in range(len(r))]
        Gs.append(np.stack([g for g, _ in blocks], axis=1))
        los.append(sum(o for _, o in blocks))
    m1, m2 = measure(Gs, los, w, 1.0)
    d1, d2 = prop.mixed_measure(ws)
    e1, e2 = rel(afq.host(d1), m1), rel(afq.host(d2), m2)
    say(f"  mixed: OneRDM {e1:.2e}  DiagonalTwoRDM {e2:.2e}")
    # back-propagation estimator
    fields, facs = prop.bp_rings(W)
    P1 = afq.host(prop.P1())
    Vs = [orc.vhs(cs.p.Likn, fields[:, (step - (nbp - ni)) % nback], dt) for ni in range(nbp)]
    Gs, los = [], []
    for r in refs:
        Ls = [np.broadcast_to(x.conj().T, (W,) + x.conj().T.shape).copy() for x in r]
        ldR = np.zeros(W, np.complex128)
        for ni in range(nbp - 1, -1, -1):
            for sp in range(len(r)):
                Ls[sp] = orc.propagate(P1, Vs[ni], Ls[sp], 6, "H")
                if ni == 0 or ni % ortho == 0:
                    Ls[sp], ld = orc.orthogonalize(Ls[sp])
                    ldR += ld
        G = np.empty((W, len(r), cs.p.N, cs.p.N), np.complex128)
        lo = np.conj(ldR).copy()
        for sp in range(len(r)):
            for k in range(W):
                A, B = Ls[sp][k], anchors[sp][k]
                T = A.conj().T @ B
                G[k, sp] = (B @ np.linalg.solve(T, A.conj().T)).T
                sg, la = np.linalg.slogdet(T)
                lo[k] += np.log(sg) + la
        Gs.append(G)
        los.append(lo)
    wr = w * np.prod([facs[:, (step - 1 - k) % (3 * nback)] for k in range(nbp)], axis=0)
    b1, b2 = measure(Gs, los, wr, 2.0 if closed else 1.0)
    d1, d2 = prop.bp_measure(ws, nbp, ortho=ortho, restore=1)
    e3, e4 = rel(afq.host(d1), b1), rel(afq.host(d2), b2)
    say(f"  back-propagated: OneRDM {e3:.2e}  DiagonalTwoRDM {e4:.2e}")
    prop.close()
    trial.close()
    assert max(e1, e2, e3, e4) < TOL
