"""SAFIRE's walker buffer passed in place through the C ABI.

The reference stores a walker set as one complex row per walker, [Phi_a | Phi_b | WEIGHT, PHASE,
PSEUDO_ELOC_, E1_, EXX_, EJ_, OVLP, LOGSCL_UP, LOGSCL_DN, IS_UNITARY, THETA]
(WalkerSetBase.cpp:75-121); SlaterMatrices(s) and getProperty are views strided by the row
(WalkerSetBase.h:256-276, 463-481).  afq.WalkerBuffer builds that layout (plus padding, so no
code can assume the row length) and passes the views with afq_walkers.phi_stride /
prop_stride.  Every result must be bit-identical to the same walkers in contiguous arrays:
the step with injected and in-kernel fields, QR, comb population control, the captured step,
energies, the det_ops entry points and the back-propagation estimator, for CLOSED, COLLINEAR and
PHMSD walkers.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd.synthetic import make_phmsd_problem, make_problem, random_walkers  # noqa: E402

PROPS = ("Phi", "weight", "phase", "eloc", "ovlp", "theta")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import afq
    return afq, torch


def _same(afq, a, b, names=PROPS):
    for k in names:
        x, y = afq.host(getattr(a, k)), afq.host(getattr(b, k))
        assert np.array_equal(x, y), k
    if a.Phi_b is not None:
        assert np.array_equal(afq.host(a.Phi_b), afq.host(b.Phi_b)), "Phi_b"


def _pair(afq, torch, prop, Pa, Pb=None):
    """The same walkers contiguous and in the reference's row layout (3 padding elements)."""
    ws = afq.WalkerSet(Pa, Pb)
    wb = afq.WalkerBuffer(Pa, Pb, pad=3)
    W = Pa.shape[0]
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    prop.Energy(ws, E, ov)
    for w in (ws, wb):
        w.eloc.copy_(E.sum(1))
        w.ovlp.copy_(ov)
        w.theta.copy_(0.01 * torch.arange(W, dtype=torch.float64))
    return ws, wb


def _drive(afq, torch, prop, ws, wb, M, Esh_value):
    W = ws.W
    Esh = {id(w): afq.device_tensor(np.array([Esh_value])) for w in (ws, wb)}
    E = {id(w): afq.empty((W, 3), torch.complex128) for w in (ws, wb)}
    ov = {id(w): afq.empty(W, torch.complex128) for w in (ws, wb)}
    for step in range(3):
        u = afq.device_tensor(orc.field_uniforms(9, 1, step, W, M))
        for w in (ws, wb):
            prop.set_step(step)
            prop.Propagate(w, Esh[id(w)], u=u)
            if step == 1:
                prop.Orthogonalize(w)
                prop.population_control(w, 0.37)
        _same(afq, ws, wb)
    e_out = afq.empty(1, torch.float64)
    for step in range(3, 9):
        for w in (ws, wb):
            prop.set_step(step)
            prop.step_graph(w, Esh[id(w)], ortho=(step % 3 == 0), popctl=(step % 4 == 0),
                            relax=0.1, energy_out=e_out)
        _same(afq, ws, wb)
        assert np.array_equal(afq.host(Esh[id(ws)]), afq.host(Esh[id(wb)]))
    for w in (ws, wb):
        prop.Energy(w, E[id(w)], ov[id(w)])
    assert np.array_equal(afq.host(E[id(ws)]), afq.host(E[id(wb)]))
    assert np.array_equal(afq.host(ov[id(ws)]), afq.host(ov[id(wb)]))


def test_closed_walker_buffer(gpu):
    afq, torch = gpu
    say("case: walker buffer CLOSED N=24 n=5 M=96 W=40")
    p = make_problem(24, 5, 96, seed=1001)
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, p.nup)
    prop = afq.Propagator(ham, 40, dt=0.005, seed=5)
    ws, wb = _pair(afq, torch, prop, random_walkers(p, 40, seed=3))
    _drive(afq, torch, prop, ws, wb, p.M, -10.0)
    # back-propagation over strided walkers (the estimator packs them)
    prop2 = afq.Propagator(ham, 40, dt=0.005, seed=6)
    prop2.enable_backprop(4)
    res = []
    for w in (ws, wb):
        prop2.set_step(0)
        prop2.bp_anchor(w)
        Esh = afq.device_tensor(np.array([-10.0]))
        for k in range(4):
            prop2.step_graph(w, Esh, ortho=(k == 2))
        G, wt = prop2.backpropagate(w, 4, ortho=2, restore=1)
        res.append((afq.host(G), afq.host(wt)))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_collinear_walker_buffer(gpu):
    afq, torch = gpu
    say("case: walker buffer COLLINEAR N=24 n=5+4 M=96 W=33")
    p = make_problem(24, 5, 96, seed=1001)
    PsiT = np.concatenate([p.PsiT, p.PsiT[:4]])
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, PsiT, 5, 4, walker_type=afq.AFQ_COLLINEAR)
    prop = afq.Propagator(ham, 33, dt=0.005, seed=7)
    Pa = random_walkers(p, 33, seed=4)
    Pb = np.ascontiguousarray(random_walkers(p, 33, seed=5)[:, :, :4])
    ws, wb = _pair(afq, torch, prop, Pa, Pb)
    _drive(afq, torch, prop, ws, wb, p.M, -10.0)


def test_phmsd_walker_buffer(gpu):
    afq, torch = gpu
    from safire_amd.phmsd import PHMSD, build_ph_struct
    say("case: walker buffer PHMSD N=24 n=5+5 M=48 60 determinants W=24")
    P = make_phmsd_problem(N=24, nel=5, M=48, ndet=60, ncore=2, nactive=6, max_rank=3, seed=77)
    ph = build_ph_struct(P.occa, P.occb, P.ci)
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, P.h1, P.Likn, P.PsiT, P.nact, P.nact,
                                    walker_type=afq.AFQ_COLLINEAR)
    trial = PHMSD(ham, ph)
    prop = afq.Propagator(ham, 24, dt=0.005, seed=8, trial=trial)
    Pa, Pb = P.walkers(24, seed=2, noise=0.1)
    ws, wb = _pair(afq, torch, prop, Pa, Pb)
    _drive(afq, torch, prop, ws, wb, P.M, -5.0)
    G = {}
    for w in (ws, wb):
        G[id(w)] = afq.empty((24, 2 * P.nact, P.N), torch.complex128)
        Ov = afq.empty(24, torch.complex128)
        trial.MixedDensityMatrix(w, G[id(w)], Ov)
    assert np.array_equal(afq.host(G[id(ws)]), afq.host(G[id(wb)]))


def test_det_ops_strided(gpu):
    afq, torch = gpu
    say("case: det_ops on strided Slater matrices N=37 n=6 W=11")
    p = make_problem(37, 6, 8, seed=12)
    Phi = random_walkers(p, 11, seed=6)
    ctx = afq.Context(0)
    dP = afq.device_tensor(p.PsiT)
    wb = afq.WalkerBuffer(Phi, pad=5)
    dc = afq.device_tensor(Phi)
    for compact in (True, False):
        shape = (11, 6, 37) if compact else (11, 37, 37)
        Ga, Gb = afq.empty(shape, torch.complex128), afq.empty(shape, torch.complex128)
        oa, ob = afq.zeros(11, torch.complex128), afq.zeros(11, torch.complex128)
        afq.det_ops.MixedDensityMatrix(ctx, dP, dc, Ga, oa, compact=compact)
        afq.det_ops.MixedDensityMatrix(ctx, dP, wb.Phi, Gb, ob, compact=compact)
        assert np.array_equal(afq.host(Ga), afq.host(Gb))
        assert np.array_equal(afq.host(oa), afq.host(ob))
    rng = np.random.default_rng(1)
    V = afq.device_tensor(0.05 * (rng.normal(size=(11, 37, 37)) + 1j * rng.normal(size=(11, 37, 37))))
    P1 = afq.device_tensor(np.eye(37) + 0.01 * rng.normal(size=(37, 37)) + 0j)
    for prec in (afq.AFQ_FP64, afq.AFQ_MIXED):
        Vp = V if prec == afq.AFQ_FP64 else V.to(torch.complex64)
        afq.det_ops.Propagate(ctx, P1, Vp, dc, precision=prec)
        afq.det_ops.Propagate(ctx, P1, Vp, wb.Phi, precision=prec)
        assert np.array_equal(afq.host(dc), afq.host(wb.Phi)), prec
    la, lb = afq.zeros(11, torch.complex128), afq.zeros(11, torch.complex128)
    afq.det_ops.orthogonalize(ctx, dc, la)
    afq.det_ops.orthogonalize(ctx, wb.Phi, lb)
    assert np.array_equal(afq.host(dc), afq.host(wb.Phi))
    assert np.array_equal(afq.host(la), afq.host(lb))
    oa, ob = afq.zeros(11, torch.complex128), afq.zeros(11, torch.complex128)
    afq.det_ops.Log_Overlap(ctx, dP, dc, oa)
    afq.det_ops.Log_Overlap(ctx, dP, wb.Phi, ob)
    assert np.array_equal(afq.host(oa), afq.host(ob))
    # the padding after each row is never touched
    pad = afq.host(wb.buf)[:, -5:]
    assert np.all(pad == 0)
