"""Back-propagation by operators evolved forward in time on the device
(afq_prop_te_anchor / te_propagate / te_measure; TimeEvolvedBPEstimator.hpp,
AFQMCBasePropagator::PropagateOperators, AFQMCBasePropagator.cpp:538-684).

- The inverse one-body propagator equals the oracle's exp(+1/2 H1) (1e-10) and inverts P1.
- X, Y, M after two segments of stored fields (lengths 1 and 3, the walkers orthogonalized
  after every step) equal the oracle's operator propagation and regularization through the
  same fields (1e-10).
- The OneRDM / DiagonalTwoRDM of te_measure equal the oracle's sums over the dressed density
  matrices with the path-restored weights, in the reference's orientation of M and in the
  transposed one; the transposed one equals the back-propagated estimator (bp_measure) over
  the same window.  That identity needs Y = X^{-H}, which the truncated Taylor series of
  expM(V) and expM(-V)^H satisfy only to O(|V|^(K+1) / (K+1)!) (2.6e-4 here at the default
  order 6), so these cases run the propagator at order 16.
- A 60-determinant PHMSD trial: the sum over references with their overlap coefficients.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd.synthetic import make_problem  # noqa: E402

TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    d = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (d if d > 0 else 1.0))


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import afq
    return afq, torch


def _restored(w, facs, step, k, nback):
    return w * np.prod([facs[:, (step - 1 - j) % (3 * nback)] for j in range(k)], axis=0)


def test_time_evolved_operators_vs_oracle(gpu):
    afq, torch = gpu
    p = make_problem(24, 5, 96, seed=1001)
    W, seed, nback, dt, order = 16, 29, 4, 0.005, 16
    say(f"case: time-evolved operators N=24 n=5 M=96 W={W}, segments 1 + 2 steps, order {order}")
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, p.nup)
    prop = afq.Propagator(ham, W, dt=dt, seed=seed, order=order)
    prop.enable_backprop(nback)
    osys = orc.OracleSystem(p, dt=dt)
    st, Eshift = osys.init_walkers(p.walkers(W))
    ws = afq.WalkerSet(st["Phi"])
    for k in ("weight", "phase", "eloc", "ovlp", "theta"):
        getattr(ws, k).copy_(torch.from_numpy(st[k]))
    prop.set_step(0)
    Esh = afq.device_tensor(np.array([Eshift]))
    prop.Propagate(ws, Esh)
    prop.Orthogonalize(ws)
    step = 1
    prop.te_anchor()
    prop.bp_anchor(ws)
    X0, Y0, M0, P1i = prop.te_operators()
    eye = np.broadcast_to(np.eye(p.N), X0.shape)
    assert np.array_equal(X0, eye) and np.array_equal(Y0, eye) and not M0.any()
    P1 = afq.host(prop.P1())
    P1o = orc.generate_p1(p.h1, orc.vexx(p.Likn), p.Likn, osys.vMF, dt, inverse=True)
    eP = rel(P1i, P1o)
    eI = float(np.abs(P1 @ P1i - np.eye(p.N)).max())
    say(f"  P1 inverse {eP:.2e}, |P1 P1inv - I| {eI:.2e}")
    assert eP < TOL and eI < 1e-7
    X, Y, M = X0[:, 0], Y0[:, 0], M0[:, 0]
    for seg in (1, 2):
        for _ in range(seg):
            prop.Propagate(ws, Esh)
            prop.Orthogonalize(ws)
            step += 1
        prop.te_propagate(ws, seg)
        fields, facs = prop.bp_rings(W)
        Vs = [orc.vhs(p.Likn, fields[:, (step - seg + i) % nback], dt) for i in range(seg)]
        X, Y, M = orc.propagate_operators(P1, P1i, Vs, X, Y, M, afq.host(ws.Phi), order)
        Xd, Yd, Md, _ = prop.te_operators()
        eX, eY, eM = rel(Xd[:, 0], X), rel(Yd[:, 0], Y), rel(Md[:, 0], M)
        say(f"  after {seg} more step(s): X {eX:.2e}  Y {eY:.2e}  M {eM:.2e}")
        assert max(eX, eY, eM) < TOL
    nbp = 3
    wr = _restored(afq.host(ws.weight), facs, step, nbp, nback)
    Phi = afq.host(ws.Phi)
    for tm in (False, True):
        G = orc.time_evolved_density_matrix(osys.PsiT, X, Y, M, Phi, transpose_m=tm)
        o1, o2 = orc.rdm_measure(G[:, None], wr, True)
        d1, d2 = prop.te_measure(ws, nbp, transpose_m=tm)
        e1, e2 = rel(afq.host(d1), o1), rel(afq.host(d2), o2)
        say(f"  te_measure transpose_m={tm}: OneRDM {e1:.2e}  DiagonalTwoRDM {e2:.2e}")
        assert max(e1, e2) < TOL
    # the transposed orientation is the back-propagated estimator over the same window
    b1, b2 = prop.bp_measure(ws, nbp, ortho=1, restore=1)
    e1, e2 = rel(afq.host(d1), afq.host(b1)), rel(afq.host(d2), afq.host(b2))
    say(f"  te_measure(transpose_m) vs bp_measure: {e1:.2e} {e2:.2e}")
    assert max(e1, e2) < 1e-9
    # no restoration: the plain weights
    o1, _ = orc.rdm_measure(orc.time_evolved_density_matrix(osys.PsiT, X, Y, M, Phi)[:, None],
                            afq.host(ws.weight), True)
    d1, _ = prop.te_measure(ws, 0, diag2rdm=False)
    assert rel(afq.host(d1), o1) < TOL
    prop.close()


def test_time_evolved_phmsd_vs_oracle(gpu):
    """Multi-reference dressing: every reference's G_d = M + (Ref_d^H X)^T (Gc_d conj(Y)) for
    both spins, weighted by conj(c_d) <Ref_d|Phi> normalized over the references."""
    afq, torch = gpu
    from safire_amd.phmsd import PHMSD, build_ph_struct
    from safire_amd.synthetic import make_phmsd_problem
    P = make_phmsd_problem(N=24, nel=6, M=40, ndet=60, ncore=2, nactive=8, max_rank=4, seed=21)
    W, seed, nback, nbp, dt = 8, 41, 3, 3, 0.005
    say(f"case: time-evolved operators, PHMSD N=24 nel=6 ndet=60 W={W}, window {nbp}")
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, P.h1, P.Likn, P.PsiT, P.nact, P.nact,
                                    walker_type=afq.AFQ_COLLINEAR)
    trial = PHMSD(ham, build_ph_struct(P.occa, P.occb, P.ci))
    prop = afq.Propagator(ham, W, dt=dt, seed=seed, trial=trial)
    prop.enable_backprop(nback)
    ws = afq.WalkerSet(*P.walkers(W, seed=3, noise=0.3))
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    prop.Energy(ws, E, ov)
    ws.eloc.copy_(E.sum(1))
    ws.ovlp.copy_(ov)
    Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).real.mean())]))
    prop.set_step(0)
    prop.Propagate(ws, Esh)
    prop.te_anchor()
    for _ in range(nbp):
        prop.Propagate(ws, Esh)
        prop.Orthogonalize(ws)
    step = 1 + nbp
    prop.te_propagate(ws, nbp)
    fields, facs = prop.bp_rings(W)
    P1 = afq.host(prop.P1())
    _, _, _, P1i = prop.te_operators()
    Vs = [orc.vhs(P.Likn, fields[:, (step - nbp + i) % nback], dt) for i in range(nbp)]
    I = np.broadcast_to(np.eye(P.N, dtype=complex), (W, P.N, P.N))
    Phis = (afq.host(ws.Phi), afq.host(ws.Phi_b))
    ops = [orc.propagate_operators(P1, P1i, Vs, I, I, np.zeros_like(I), Phis[sp]) for sp in range(2)]
    Xd, Yd, Md, _ = prop.te_operators()
    eO = max(rel(Xd[:, sp], ops[sp][0]) + rel(Yd[:, sp], ops[sp][1]) + rel(Md[:, sp], ops[sp][2])
             for sp in range(2))
    say(f"  X, Y, M of both spins {eO:.2e}")
    assert eO < TOL
    refs = [(P.orbs[list(a)], P.orbs[list(b)], c) for a, b, c in zip(P.occa, P.occb, P.ci)]
    lo = np.array([orc.log_overlap(ra, Phis[0]) + orc.log_overlap(rb, Phis[1])
                   for ra, rb, _ in refs])
    x = np.conj(np.array([c for _, _, c in refs]))[:, None] * np.exp(lo - lo.real.max(0))
    cf = x / x.sum(0)
    wr = _restored(afq.host(ws.weight), facs, step, nbp, nback)
    one = two = 0
    for d, (ra, rb, _) in enumerate(refs):
        G = np.stack([orc.time_evolved_density_matrix(r, *ops[sp], Phis[sp])
                      for sp, r in enumerate((ra, rb))], axis=1)
        o, t = orc.rdm_measure(G, wr, False, cf[d])
        one, two = one + o, two + t
    d1, d2 = prop.te_measure(ws, nbp)
    e1, e2 = rel(afq.host(d1), one), rel(afq.host(d2), two)
    say(f"  te_measure over 60 references: OneRDM {e1:.2e}  DiagonalTwoRDM {e2:.2e}")
    prop.close()
    trial.close()
    assert max(e1, e2) < TOL


def test_driver_schedule_matches_backpropagation(gpu):
    """run_afqmc with both back-propagation estimators over lengths {1, 2} and no
    equilibration: anchored before the first step and re-anchored at the longest length, each
    measures Steps=1 on steps 0, 2, 4 and Steps=2 on steps 1, 3, 5 (the bin counts the
    reference's estimator test expects, test_estimators.cpp:361-366), and with the walkers
    orthogonalized every step (and the Taylor series at order 16) the time-evolved estimator
    with M transposed reproduces the back-propagated one window by window."""
    afq, torch = gpu
    from safire_amd.driver import run_afqmc
    p = make_problem(24, 5, 96, seed=1001)
    W = 16
    say(f"case: driver with both back-propagation estimators, N=24 W={W}, 6 steps")
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, p.nup)
    prop = afq.Propagator(ham, W, dt=0.005, seed=7, order=16)
    ws = afq.WalkerSet(p.walkers(W))
    obs = {"onerdm": {}, "diag_twordm": {}}
    est = {"backprop": {"propagation_steps": [2, 1], "walker_ortho_interval": 1, **obs},
           "time_evolved_bp": {"propagation_steps": [1, 2], "transpose_m": True, **obs}}
    tr = run_afqmc(prop, ws, steps=6, equilibration=0, ortho_interval=1, popctl_interval=0,
                   measure_interval=1, estimators=est)
    worst = 0.0
    for L in (1, 2):
        for name in ("OneRDM", "DiagonalTwoRDM"):
            te, bp = tr.time_evolved_bp[L][name], tr.backprop[L][name]
            assert len(te) == len(bp) == 3, (L, name, len(te), len(bp))
            worst = max([worst] + [rel(a, b) for a, b in zip(te, bp)])
    say(f"  3 + 3 windows, time-evolved vs back-propagated {worst:.2e}")
    prop.close()
    assert worst < 1e-9


def test_time_evolved_collinear_vs_backpropagation(gpu):
    """COLLINEAR walkers with different up and down trials (5 + 4 electrons): both spins'
    operators against the oracle, and the transposed orientation against bp_measure over the
    same window (Taylor order 16)."""
    afq, torch = gpu
    p = make_problem(24, 5, 96, seed=1003)
    nup, ndn, W, seed, nback, nbp, order = 5, 4, 8, 37, 4, 3, 16
    say(f"case: time-evolved operators, COLLINEAR N=24 nup={nup} ndown={ndn} W={W}, window {nbp}")
    rng = np.random.default_rng(2)
    evecs = np.linalg.eigh(p.h1)[1]
    Q = np.linalg.qr(rng.normal(size=(6, 6)))[0]
    dn = evecs[:, :6] @ Q[:, :ndn]
    PsiT = np.ascontiguousarray(np.vstack([p.PsiT, dn.T]).astype(np.complex128))
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, PsiT, nup, ndn, walker_type=afq.AFQ_COLLINEAR)
    prop = afq.Propagator(ham, W, dt=0.005, seed=seed, order=order)
    prop.enable_backprop(nback)
    Pa = np.broadcast_to(PsiT[:nup].conj().T, (W, p.N, nup)) + 0.05 * (
        rng.normal(size=(W, p.N, nup)) + 1j * rng.normal(size=(W, p.N, nup)))
    Pb = np.broadcast_to(PsiT[nup:].conj().T, (W, p.N, ndn)) + 0.05 * (
        rng.normal(size=(W, p.N, ndn)) + 1j * rng.normal(size=(W, p.N, ndn)))
    ws = afq.WalkerSet(Pa, Pb)
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    prop.Energy(ws, E, ov)
    ws.eloc.copy_(E.sum(1))
    ws.ovlp.copy_(ov)
    Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).real.mean())]))
    prop.set_step(0)
    prop.Propagate(ws, Esh)
    prop.Orthogonalize(ws)
    prop.bp_anchor(ws)
    prop.te_anchor()
    for _ in range(nbp):
        prop.Propagate(ws, Esh)
        prop.Orthogonalize(ws)
    step = 1 + nbp
    prop.te_propagate(ws, nbp)
    fields, facs = prop.bp_rings(W)
    P1 = afq.host(prop.P1())
    Xd, Yd, Md, P1i = prop.te_operators()
    Vs = [orc.vhs(p.Likn, fields[:, (step - nbp + i) % nback], 0.005) for i in range(nbp)]
    I = np.broadcast_to(np.eye(p.N, dtype=complex), (W, p.N, p.N))
    Phis = (afq.host(ws.Phi), afq.host(ws.Phi_b))
    ops = [orc.propagate_operators(P1, P1i, Vs, I, I, np.zeros_like(I), Phis[sp], order)
           for sp in range(2)]
    eO = max(max(rel(Xd[:, sp], ops[sp][0]), rel(Yd[:, sp], ops[sp][1]), rel(Md[:, sp], ops[sp][2]))
             for sp in range(2))
    wr = _restored(afq.host(ws.weight), facs, step, nbp, nback)
    G = np.stack([orc.time_evolved_density_matrix(r, *ops[sp], Phis[sp])
                  for sp, r in enumerate((PsiT[:nup], PsiT[nup:]))], axis=1)
    o1, o2 = orc.rdm_measure(G, wr, False)
    d1, d2 = prop.te_measure(ws, nbp)
    eM = max(rel(afq.host(d1), o1), rel(afq.host(d2), o2))
    t1, t2 = prop.te_measure(ws, nbp, transpose_m=True)
    b1, b2 = prop.bp_measure(ws, nbp, ortho=1, restore=1)
    eB = max(rel(afq.host(t1), afq.host(b1)), rel(afq.host(t2), afq.host(b2)))
    say(f"  X, Y, M {eO:.2e}; te_measure {eM:.2e}; transposed vs bp_measure {eB:.2e}")
    prop.close()
    assert eO < TOL and eM < TOL and eB < 1e-9
