"""Parity of the HIP walker step with the FP64 oracle, through the C ABI.

FP64 bar: every array agrees with the oracle to 1e-10 relative, measured per walker (each
walker's max |gpu - oracle| over its own max |oracle|, floored at 1e-6 of the array's) as well as
norm-wise per array, on the same seeded inputs.  Integer-exact outputs (the
Philox uniforms) are compared bit for bit.  Mixed precision: per-walker local energies within
0.1 mHa of the FP64 path from the same walker state.
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd.synthetic import make_problem, random_walkers  # noqa: E402

TOL = 1e-10
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sdet_ops.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import afq
    ctx = afq.Context(0)
    return afq, ctx, torch


def rel(a, b):
    """Relative error over the finite entries (NaN/inf positions must agree): the larger of
    the norm-wise max |a - b| / max |b| and the per-walker one, max over the leading (walker)
    axis of max |a_w - b_w| / max |b_w|, where a walker's scale is floored at 1e-6 max |b| so
    that entries that are zero by construction do not turn rounding into a relative error."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    fa, fb = np.isfinite(a), np.isfinite(b)
    assert np.array_equal(fa, fb), "non-finite entries differ"
    if not fb.any():
        return 0.0
    d = np.max(np.abs(b[fb]))
    d = d if d > 0 else 1.0
    diff = np.abs(np.where(fb, a, 0) - np.where(fb, b, 0))
    mag = np.abs(np.where(fb, b, 0))
    if a.ndim == 0:
        return float(diff / d)
    dw = diff.reshape(len(diff), -1).max(axis=1)
    sw = np.maximum(mag.reshape(len(mag), -1).max(axis=1), 1e-6 * d)
    return float(max(dw.max() / d, np.max(dw / sw)))


def cplx(a):
    a = np.asarray(a, dtype=np.float64)
    return a[..., 0] + 1j * a[..., 1]


# ---- det_ops ---------------------------------------------------------------------------------

def test_golden_sdet_ops_on_gpu(gpu):
    afq, ctx, torch = gpu
    say("case: golden sdet ops (test_sdet_ops.cpp values) W=5 N=4 n=3")
    A = cplx(GOLD["A"])
    B = np.broadcast_to(cplx(GOLD["B"]), (5, 4, 3)).copy()
    dA, dB = afq.device_tensor(A), afq.device_tensor(B)
    ov = afq.zeros(5, torch.complex128)
    afq.det_ops.Log_Overlap(ctx, dA, dB, ov)
    ref = np.log(complex(*GOLD["ov_base_arg"]))
    assert rel(afq.host(ov), np.full(5, ref)) < 1e-12
    G = afq.empty((5, 4, 4), torch.complex128)
    ov.zero_()
    afq.det_ops.MixedDensityMatrix(ctx, dA, dB, G, ov, compact=False)
    assert rel(afq.host(G), np.broadcast_to(cplx(GOLD["v_ref"]), (5, 4, 4))) < 1e-12
    Gc = afq.empty((5, 3, 4), torch.complex128)
    ov.zero_()
    afq.det_ops.MixedDensityMatrix(ctx, dA, dB, Gc, ov, compact=True)
    assert rel(afq.host(Gc), np.broadcast_to(cplx(GOLD["vc_ref"]), (5, 3, 4))) < 1e-12
    assert rel(afq.host(ov), np.full(5, ref)) < 1e-12
    # sub-block (A_ = A[0:2,0:3], B_ = B[:,0:3,0:2])
    dA2 = afq.device_tensor(A[0:2, 0:3].copy())
    dB2 = afq.device_tensor(B[:, 0:3, 0:2].copy())
    G2 = afq.empty((5, 3, 3), torch.complex128)
    ov2 = afq.zeros(5, torch.complex128)
    afq.det_ops.MixedDensityMatrix(ctx, dA2, dB2, G2, ov2, compact=False)
    assert rel(afq.host(G2), np.broadcast_to(cplx(GOLD["v_ref_2"]), (5, 3, 3))) < 1e-12
    assert rel(afq.host(ov2), np.full(5, np.log(complex(*GOLD["ov2_base_arg"])))) < 1e-12


def test_golden_inverse_logdet_on_gpu(gpu):
    afq, ctx, torch = gpu
    say("case: golden inverse_logdet 5 x (4x4)")
    t = GOLD["inverse_test"]
    A = np.broadcast_to(cplx(t["A"]), (5, 4, 4)).copy()
    dA = afq.device_tensor(A)
    inv = afq.empty((5, 4, 4), torch.complex128)
    ov = afq.zeros(5, torch.complex128)
    afq.det_ops.inverse_logdet(ctx, dA, ov, inv)
    assert np.max(np.abs(afq.host(inv) - cplx(t["inv_ref"]))) < 1e-7
    assert rel(afq.host(ov), np.full(5, complex(*t["logdet_ref"]))) < 1e-12
    ref_inv, ref_ov = orc.inverse_logdet(A)
    assert rel(afq.host(inv), ref_inv) < 1e-13


@pytest.mark.parametrize("n", [3, 12, 40])
def test_inverse_logdet_singular_and_random(gpu, n):
    """getri_or_zero: a singular matrix gives a zero inverse; regular ones match the oracle."""
    afq, ctx, torch = gpu
    say(f"case: inverse_logdet n={n} (one singular matrix in the batch)")
    rng = np.random.default_rng(n)
    A = rng.normal(size=(6, n, n)) + 1j * rng.normal(size=(6, n, n))
    A[2, 1, :] = 0.0  # singular
    dA = afq.device_tensor(A)
    inv = afq.empty((6, n, n), torch.complex128)
    ov = afq.zeros(6, torch.complex128)
    afq.det_ops.inverse_logdet(ctx, dA, ov, inv)
    ref_inv, ref_ov = orc.inverse_logdet(A)
    gi, go = afq.host(inv), afq.host(ov)
    assert np.all(gi[2] == 0) and np.all(ref_inv[2] == 0)
    assert go[2].real < -600
    keep = [0, 1, 3, 4, 5]
    assert rel(gi[keep], ref_inv[keep]) < TOL
    assert rel(go[keep], ref_ov[keep]) < TOL


@pytest.mark.parametrize("N,n,W", [(24, 5, 17), (108, 24, 33), (200, 40, 9)])
def test_det_ops_vs_oracle(gpu, N, n, W):
    afq, ctx, torch = gpu
    say(f"case: det_ops N={N} n={n} W={W}")
    p = make_problem(N, n, 8, seed=N + n)
    Phi = random_walkers(p, W, seed=5)
    dP, dPhi = afq.device_tensor(p.PsiT), afq.device_tensor(Phi)
    ov = afq.zeros(W, torch.complex128)
    G = afq.empty((W, n, N), torch.complex128)
    afq.det_ops.MixedDensityMatrix(ctx, dP, dPhi, G, ov, compact=True)
    Gr, ovr = orc.mixed_density_matrix(p.PsiT, Phi, compact=True)
    assert rel(afq.host(G), Gr) < TOL
    assert rel(afq.host(ov), ovr) < TOL
    Gf = afq.empty((W, N, N), torch.complex128)
    ov.zero_()
    afq.det_ops.MixedDensityMatrix(ctx, dP, dPhi, Gf, ov, compact=False)
    Gfr, _ = orc.mixed_density_matrix(p.PsiT, Phi, compact=False)
    assert rel(afq.host(Gf), Gfr) < TOL
    ov.zero_()
    afq.det_ops.Log_Overlap(ctx, dP, dPhi, ov)
    assert rel(afq.host(ov), orc.log_overlap(p.PsiT, Phi)) < TOL


@pytest.mark.parametrize("N,n,W", [(24, 5, 7), (108, 24, 16)])
def test_propagate_vs_oracle(gpu, N, n, W):
    afq, ctx, torch = gpu
    say(f"case: propagate N={N} n={n} W={W}")
    rng = np.random.default_rng(N)
    p = make_problem(N, n, 8, seed=3)
    Phi = random_walkers(p, W, seed=2)
    V = 0.05 * (rng.normal(size=(W, N, N)) + 1j * rng.normal(size=(W, N, N)))
    P1 = (np.eye(N) + 0.01 * rng.normal(size=(N, N))) + 0.003j * rng.normal(size=(N, N))
    for TA in ("N", "H"):
        dPhi = afq.device_tensor(Phi)
        afq.det_ops.Propagate(ctx, afq.device_tensor(P1), afq.device_tensor(V), dPhi, order=6, TA=TA)
        ref = orc.propagate(P1, V, Phi, order=6, TA=TA)
        assert rel(afq.host(dPhi), ref) < TOL, TA


@pytest.mark.parametrize("N,n,W", [(24, 5, 7), (108, 24, 16), (200, 40, 4)])
def test_orthogonalize_vs_oracle(gpu, N, n, W):
    afq, ctx, torch = gpu
    say(f"case: orthogonalize N={N} n={n} W={W}")
    p = make_problem(N, n, 4, seed=9)
    Phi = random_walkers(p, W, seed=4, noise=1.0)
    dPhi = afq.device_tensor(Phi)
    ld = afq.zeros(W, torch.complex128)
    afq.det_ops.orthogonalize(ctx, dPhi, ld)
    Q, ldr = orc.orthogonalize(Phi)
    assert rel(afq.host(dPhi), Q) < TOL
    assert rel(afq.host(ld), ldr) < TOL


# ---- hamiltonian operations ------------------------------------------------------------------

@pytest.fixture(scope="module", params=[(24, 5, 96), (108, 24, 600)], ids=["N24_n5_M96", "N108_n24_M600"])
def system(request, gpu):
    afq, ctx, torch = gpu
    N, n, M = request.param
    say(f"case: hamiltonian setup N={N} n={n} M={M}")
    p = make_problem(N, n, M, seed=1000 + N)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, n, precision=afq.AFQ_MIXED)
    osys = orc.OracleSystem(p, dt=0.005)
    return p, ham, osys


def test_setup_vs_oracle(gpu, system):
    afq, ctx, torch = gpu
    p, ham, osys = system
    L, haj, vexx = ham.tensors()
    assert rel(L.reshape(osys.Lnak.shape), osys.Lnak) < TOL
    assert rel(haj, osys.haj) < TOL
    assert rel(vexx, osys.vexx) < TOL
    vmf = afq.empty(p.M, torch.complex128)
    ham.vMF(0.005, vmf)
    assert rel(afq.host(vmf), osys.vMF) < TOL
    P1 = afq.empty((p.N, p.N), torch.complex128)
    ham.getOneBodyPropagator(vmf, 0.005, P1)
    assert rel(afq.host(P1), osys.P1) < TOL


def test_vbias_vhs_energy_vs_oracle(gpu, system):
    afq, ctx, torch = gpu
    p, ham, osys = system
    W = 13
    Phi = random_walkers(p, W, seed=8)
    Gr, _ = orc.mixed_density_matrix(p.PsiT, Phi)
    dG = afq.device_tensor(Gr)
    v = afq.empty((W, p.M), torch.complex128)
    ham.vbias(dG, v, 0.005, precision=afq.AFQ_FP64)
    assert rel(afq.host(v), orc.vbias(osys.Lnak, Gr, 0.005)) < TOL
    rng = np.random.default_rng(1)
    X = rng.normal(size=(W, p.M)) + 1j * rng.normal(size=(W, p.M))
    V = afq.empty((W, p.N, p.N), torch.complex128)
    ham.vHS(afq.device_tensor(X), V, 0.005, precision=afq.AFQ_FP64)
    assert rel(afq.host(V), orc.vhs(p.Likn, X, 0.005)) < TOL
    E = afq.empty((W, 3), torch.complex128)
    ham.energy(E, dG, precision=afq.AFQ_FP64)
    Er = orc.energy(osys.haj, osys.Lnak, Gr, p.nup)
    for col in range(3):
        assert rel(afq.host(E)[:, col], Er[:, col]) < TOL, col


def test_mixed_contractions_close_to_fp64(gpu, system):
    afq, ctx, torch = gpu
    p, ham, osys = system
    W = 13
    Phi = random_walkers(p, W, seed=8)
    Gr, _ = orc.mixed_density_matrix(p.PsiT, Phi)
    dG = afq.device_tensor(Gr)
    v = afq.empty((W, p.M), torch.complex128)
    ham.vbias(dG, v, 0.005, precision=afq.AFQ_MIXED)
    assert rel(afq.host(v), orc.vbias(osys.Lnak, Gr, 0.005)) < 1e-5
    E = afq.empty((W, 3), torch.complex128)
    ham.energy(E, dG, precision=afq.AFQ_MIXED)
    Er = orc.energy(osys.haj, osys.Lnak, Gr, p.nup)
    assert np.max(np.abs(afq.host(E).sum(1) - Er.sum(1))) < 1e-4  # 0.1 mHa


@pytest.mark.parametrize("sym", [True, False], ids=["symmetric-L", "general-L"])
def test_vhs_paths_vs_oracle(gpu, sym):
    """vHS with symmetric Cholesky vectors (triangle + mirror) and with a general tensor."""
    afq, ctx, torch = gpu
    N, n, M, W = 37, 6, 50, 9
    say(f"case: vHS N={N} M={M} W={W} symmetric={sym}")
    p = make_problem(N, n, M, seed=21)
    Likn = p.Likn.copy()
    if not sym:
        Likn += 0.01 * np.random.default_rng(2).normal(size=Likn.shape)
    ham = afq.HamiltonianOperations(ctx, p.h1, Likn, p.PsiT, n, precision=afq.AFQ_MIXED)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(W, M)) + 1j * rng.normal(size=(W, M))
    V = afq.empty((W, N, N), torch.complex128)
    ham.vHS(afq.device_tensor(X), V, 0.01, precision=afq.AFQ_FP64)
    Vr = orc.vhs(Likn, X, 0.01)
    assert rel(afq.host(V), Vr) < TOL
    V32 = afq.empty((W, N, N), torch.complex64)
    ham.vHS(afq.device_tensor(X.astype(np.complex64)), V32, 0.01, precision=afq.AFQ_MIXED)
    assert rel(afq.host(V32).astype(np.complex128), Vr) < 1e-5


def test_mixed_vhs_and_propagate_close_to_fp64(gpu, system):
    afq, ctx, torch = gpu
    p, ham, osys = system
    W = 5
    rng = np.random.default_rng(3)
    X = rng.normal(size=(W, p.M)) + 1j * rng.normal(size=(W, p.M))
    V32 = afq.empty((W, p.N, p.N), torch.complex64)
    ham.vHS(afq.device_tensor(X.astype(np.complex64)), V32, 0.005, precision=afq.AFQ_MIXED)
    Vr = orc.vhs(p.Likn, X, 0.005)
    ev = rel(afq.host(V32).astype(np.complex128), Vr)
    say(f"mixed vHS rel err {ev:.2e}")
    assert ev < 1e-5
    Phi = random_walkers(p, W, seed=1)
    P1 = osys.P1
    ref = orc.propagate(P1, Vr, Phi)
    dPhi = afq.device_tensor(Phi)
    afq.det_ops.Propagate(ctx, afq.device_tensor(P1), afq.device_tensor(Vr.astype(np.complex64)),
                          dPhi, precision=afq.AFQ_MIXED)
    ep = rel(afq.host(dPhi), ref)
    say(f"mixed propagate rel err {ep:.2e}")
    assert ep < 1e-5


# ---- fields and weights ----------------------------------------------------------------------

def test_philox_uniforms_bit_exact(gpu):
    afq, ctx, torch = gpu
    say("case: philox uniforms W=37 M=129")
    u = afq.empty((37, 129), torch.float64)
    afq.field_uniforms(ctx, 0x1234_5678_9ABC, 1, 17, u)
    ref = orc.field_uniforms(0x1234_5678_9ABC, 1, 17, 37, 129)
    assert np.array_equal(afq.host(u), ref)


def test_construct_x_and_update_vs_oracle(gpu):
    afq, ctx, torch = gpu
    say("case: construct_x + walker_update W=19 M=41")
    rng = np.random.default_rng(6)
    W, M = 19, 41
    u = rng.uniform(size=(W, M))
    vb = rng.normal(size=(W, M)) + 1j * rng.normal(size=(W, M))
    vb[0, 0] = 70 + 30j
    vmf = rng.normal(size=M).astype(complex)
    dX = afq.device_tensor(vb)
    HW, mf = afq.empty(W, torch.complex128), afq.empty(W, torch.complex128)
    afq.construct_x(ctx, afq.device_tensor(u), afq.device_tensor(vmf), dX, HW, mf)
    Xr, HWr, mfr = orc.construct_x(u, vb, vmf)
    assert rel(afq.host(dX), Xr) < TOL
    assert rel(afq.host(HW), HWr) < TOL
    assert rel(afq.host(mf), mfr) < TOL
    # walker update on the same inputs, including a non-finite overlap and a phase kill
    st = dict(weight=(1 + 0.1 * rng.normal(size=W)).astype(complex), phase=np.ones(W, complex),
              eloc=rng.normal(size=W) - 3 + 0j, ovlp=0.1 * rng.normal(size=W) + 0j,
              theta=np.zeros(W, complex))
    new = 0.02 * rng.normal(size=W) + 0.05j * rng.normal(size=W)
    new[3] = np.nan
    new[5] = 3.0j
    dst = {k: afq.device_tensor(v) for k, v in st.items()}
    Esh = afq.device_tensor(np.array([-3.0]))
    prm = afq.UpdateParams(0.005, 0.7, 1.0, 10.0, 100.0, 0.1, W, 0, 0)
    afq.walker_update(ctx, prm, Esh, afq.device_tensor(new), mf, HW, dst["weight"], dst["phase"],
                      dst["eloc"], dst["ovlp"], dst["theta"])
    orc.walker_update(st, 0.005, -3.0, 0.7, new, mfr, HWr, target_pop=W)
    for k in st:
        assert rel(afq.host(dst[k]), st[k]) < TOL, k
    # growth estimator on a regular population
    w = (1 + 0.2 * rng.normal(size=W)) * np.exp(0.3j * rng.normal(size=W))
    e = -3 + 0.5 * rng.normal(size=W) + 0.1j * rng.normal(size=W)
    out = afq.empty(1, torch.float64)
    afq.average_pseudo_energy(ctx, afq.device_tensor(w), afq.device_tensor(e), 0.005, out)
    assert abs(afq.host(out)[0] - orc.average_pseudo_energy(w, e, 0.005)) < 1e-9


# ---- full step -------------------------------------------------------------------------------

def _run_pair(afq, torch, p, W, steps, seed=11, ortho_every=5, precision=0, local_energy=False):
    """Advance W walkers `steps` steps on the GPU and in the oracle from the same state with the
    same injected uniforms; yields (step, gpu_state, oracle_state) after every step."""
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, p.nup, precision=precision)
    prop = afq.Propagator(ham, W, dt=0.005, local_energy=local_energy)
    osys = orc.OracleSystem(p, dt=0.005)
    st, Eshift = osys.init_walkers(p.walkers(W))
    ws = afq.WalkerSet(st["Phi"])
    for k in ("weight", "phase", "eloc", "ovlp", "theta"):
        getattr(ws, k).copy_(torch.from_numpy(st[k]))
    Esh = afq.device_tensor(np.array([Eshift]))
    for step in range(steps):
        u = orc.field_uniforms(seed, 1, step, W, p.M)
        prop.Propagate(ws, Esh, u=afq.device_tensor(u))
        osys.step(st, u, Eshift, local_energy=local_energy)
        if step % ortho_every == 0:
            prop.Orthogonalize(ws)
            osys.orthogonalize(st)
        yield step, ws, st, ctx


def test_step_vs_oracle_N24(gpu):
    afq, ctx, torch = gpu
    p = make_problem(24, 5, 96, seed=1001)
    say("case: full step N=24 n=5 M=96 W=64, 10 steps, QR at 0 and 5")
    for step, ws, st, c in _run_pair(afq, torch, p, 64, 10):
        assert rel(afq.host(ws.Phi), st["Phi"]) < TOL, step
        for k in ("ovlp", "eloc", "weight", "phase", "theta"):
            assert rel(afq.host(getattr(ws, k)), st[k]) < TOL, (step, k)


def test_local_energy_step_vs_oracle(gpu):
    """hybrid = false: the local energy of the propagated walkers (Energy instead of
    Log_Overlap) drives the weights with the midpoint rule."""
    afq, ctx, torch = gpu
    p = make_problem(24, 5, 96, seed=1001)
    say("case: local-energy step N=24 n=5 M=96 W=32, 6 steps, QR at 0 and 5")
    for step, ws, st, c in _run_pair(afq, torch, p, 32, 6, local_energy=True):
        assert rel(afq.host(ws.Phi), st["Phi"]) < TOL, step
        for k in ("ovlp", "eloc", "weight", "phase", "theta"):
            assert rel(afq.host(getattr(ws, k)), st[k]) < TOL, (step, k)


def test_step_vs_oracle_N108(gpu):
    afq, ctx, torch = gpu
    p = make_problem(108, 24, 600, seed=1002)
    say("case: full step N=108 n=24 M=600 W=16, 2 steps")
    for step, ws, st, c in _run_pair(afq, torch, p, 16, 2):
        assert rel(afq.host(ws.Phi), st["Phi"]) < TOL, step
        for k in ("ovlp", "eloc", "weight", "phase"):
            assert rel(afq.host(getattr(ws, k)), st[k]) < TOL, (step, k)


def test_mixed_step_energy_error(gpu):
    """One mixed-precision step from the FP64 state: per-walker local energies (computed in FP64
    from the propagated walkers) against the FP64 step.

    The synthetic hamiltonians have local energies of 1e3-1e5 Ha after one step (the exchange
    term alone is ~ -1.4e3 Ha at N=108), and E_loc responds to a relative change eps of the walker
    matrices with |dE| ~ eps * |E| * cond(Psi^H Phi).  FP32 propagation (unit roundoff 6e-8) can
    therefore reach only a relative deviation of ~1e-6-1e-5; the bound below is relative, and
    the absolute deviation is printed (it is reported, not asserted, against 0.1 mHa)."""
    afq, ctx, torch = gpu
    p = make_problem(108, 24, 600, seed=1002)
    W = 32
    say("case: mixed vs FP64 one step N=108 n=24 M=600 W=32")
    res = {}
    for prec in (afq.AFQ_FP64, afq.AFQ_MIXED):
        c = afq.Context(0)
        ham = afq.HamiltonianOperations(c, p.h1, p.Likn, p.PsiT, p.nup, precision=prec)
        prop = afq.Propagator(ham, W, dt=0.005)
        ws = afq.WalkerSet(p.walkers(W))
        E = afq.empty((W, 3), torch.complex128)
        ov = afq.empty(W, torch.complex128)
        prop.Energy(ws, E, ov)
        ws.eloc.copy_(E.sum(1))
        ws.ovlp.copy_(ov)
        Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).real.mean())]))
        u = orc.field_uniforms(3, 1, 0, W, p.M)
        prop.Propagate(ws, Esh, u=afq.device_tensor(u))
        phis = afq.host(ws.Phi)
        prop.Energy(ws, E, ov)
        res[prec] = afq.host(E.sum(1))
        res[("phi", prec)] = phis
    say(f"mixed vs fp64 Phi rel err {rel(res[('phi', 1)], res[('phi', 0)]):.3e}")
    err = np.abs(res[afq.AFQ_MIXED] - res[afq.AFQ_FP64])
    relerr = err / np.abs(res[afq.AFQ_FP64])
    say(f"mixed vs fp64 per-walker energy: max |dE| {err.max():.3e} Ha, max |dE|/|E| "
        f"{relerr.max():.3e}, median |E| {np.median(np.abs(res[afq.AFQ_FP64])):.3e} Ha")
    assert relerr.max() < 1e-4


def test_graph_replay_matches_eager(gpu):
    afq, ctx, torch = gpu
    p = make_problem(24, 5, 96, seed=1001)
    W = 32
    say("case: hipGraph replay vs eager, N=24 n=5 M=96 W=32, 7 steps")
    out = {}
    for mode in ("eager", "graph"):
        c = afq.Context(0)
        ham = afq.HamiltonianOperations(c, p.h1, p.Likn, p.PsiT, p.nup)
        prop = afq.Propagator(ham, W, dt=0.005, seed=77)
        ws = afq.WalkerSet(p.walkers(W))
        E = afq.empty((W, 3), torch.complex128)
        ov = afq.empty(W, torch.complex128)
        prop.Energy(ws, E, ov)
        ws.eloc.copy_(E.sum(1))
        ws.ovlp.copy_(ov)
        Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).real.mean())]))
        e_out = afq.empty(1, torch.float64)
        relax = 1 - np.exp(-0.1)
        for step in range(7):
            ortho = step % 5 == 0
            if mode == "eager":
                prop.Propagate(ws, Esh)
                if ortho:
                    prop.Orthogonalize(ws)
                afq.average_pseudo_energy(c, ws.weight, ws.eloc, 0.005, e_out, relax, Esh)
            else:
                prop.step_graph(ws, Esh, ortho=ortho, relax=relax, energy_out=e_out)
            torch.cuda.synchronize()
        out[mode] = (afq.host(ws.Phi), afq.host(ws.weight), afq.host(Esh))
    for a, b in zip(out["eager"], out["graph"]):
        assert np.array_equal(a, b)


def test_population_control_vs_oracle(gpu):
    afq, ctx, torch = gpu
    say("case: comb population control N=24 n=5 W=97 (zero weights included)")
    p = make_problem(24, 5, 12, seed=5)
    W = 97
    rng = np.random.default_rng(3)
    c = afq.Context(0)
    ham = afq.HamiltonianOperations(c, p.h1, p.Likn, p.PsiT, p.nup)
    prop = afq.Propagator(ham, W, dt=0.005)
    st = dict(Phi=random_walkers(p, W, seed=7),
              weight=rng.exponential(size=W) * np.exp(1j * rng.uniform(0, 6.28, size=W)),
              phase=np.exp(1j * rng.uniform(0, 6.28, size=W)), eloc=rng.normal(size=W) + 0j,
              ovlp=rng.normal(size=W) + 1j * rng.normal(size=W), theta=rng.normal(size=W) + 0j)
    st["weight"][[0, 5, 96]] = 0.0
    ws = afq.WalkerSet(st["Phi"])
    for k in ("weight", "phase", "eloc", "ovlp", "theta"):
        getattr(ws, k).copy_(torch.from_numpy(st[k]))
    prop.population_control(ws, 0.618)
    orc.population_control(st, 0.618)
    assert not prop.collapsed()
    assert np.array_equal(afq.host(ws.Phi), st["Phi"])
    for k in ("phase", "eloc", "ovlp", "theta"):
        assert np.array_equal(afq.host(getattr(ws, k)), st[k]), k
    assert rel(afq.host(ws.weight), st["weight"]) < 1e-15
    # a collapsed population is reported and left unchanged
    ws.weight.zero_()
    before = afq.host(ws.Phi)
    prop.population_control(ws, 0.5)
    assert prop.collapsed()
    assert np.array_equal(afq.host(ws.Phi), before)


def test_graph_with_population_control_matches_eager(gpu):
    """The captured step with QR and comb equals the eager sequence with the comb's uniform taken
    from the same Philox counter (stream + 1000, step counter)."""
    afq, ctx, torch = gpu
    p = make_problem(24, 5, 96, seed=1001)
    W = 40
    say("case: graph step + QR + comb vs eager, N=24 n=5 M=96 W=40, 12 steps")
    out = {}
    seed = 99
    for mode in ("eager", "graph"):
        c = afq.Context(0)
        ham = afq.HamiltonianOperations(c, p.h1, p.Likn, p.PsiT, p.nup)
        prop = afq.Propagator(ham, W, dt=0.005, seed=seed)
        ws = afq.WalkerSet(p.walkers(W))
        E = afq.empty((W, 3), torch.complex128)
        ov = afq.empty(W, torch.complex128)
        prop.Energy(ws, E, ov)
        ws.eloc.copy_(E.sum(1))
        ws.ovlp.copy_(ov)
        Esh = afq.device_tensor(np.array([float(afq.host(E.sum(1)).real.mean())]))
        e_out = afq.empty(1, torch.float64)
        relax = 1 - np.exp(-0.1)
        for step in range(12):
            ortho, pc = step % 5 == 0, step % 4 == 0
            if mode == "eager":
                prop.Propagate(ws, Esh)
                if ortho:
                    prop.Orthogonalize(ws)
                if pc:
                    ctr = orc.philox4x32_10([0, 0, step, 1 + 1000], [seed & 0xffffffff, seed >> 32])
                    bits = (int(ctr[0]) << 32) | int(ctr[1])
                    prop.population_control(ws, ((bits >> 11) + 0.5) * 2.0 ** -53)
                afq.average_pseudo_energy(c, ws.weight, ws.eloc, 0.005, e_out, relax, Esh)
            else:
                prop.step_graph(ws, Esh, ortho=ortho, popctl=pc, relax=relax, energy_out=e_out)
            torch.cuda.synchronize()
        out[mode] = (afq.host(ws.Phi), afq.host(ws.weight), afq.host(Esh))
    for a, b in zip(out["eager"], out["graph"]):
        assert np.array_equal(a, b)


# ---- shapes at the edges of the tilings -------------------------------------------------------

ODD = [(7, 1, 3, 1), (17, 3, 5, 5), (33, 16, 40, 3), (65, 64, 10, 2), (16, 16, 16, 4)]


@pytest.mark.parametrize("N,n,M,W", ODD, ids=[f"N{a}-n{b}-M{c}-W{d}" for a, b, c, d in ODD])
def test_edge_shapes_vs_oracle(gpu, N, n, M, W):
    """Orbitals, electrons, Cholesky vectors and walkers that fill no 16-wide tile (a single
    electron, a single walker, n = N, n = 64, N one past a tile): the energy and vbias of
    perturbed walkers and three full steps with a QR in the first, against the oracle."""
    afq, ctx0, torch = gpu
    say(f"case: edge shape N={N} n={n} M={M} W={W}")
    p = make_problem(N, n, M, seed=2000 + N + n)
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, n)
    osys = orc.OracleSystem(p, dt=0.005)
    Phi = random_walkers(p, W, seed=9)
    Gr, _ = orc.mixed_density_matrix(p.PsiT, Phi)
    dG = afq.device_tensor(Gr)
    v = afq.empty((W, M), torch.complex128)
    ham.vbias(dG, v, 0.005, precision=afq.AFQ_FP64)
    E = afq.empty((W, 3), torch.complex128)
    ham.energy(E, dG, precision=afq.AFQ_FP64)
    Er = orc.energy(osys.haj, osys.Lnak, Gr, n)
    ev = rel(afq.host(v), orc.vbias(osys.Lnak, Gr, 0.005))
    ee = max(rel(afq.host(E)[:, c], Er[:, c]) for c in range(3))
    say(f"  vbias {ev:.2e}  energy {ee:.2e}")
    assert ev < TOL and ee < TOL
    prop = afq.Propagator(ham, W, dt=0.005)
    st, Eshift = osys.init_walkers(Phi)
    ws = afq.WalkerSet(st["Phi"])
    for k in ("weight", "phase", "eloc", "ovlp", "theta"):
        getattr(ws, k).copy_(torch.from_numpy(st[k]))
    Esh = afq.device_tensor(np.array([Eshift]))
    for step in range(3):
        u = orc.field_uniforms(5, 1, step, W, M)
        prop.Propagate(ws, Esh, u=afq.device_tensor(u))
        osys.step(st, u, Eshift)
        if step == 0:
            prop.Orthogonalize(ws)
            osys.orthogonalize(st)
        e = max([rel(afq.host(ws.Phi), st["Phi"])] +
                [rel(afq.host(getattr(ws, k)), st[k]) for k in ("ovlp", "eloc", "weight")])
        say(f"  step {step}: {e:.2e}")
        assert e < TOL, step
    prop.close()
