"""The device driver loop (run_afqmc with QR, comb population control, Eshift relaxation and the
energy estimator) against the oracle driven with the same Philox draws, and the mixed-precision
energy trace against the FP64 one."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd.synthetic import make_problem  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def comb_uniform(seed, stream, step):
    c = orc.philox4x32_10([0, 0, step, stream + 1000], [seed & 0xffffffff, seed >> 32])
    return (((int(c[0]) << 32) | int(c[1])) >> 11) * 2.0 ** -53 + 2.0 ** -54


def oracle_driver(p, W, steps, seed, stream=1, equil=10, ortho=5, popctl=10, measure=5, dt=0.005):
    osys = orc.OracleSystem(p, dt=dt)
    st, Eshift = osys.init_walkers(p.walkers(W))
    relax = 1.0 - np.exp(-10.0 / equil)
    samples = []
    for step in range(steps):
        u = orc.field_uniforms(seed, stream, step, W, p.M)
        osys.step(st, u, Eshift)
        if step % ortho == 0:
            osys.orthogonalize(st)
        if step % popctl == 0:
            orc.population_control(st, comb_uniform(seed, stream, step))
        if step < equil:
            e = orc.average_pseudo_energy(st["weight"], st["eloc"], dt)
            Eshift += relax * (e - Eshift)
        elif step % measure == 0:
            G, _ = orc.mixed_density_matrix(osys.PsiT, st["Phi"])
            E = orc.energy(osys.haj, osys.Lnak, G, p.nup).sum(1)
            samples.append(np.sum(st["weight"] * E) / np.sum(st["weight"]))
    return np.array(samples), Eshift


def device_driver(p, W, steps, seed, precision=0, **kw):
    from safire_amd import afq, driver
    ctx = afq.Context(0)
    ham = afq.HamiltonianOperations(ctx, p.h1, p.Likn, p.PsiT, p.nup, precision=precision)
    prop = afq.Propagator(ham, W, dt=0.005, seed=seed)
    ws = afq.WalkerSet(p.walkers(W))
    tr = driver.run_afqmc(prop, ws, steps, **kw)
    assert not prop.collapsed()
    return tr


def test_driver_matches_oracle_trajectory():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    # weaker interactions (lscale 0.1) keep a phaseless population alive for long runs
    p = make_problem(24, 5, 96, seed=1001, lscale=0.1)
    W, steps, seed = 32, 31, 4242
    say(f"case: driver vs oracle driver N=24 n=5 M=96 W={W} steps={steps}")
    ref, esh = oracle_driver(p, W, steps, seed)
    tr = device_driver(p, W, steps, seed, equilibration=10, ortho_interval=5,
                       popctl_interval=10, measure_interval=5)
    got = np.array(tr.energy)
    assert got.shape == ref.shape and got.size >= 4
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    say(f"energy samples {np.real(ref)}; max rel diff {err:.2e}")
    assert err < 1e-8
    assert abs(tr.eshift - esh) < 1e-8 * abs(esh)


def test_mixed_trace_within_error_bar():
    """Fixed-seed runs of the same workload in FP64 and in mixed precision: the mixed energy
    estimate lies within the FP64 estimate's error bar."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from safire_amd import afq
    p = make_problem(24, 5, 96, seed=1001, lscale=0.1)
    W, steps, seed = 256, 400, 777
    say(f"case: mixed vs FP64 energy trace N=24 n=5 M=96 W={W} steps={steps}")
    res = {}
    for prec in (afq.AFQ_FP64, afq.AFQ_MIXED):
        tr = device_driver(p, W, steps, seed, precision=prec, equilibration=50, ortho_interval=5,
                           popctl_interval=10, measure_interval=2)
        res[prec] = tr.mean_error()
    (m64, e64), (m32, e32) = res[afq.AFQ_FP64], res[afq.AFQ_MIXED]
    say(f"FP64 {m64:.6f} +- {e64:.6f}   mixed {m32:.6f} +- {e32:.6f}   diff {m32 - m64:.3e} Ha")
    assert np.isfinite(m64) and np.isfinite(m32) and e64 > 0
    assert abs(m32 - m64) < e64
