"""The device path on a real molecule read from the reference's own inputs: BH (19 orbitals,
78 Cholesky vectors, 3+3 electrons; tests/golden/bh, copied from the reference's functional
test inputs).

- Deterministic: the local energy at the RHF trial equals the closed-form Hartree-Fock energy
  of the integrals (1e-10), and the PHMSD (CASCI, 44 determinants) overlap, density matrix and
  local energy of perturbed walkers equal the explicit sum over determinants of the oracle.
- Statistical: seeded hybrid AFQMC runs at the reference's parameters (dt = 0.01, 1536 walkers,
  3000 equilibration steps, measurements, population control and orthogonalization every 10
  steps) agree with the reference's statistical results
  (tests/functional/statistical_references/BH/..., extracted into statistical_refs.json) by the
  reference runner's criterion (run_functional.py:385-410): |E - E_ref| <= z sqrt(s^2 + s_ref^2)
  with z the two-sided 0.001 quantile (3.29), s the blocking error.  The runs are shorter than
  the reference's (9000 measured steps instead of 45000), so s is the larger error.
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd import h5inputs  # noqa: E402

G = os.path.join(os.path.dirname(__file__), "golden", "bh")
Z_CRIT = 3.2905267314919255   # scipy.stats.norm.ppf(1 - 0.001 / 2)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(h5inputs.LIB_PATH):
        pytest.skip("libafq_h5.so not built")
    from safire_amd import afq
    return afq, afq.Context(0), torch


def test_bh_rhf_trial_energy(gpu):
    afq, ctx, torch = gpu
    say("case: BH RHF trial energy vs closed-form Hartree-Fock")
    from test_h5inputs import rhf_energy
    s = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"),
                             os.path.join(G, "afqmc_rhf_nomsd.h5"))
    ref = rhf_energy(s)
    for prec in (afq.AFQ_FP64, afq.AFQ_MIXED):
        ham = afq.HamiltonianOperations(ctx, s.h1, s.Likn, s.PsiT, s.nup, E0=s.E0, precision=prec)
        prop = afq.Propagator(ham, 4, dt=0.01)
        ws = afq.WalkerSet(s.walkers(4)[0])
        E = afq.empty((4, 3), torch.complex128)
        ov = afq.empty(4, torch.complex128)
        prop.Energy(ws, E, ov, prec)
        e = afq.host(E).sum(1)
        tol = 1e-10 if prec == afq.AFQ_FP64 else 1e-6
        say(f"  precision {prec}: E = {e[0].real:.12f}  closed form {ref.real:.12f}")
        assert np.abs(e - ref).max() < tol * abs(ref)
        prop.close()


def test_bh_phmsd_vs_explicit_sum(gpu):
    afq, ctx, torch = gpu
    say("case: BH CASCI PHMSD trial ops vs explicit determinant sum")
    from safire_amd.phmsd import PHMSD
    s = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"),
                             os.path.join(G, "afqmc_casci_rhf_phmsd.h5"))
    ham = afq.HamiltonianOperations(ctx, s.h1, s.Likn, s.PsiT, s.nact, s.nact,
                                    walker_type=afq.AFQ_COLLINEAR, E0=s.E0)
    trial = PHMSD(ham, s.ph)
    W = 6
    rng = np.random.default_rng(7)
    up = s.PsiT[:s.nup].conj().T
    Pa = up[None] + 0.2 * (rng.standard_normal((W, 19, 3)) + 1j * rng.standard_normal((W, 19, 3)))
    Pb = up[None] + 0.2 * (rng.standard_normal((W, 19, 3)) + 1j * rng.standard_normal((W, 19, 3)))
    ws = afq.WalkerSet(Pa, Pb)
    Gd = afq.empty((W, 2 * s.nact, 19), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    trial.MixedDensityMatrix(ws, Gd, ov)
    E = afq.empty((W, 3), torch.complex128)
    ov2 = afq.empty(W, torch.complex128)
    trial.Energy(ws, E, ov2)
    orbs = s.PsiT[:s.nact]
    L1, h1 = orc.half_rotate(orbs, s.Likn, s.h1)
    ci = s.ci * (-1.0 if s.ci[0].real < 0 else 1.0)
    Gb, Eb, ovb = orc.msd_brute(orbs, s.occa, s.occb, ci, np.concatenate([h1, h1]),
                                np.concatenate([L1, L1]), Pa, Pb, E0=s.E0)
    eG = np.abs(afq.host(Gd) - Gb).max() / np.abs(Gb).max()
    eE = np.abs(afq.host(E) - Eb).max() / np.abs(Eb).max()
    eo = np.abs(np.exp(afq.host(ov) - ovb) - 1).max()
    say(f"  G {eG:.2e}  E {eE:.2e}  overlap {eo:.2e}")
    assert eG < 1e-10 and eE < 1e-10 and eo < 1e-10
    assert np.abs(afq.host(ov2) - afq.host(ov)).max() < 1e-12


CASES = [("rhf_nomsd/closed", "fp64"), ("rhf_nomsd/closed", "mixed"),
         ("uhf_nomsd/collinear", "fp64"), ("rcasci_rhf_phmsd/collinear", "fp64"),
         ("rcasci_uhf_phmsd/collinear", "fp64")]


def run_case(afq, ctx, case, prec, measured_steps):
    from safire_amd import driver
    from safire_amd.phmsd import PHMSD
    refs = json.load(open(os.path.join(G, "statistical_refs.json")))
    r = refs[case]
    wt = afq.AFQ_CLOSED if r["walker_type"] == "CLOSED" else afq.AFQ_COLLINEAR
    s = h5inputs.load_system(os.path.join(G, r["hamiltonian"]), os.path.join(G, r["wavefunction"]),
                             walker_type=wt)
    W = r["walkers"]
    p = afq.AFQ_MIXED if prec == "mixed" else afq.AFQ_FP64
    trial = None
    if s.ph is None:
        ham = afq.HamiltonianOperations(ctx, s.h1, s.Likn, s.PsiT, s.nup, s.ndown, walker_type=wt,
                                        E0=s.E0, precision=p)
        Pa, Pb = s.walkers(W)
    else:
        ham = afq.HamiltonianOperations(ctx, s.h1, s.Likn, s.PsiT, s.nact, s.nact,
                                        walker_type=afq.AFQ_COLLINEAR, E0=s.E0, precision=p)
        trial = PHMSD(ham, s.ph)
        up = np.ascontiguousarray(s.PsiT[:s.nup].conj().T)
        Pa = np.broadcast_to(up, (W,) + up.shape).copy()
        Pb = Pa.copy()
    prop = afq.Propagator(ham, W, dt=r["timestep"], seed=42, trial=trial)
    ws = afq.WalkerSet(Pa, Pb)
    eq = r["equilibration_steps"]
    tr = driver.run_afqmc(prop, ws, steps=eq + measured_steps, equilibration=eq,
                          ortho_interval=r["walker_ortho_interval"],
                          popctl_interval=r["population_control_interval"],
                          measure_interval=r["measure_interval"],
                          progress=lambda st, t: say(f"  step {st}"), progress_interval=3000)
    e = np.asarray(tr.energy)
    mean = complex(e.mean())
    err = complex(driver.blocking_error(e.real), driver.blocking_error(e.imag))
    prop.close()
    if trial is not None:
        trial.close()
    return r, mean, err, np.real(np.mean(np.asarray(tr.components), axis=0))


@pytest.mark.parametrize("case,prec", CASES)
def test_bh_statistical_reference(gpu, case, prec):
    afq, ctx, torch = gpu
    say(f"case: BH statistical {case} {prec}")
    r, mean, err, comps = run_case(afq, ctx, case, prec, measured_steps=9000)
    ref = r["observables"]["Energy"]
    rmean = complex(*ref["mean"])
    rerr = complex(*ref["error"])
    d = abs(mean - rmean)
    tol = Z_CRIT * np.hypot(abs(err), abs(rerr))
    say(f"  E = {mean.real:.6f} +- {err.real:.6f}   reference {rmean.real:.6f} +- {rerr.real:.6f}"
        f"   |dE| = {d:.2e}, tolerance {tol:.2e}")
    say(f"  (E1, EX, EJ) = {comps}   reference "
        f"{[r['observables'][k]['mean'][0] for k in ('OnebodyEnergy', 'ExchangeEnergy', 'CoulombEnergy')]}")
    assert np.isfinite(mean) and err.real > 0
    assert d <= tol
