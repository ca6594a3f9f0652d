"""The device path on real molecules read from the reference's own inputs (tests/golden/<molecule>,
copied from the reference's functional test inputs by tests/golden/make_molecule_fixtures.py):
BH (19 orbitals, 78 Cholesky vectors, 3+3 electrons), N2 (28 orbitals, 134 vectors, 7+7
electrons, a 50-determinant CAS trial) and Li (14 orbitals, 70 vectors, a fully polarized ROHF
trial with 3+0 electrons).

- Deterministic: the local energy at the BH RHF trial equals the closed-form Hartree-Fock energy
  of the integrals (1e-10), the PHMSD (CASCI, 44 determinants) overlap, density matrix and
  local energy of perturbed walkers equal the explicit sum over determinants of the oracle, and
  the NOMSD form of the UHF CASCI trial (34 non-orthogonal determinants) gives the overlaps and
  local energies of its PHMSD form.
- Statistical: seeded hybrid AFQMC runs at the reference's parameters (dt = 0.01, 1536 walkers,
  3000 equilibration steps, measurements, population control and orthogonalization every 10
  steps, the reference's estimator blocks) agree with the reference's statistical results
  (tests/functional/statistical_references/<molecule>/..., extracted into
  statistical_refs.json) observable by observable (Energy, its three components, Overlap, and the
  mixed and back-propagated OneRDM and DiagonalTwoRDM where the reference measured them) by
  the reference runner's criterion (run_functional.py:385-410): |A - B| <= z sqrt(s^2 + s_ref^2)
  element-wise, z the two-sided 0.001 quantile Bonferroni-adjusted over the elements (3.29 for
  a scalar), s the blocking error, with the 1e-9 (1 + |B|) floor for elements without
  stochastic error.  The runs are shorter than the reference's (9000 measured steps instead of
  9000-45000), so s is usually the larger error.
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from safire_amd import h5inputs  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden")
G = os.path.join(GOLD, "bh")


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


def test_bh_casci_nomsd_equals_phmsd_form(gpu):
    """The reference holds the UHF CASCI trial of BH twice: as 34 non-orthogonal determinants
    (NOMSD) and as particle-hole excitations (PHMSD).  The device NOMSD trial's log overlap,
    mixed density matrix, vbias and local energy on perturbed walkers equal the oracle's PHMSD
    determinant sum (overlaps up to the common normalisation, 1e-10)."""
    afq, ctx, torch = gpu
    from safire_amd.nomsd import NOMSD
    say("case: BH CASCI UHF NOMSD (34 determinants) vs its PHMSD form")
    sn = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"),
                              os.path.join(G, "afqmc_casci_uhf_nomsd.h5"))
    sp = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"),
                              os.path.join(G, "afqmc_casci_uhf_phmsd.h5"))
    ham = afq.HamiltonianOperations(ctx, sn.h1, sn.Likn, sn.PsiT, 3, 3,
                                    walker_type=afq.AFQ_COLLINEAR, E0=sn.E0)
    trial = NOMSD(ham, afq.device_tensor(sn.dets, np.complex128), sn.ci)
    W = 6
    rng = np.random.default_rng(11)
    Pa = sn.PsiT[:3].conj().T[None] + 0.2 * (rng.standard_normal((W, 19, 3)) +
                                             1j * rng.standard_normal((W, 19, 3)))
    Pb = sn.PsiT[3:].conj().T[None] + 0.2 * (rng.standard_normal((W, 19, 3)) +
                                             1j * rng.standard_normal((W, 19, 3)))
    ws = afq.WalkerSet(Pa, Pb)
    orbs = sp.PsiT[:sp.nact]
    L1, h1 = orc.half_rotate(orbs, sp.Likn, sp.h1)
    Gb, Eb, ovb = orc.msd_brute(orbs, sp.occa, sp.occb, sp.ci, np.concatenate([h1, h1]),
                                np.concatenate([L1, L1]), Pa, Pb, E0=sp.E0)
    ov = afq.empty(W, torch.complex128)
    trial.Log_Overlap(ws, ov)
    eo = np.abs(np.exp(afq.host(ov) - ovb) - 1).max()
    E = afq.empty((W, 3), torch.complex128)
    ov2 = afq.empty(W, torch.complex128)
    trial.Energy(ws, E, ov2)
    eE = np.abs(afq.host(E) - Eb).max() / np.abs(Eb).max()
    # full density matrix: the PHMSD compact G rotated back to the orbital basis
    ws.ovlp.copy_(ov)
    Gd = afq.empty((W, 2 * 19 * 19), torch.complex128)
    Ov = afq.empty(W, torch.complex128)
    trial.MixedDensityMatrix(ws, Gd, Ov)
    Gfull = np.einsum("ai,wsak->wsik", orbs, Gb.reshape(W, 2, sp.nact, 19)).reshape(W, -1)
    eG = np.abs(afq.host(Gd) - Gfull).max() / np.abs(Gfull).max()
    v = afq.empty((W, sn.Likn.shape[2]), torch.complex128)
    trial.vbias(ws, v, 0.01)
    vr = orc.vbias_full(sn.Likn, Gfull, 0.01, 1.0)
    ev = np.abs(afq.host(v) - vr).max() / np.abs(vr).max()
    say(f"  overlap {eo:.2e}  E {eE:.2e}  G {eG:.2e}  vbias {ev:.2e}")
    assert max(eo, eE, eG, ev) < 1e-10
    trial.close()


CASES = [("BH", "rhf_nomsd/closed", "fp64"), ("BH", "rhf_nomsd/closed", "mixed"),
         ("BH", "uhf_nomsd/collinear", "fp64"), ("BH", "rcasci_rhf_phmsd/collinear", "fp64"),
         ("BH", "rcasci_uhf_phmsd/collinear", "fp64"), ("BH", "rcasci_rhf_nomsd/closed", "fp64"),
         ("BH", "rcasci_uhf_nomsd/collinear", "fp64"), ("N2", "cas_wfn_collinear/collinear", "fp64"),
         ("Li", "rohf_wfn_polarized/collinear", "fp64")]
SCALARS = ("Energy", "OnebodyEnergy", "ExchangeEnergy", "CoulombEnergy", "Overlap")


def run_case(afq, ctx, molecule, case, prec, measured_steps):
    from safire_amd import driver
    from safire_amd.phmsd import PHMSD
    gdir = os.path.join(GOLD, molecule.lower())
    refs = json.load(open(os.path.join(gdir, "statistical_refs.json")))
    r = refs[case]
    wt = afq.AFQ_CLOSED if r["walker_type"] == "CLOSED" else afq.AFQ_COLLINEAR
    s = h5inputs.load_system(os.path.join(gdir, r["hamiltonian"]),
                             os.path.join(gdir, r["wavefunction"]), walker_type=wt)
    W = r["walkers"]
    p = afq.AFQ_MIXED if prec == "mixed" else afq.AFQ_FP64
    trial = None
    if s.ph is None:
        ham = afq.HamiltonianOperations(ctx, s.h1, s.Likn, s.PsiT, s.nup, s.ndown, walker_type=wt,
                                        E0=s.E0, precision=p)
        if s.dets is not None:
            from safire_amd.nomsd import NOMSD
            trial = NOMSD(ham, afq.device_tensor(s.dets, np.complex128), s.ci)
        Pa, Pb = s.walkers(W)
    else:
        ham = afq.HamiltonianOperations(ctx, s.h1, s.Likn, s.PsiT, s.nact, s.nact,
                                        walker_type=afq.AFQ_COLLINEAR, E0=s.E0, precision=p)
        trial = PHMSD(ham, s.ph)
        up = np.ascontiguousarray(s.PsiT[:s.nup].conj().T)
        Pa = np.broadcast_to(up, (W,) + up.shape).copy()
        Pb = Pa.copy()
    prop = afq.Propagator(ham, W, dt=r["timestep"], seed=42, trial=trial)
    say(f"  N={s.h1.shape[0]} M={s.Likn.shape[2]} electrons {s.nup}+{s.ndown}, "
        f"{prop.number_of_references()} reference determinant(s), {W} walkers")
    ws = afq.WalkerSet(Pa, Pb)
    eq = r["equilibration_steps"]
    tr = driver.run_afqmc(prop, ws, steps=eq + measured_steps, equilibration=eq,
                          ortho_interval=r["walker_ortho_interval"],
                          popctl_interval=r["population_control_interval"],
                          measure_interval=r["measure_interval"], estimators=r["estimators"],
                          progress=lambda st, t: say(f"  step {st}"), progress_interval=3000)
    prop.close()
    if trial is not None:
        trial.close()
    return r, tr


def _cplx(d):
    return (np.asarray(d["re"]) + 1j * np.asarray(d["im"])).reshape(d["shape"])


def compare(name, A, Aerr, B, Berr):
    """run_functional.py:385-410 on one observable (any rank, a scalar being the one-element
    case)."""
    from scipy.stats import norm
    A, Aerr = np.atleast_1d(A), np.abs(np.atleast_1d(Aerr))
    B, Berr = np.atleast_1d(B), np.abs(np.atleast_1d(Berr))
    z = norm.ppf(1 - 0.001 / (2 * A.size))
    tol = np.maximum(z * np.sqrt(Aerr ** 2 + Berr ** 2), 1e-9 * (1 + np.abs(B)))
    d = np.abs(A - B)
    worst = np.unravel_index(np.argmax(d / tol), d.shape)
    ok = bool(d[worst] <= tol[worst])
    say(f"  {'ok      ' if ok else 'MISMATCH'} {name}: |d| = {d[worst]:.3e}, tolerance {tol[worst]:.3e}"
        f" (z {z:.2f}): {A[worst]:.6f} +- {Aerr[worst]:.6f} vs {B[worst]:.6f} +- {Berr[worst]:.6f}"
        f" at {tuple(int(x) for x in worst)}")
    return ok


@pytest.mark.parametrize("molecule,case,prec", CASES)
def test_statistical_reference(gpu, molecule, case, prec):
    afq, ctx, torch = gpu
    from safire_amd import driver
    say(f"case: {molecule} statistical {case} {prec}")
    r, tr = run_case(afq, ctx, molecule, case, prec, measured_steps=9000)
    series = {"Energy": tr.energy, "OnebodyEnergy": [c[0] for c in tr.components],
              "ExchangeEnergy": [c[1] for c in tr.components],
              "CoulombEnergy": [c[2] for c in tr.components], "Overlap": tr.overlap}
    ok = True
    for name in SCALARS:
        mean, err = driver.mean_error(np.asarray(series[name], np.complex128)[:, None])
        ref = r["observables"][name]
        ok &= compare(name, mean[0], err[0], complex(*ref["mean"]), complex(*ref["error"]))
    e = np.asarray(tr.energy)
    assert np.all(np.isfinite(e)) and driver.blocking_error(e.real) > 0
    for group, obs in r["rdm"].items():
        store = tr.mixed if group == "mixed" else tr.backprop[int(group.split("/")[1])]
        for name, ref in obs.items():
            mean, err = driver.mean_error(store[name])
            ok &= compare(f"{group} {name} ({len(store[name])} measurements)", mean, err,
                          _cplx(ref["mean"]), _cplx(ref["error"]))
    assert ok


def test_time_evolved_matches_backprop_reference(gpu):
    """BH with the RHF trial (FP64): the time-evolved estimator with M transposed, over the
    reference's back-propagation window (100 steps, path restoration), agrees with the
    reference's back-propagated OneRDM and DiagonalTwoRDM under the runner's criterion.  Its
    regularization is exact for orthonormal walkers, so the walkers are orthogonalized every
    step, the Taylor series runs at order 12 (the estimator needs Y = X^{-H}), and population
    control copies the operators with their walkers."""
    afq, ctx, torch = gpu
    from safire_amd import driver
    gdir = os.path.join(GOLD, "bh")
    r = json.load(open(os.path.join(gdir, "statistical_refs.json")))["rhf_nomsd/closed"]
    say("case: BH time-evolved back-propagation rhf_nomsd/closed fp64")
    s = h5inputs.load_system(os.path.join(gdir, r["hamiltonian"]),
                             os.path.join(gdir, r["wavefunction"]), walker_type=afq.AFQ_CLOSED)
    W = r["walkers"]
    ham = afq.HamiltonianOperations(ctx, s.h1, s.Likn, s.PsiT, s.nup, s.ndown,
                                    walker_type=afq.AFQ_CLOSED, E0=s.E0)
    prop = afq.Propagator(ham, W, dt=r["timestep"], seed=42, order=12)
    ws = afq.WalkerSet(*s.walkers(W))
    bp = r["estimators"]["backprop"]
    te = {"propagation_steps": bp["propagation_steps"], "path_restoration": bp["path_restoration"],
          "transpose_m": True, "follow_walkers": True, "onerdm": {}, "diag_twordm": {}}
    eq = r["equilibration_steps"]
    tr = driver.run_afqmc(prop, ws, steps=eq + 9000, equilibration=eq, ortho_interval=1,
                          popctl_interval=r["population_control_interval"],
                          measure_interval=r["measure_interval"],
                          estimators={"time_evolved_bp": te},
                          progress=lambda st, t: say(f"  step {st}"), progress_interval=3000)
    prop.close()
    L = bp["propagation_steps"][0]
    ok = True
    for name, ref in r["rdm"][f"backprop/{L}"].items():
        series = tr.time_evolved_bp[L][name]
        mean, err = driver.mean_error(series)
        ok &= compare(f"time-evolved/{L} {name} ({len(series)} measurements)", mean, err,
                      _cplx(ref["mean"]), _cplx(ref["error"]))
    assert ok
