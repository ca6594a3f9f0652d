"""The HDF5 input readers (include/afq_h5.h, safire_amd/h5inputs.py) on the reference's BH
fixtures (tests/golden/bh, copied by tests/golden/make_molecule_fixtures.py): extents, the energy
offset, orthonormal trials, symmetric integrals, the PHMSD active space and excitation
structure, and the oracle's local energy at the RHF trial against the closed-form Hartree-Fock
energy of the same integrals."""
import os

import numpy as np
import pytest

import oracle as orc
from safire_amd import h5inputs

G = os.path.join(os.path.dirname(__file__), "golden", "bh")
pytestmark = pytest.mark.skipif(not os.path.exists(h5inputs.LIB_PATH),
                                reason="libafq_h5.so not built (no HDF5 headers)")
NUCLEAR = 1.42897280978613   # /Hamiltonian nuclear_energy attribute of the BH files


def rhf_energy(s):
    """E0 + 2 tr(h D) + 2 sum_n tr(L_n D)^2 - sum_n tr(L_n D L_n D), D = C C^H per spin."""
    C = s.PsiT[:s.nup].conj().T
    D = C @ C.conj().T
    L = np.moveaxis(s.Likn, 2, 0)
    j = np.einsum("nij,ji->n", L, D)
    x = np.einsum("nij,jk,nkl,li->", L, D, L, D)
    return s.E0 + 2 * np.trace(s.h1 @ D) + 2 * np.sum(j * j) - x


def test_hamiltonian_and_nomsd():
    s = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"),
                             os.path.join(G, "afqmc_rhf_nomsd.h5"))
    assert s.walker_type == h5inputs.CLOSED and (s.nup, s.ndown) == (3, 0)
    assert s.h1.shape == (19, 19) and s.Likn.shape == (19, 19, 78) and s.PsiT.shape == (3, 19)
    assert abs(s.E0 - NUCLEAR) < 1e-12
    assert np.abs(s.h1 - s.h1.T).max() < 1e-12
    assert np.abs(s.Likn - s.Likn.transpose(1, 0, 2)).max() < 1e-12
    assert np.abs(s.PsiT @ s.PsiT.conj().T - np.eye(3)).max() < 1e-12
    u = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_collinear.h5"),
                             os.path.join(G, "afqmc_uhf_nomsd.h5"))
    assert u.walker_type == h5inputs.COLLINEAR and (u.nup, u.ndown) == (3, 3)
    assert u.PsiT.shape == (6, 19)
    # the collinear Hamiltonian file holds the same spin-independent integrals
    assert np.array_equal(u.Likn, s.Likn) and np.array_equal(u.h1, s.h1)


def test_closed_trial_run_as_collinear():
    s = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"),
                             os.path.join(G, "afqmc_rhf_nomsd.h5"),
                             walker_type=h5inputs.COLLINEAR)
    assert s.PsiT.shape == (6, 19) and np.array_equal(s.PsiT[:3], s.PsiT[3:])


def test_rhf_trial_energy_closed_form():
    s = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"),
                             os.path.join(G, "afqmc_rhf_nomsd.h5"))
    Phi = s.walkers(2)[0]
    L1, h1 = orc.half_rotate(s.PsiT, s.Likn, s.h1)
    Gm, _ = orc.mixed_density_matrix(s.PsiT, Phi)
    E = orc.energy(h1, L1, Gm, s.nup, 0, closed=True, E0=s.E0)
    e = np.asarray(E).sum(axis=1)
    ref = rhf_energy(s)
    assert abs(ref.imag) < 1e-12
    assert np.abs(e - ref).max() < 1e-10 * abs(ref)
    # a Hartree-Fock energy of BH in this basis, above the AFQMC references near -25.1655
    assert -25.2 < ref.real < -24.9


@pytest.mark.parametrize("wfn,ndet", [("afqmc_casci_rhf_phmsd.h5", 44),
                                      ("afqmc_casci_uhf_phmsd.h5", 34)])
def test_phmsd_structure(wfn, ndet):
    s = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"), os.path.join(G, wfn))
    wf = h5inputs.read_wavefunction(os.path.join(G, wfn))
    ph = s.ph
    assert s.walker_type == h5inputs.COLLINEAR and ph.ndet == ndet
    assert ph.nel == (3, 3) and s.PsiT.shape == (2 * s.nact, 19)
    # every determinant rebuilt from the excitation structure occupies the file's orbitals
    used = np.unique(np.concatenate([wf.occa.ravel(), wf.occb.ravel()]))
    assert len(used) == s.nact
    for d in range(ndet):
        for sp, occ in ((0, wf.occa), (1, wf.occb)):
            u = ph.det_a[d] if sp == 0 else ph.det_b[d]
            conf = sorted(used[ph.configuration(sp, u)])
            assert conf == sorted(occ[d])
    assert np.allclose(np.abs(ph.ci), np.abs(wf.ci))
    # active orbitals are orthonormal rows of the reference rotation
    orbs = s.PsiT[:s.nact]
    assert np.abs(orbs @ orbs.conj().T - np.eye(s.nact)).max() < 1e-12


def test_missing_file_is_an_error():
    with pytest.raises(h5inputs.H5Error):
        h5inputs.read_wavefunction(os.path.join(G, "does_not_exist.h5"))


def test_casci_nomsd_loads_every_determinant():
    for wfn, wt, rows in (("afqmc_casci_rhf_nomsd.h5", h5inputs.CLOSED, 3),
                          ("afqmc_casci_uhf_nomsd.h5", h5inputs.COLLINEAR, 6)):
        s = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"), os.path.join(G, wfn))
        assert s.walker_type == wt and s.dets.shape == (34, rows, 19) and len(s.ci) == 34
        assert np.array_equal(s.dets[0], s.PsiT)
    # a CLOSED expansion run as COLLINEAR walkers repeats each determinant's orbitals per spin
    c = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"),
                             os.path.join(G, "afqmc_casci_rhf_nomsd.h5"),
                             walker_type=h5inputs.COLLINEAR)
    assert c.dets.shape == (34, 6, 19) and np.array_equal(c.dets[:, :3], c.dets[:, 3:])


def test_casci_uhf_nomsd_is_the_casci_uhf_phmsd_state():
    """The reference's UHF CASCI trial is held twice, as 34 non-orthogonal determinants
    (PsiT_{2d + spin}) and as 34 particle-hole excitations: the NOMSD restatement (determinant
    sum of log overlaps, per-determinant half-rotated energies) and the pinned PHMSD one agree on
    generic walkers, which fixes the determinant order, the spin interleaving and the weights."""
    sn = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"),
                              os.path.join(G, "afqmc_casci_uhf_nomsd.h5"))
    sp = h5inputs.load_system(os.path.join(G, "afqmc_H_rhf_closed.h5"),
                              os.path.join(G, "afqmc_casci_uhf_phmsd.h5"))
    dets = [(d[:3], d[3:]) for d in sn.dets]
    rng = np.random.default_rng(1)
    W = 4

    def walkers(rows):
        return np.stack([np.linalg.qr(rows.conj().T + 0.3 * (rng.normal(size=(19, 3)) +
                                                              1j * rng.normal(size=(19, 3))))[0]
                         for _ in range(W)])
    Pa, Pb = walkers(sn.PsiT[:3]), walkers(sn.PsiT[3:])
    ovn = orc.msd_log_overlap(dets, sn.ci, Pa, Pb)
    orbs = sp.PsiT[:sp.nact]
    L1, h1 = orc.half_rotate(orbs, sp.Likn, sp.h1)
    _, Eb, ovb = orc.msd_brute(orbs, sp.occa, sp.occb, sp.ci, np.concatenate([h1, h1]),
                               np.concatenate([L1, L1]), Pa, Pb, E0=sp.E0)
    assert np.abs(np.exp(ovn - ovb) - 1).max() < 1e-12
    En, _ = orc.msd_energy(dets, sn.ci, orc.msd_half_rotated(dets, sn.Likn, sn.h1), Pa, Pb,
                           E0=sn.E0)
    assert np.abs(En - Eb).max() < 1e-10 * np.abs(Eb).max()
