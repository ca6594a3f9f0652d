"""Particle-hole multi-determinant (PHMSD) trial: host-side construction of the excitation
structure, and the device operations behind the reference's PHMSD interface.

build_ph_struct follows the reference's reader (readWfn.cpp:161-162 and 238-340,
Excitations.hpp:148-192): determinant 0 is the reference configuration; every other
determinant's alpha and beta occupation strings are expressed as excitations of it
(hole *positions* in the reference, particle *orbitals*), the unique excitations of each spin
are numbered by rank and then by first appearance, and each coefficient takes the sign of the
permutation that sorts the reference with its holes replaced in place by the particles.  The
device kernels consume the same numbering: configuration 0 is the reference, configurations
of rank r are contiguous.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class PhStruct:
    nel: tuple                # (nup, ndown)
    refc: tuple               # per spin [nel] int32 reference occupation (active indices)
    rank: tuple               # per spin [nconf] int32 excitation rank, 0 for the reference
    exc_start: tuple          # per spin [nconf] int32 offset of configuration u in exc
    exc: tuple                # per spin flat int32: holes then particles of each configuration
    det_a: np.ndarray         # [ndet] int32 alpha configuration of each determinant
    det_b: np.ndarray         # [ndet] int32 beta configuration
    ci: np.ndarray            # [ndet] complex128 coefficient with the permutation sign

    @property
    def nconf(self):
        return (len(self.rank[0]), len(self.rank[1]))

    @property
    def ndet(self):
        return len(self.ci)

    def max_rank(self, s):
        return int(self.rank[s].max()) if len(self.rank[s]) else 0

    def configuration(self, s, u):
        """Occupied active orbitals of configuration u, in reference order with the holes
        replaced in place (ph_excitations::get_configuration, Excitations.hpp:440-481)."""
        orbs = np.array(self.refc[s], dtype=np.int64)
        r = int(self.rank[s][u])
        e = self.exc[s][self.exc_start[s][u]:self.exc_start[s][u] + 2 * r]
        orbs[e[:r]] = e[r:]
        return orbs


def excitation(ref, confg):
    """get_excitation_number (Excitations.hpp:148-192): hole positions (indices i of ref with
    ref[i] not in confg, ascending), particle orbitals (entries of confg not in ref, in confg
    order), and the sign (-1)^(inversions of ref with each hole replaced by its particle)."""
    ref = [int(x) for x in ref]
    cs = set(int(x) for x in confg)
    rs = set(ref)
    holes = [i for i, o in enumerate(ref) if o not in cs]
    parts = [int(o) for o in confg if int(o) not in rs]
    assert len(holes) == len(parts)
    work = list(ref)
    for h, p in zip(holes, parts):
        work[h] = p
    inv = sum(1 for i in range(len(work)) for j in range(i + 1, len(work)) if work[j] < work[i])
    return holes, parts, (-1.0 if inv % 2 else 1.0)


def build_ph_struct(occa, occb, ci) -> PhStruct:
    """occa [ndet][nup], occb [ndet][ndown]: sorted active-orbital occupations; ci [ndet].
    Determinant 0 is the reference."""
    occa = np.asarray(occa, dtype=np.int64)
    occb = np.asarray(occb, dtype=np.int64)
    ci = np.asarray(ci, dtype=np.complex128).copy()
    ndet = len(ci)
    if ci[0].real < 0:  # readWfn.cpp:161-162
        ci = -ci
    occ = (occa, occb)
    refs = (occa[0], occb[0])
    uniq = ({}, {})          # per spin: rank -> {excitation tuple: position}
    order = ([], [])         # per spin: rank -> list of excitation tuples (first appearance)
    for s in range(2):
        order[s].extend([] for _ in range(occ[s].shape[1] + 1))
    exc_of = np.empty((2, ndet), dtype=object)
    for d in range(ndet):
        for s in range(2):
            h, p, sign = excitation(refs[s], occ[s][d])
            ci[d] *= sign
            key = tuple(h) + tuple(p)
            r = len(h)
            exc_of[s, d] = (r, key)
            if d > 0 and r > 0 and key not in uniq[s].setdefault(r, {}):
                uniq[s][r][key] = len(order[s][r])
                order[s][r].append(key)
    rank, exc_start, exc, offset = [], [], [], []
    for s in range(2):
        rk, st, ex, off = [0], [0], [], {0: 0}
        nxt = 1
        for r in range(1, len(order[s])):
            off[r] = nxt
            for key in order[s][r]:
                rk.append(r)
                st.append(len(ex))
                ex.extend(key)
                nxt += 1
        rank.append(np.array(rk, np.int32))
        exc_start.append(np.array(st, np.int32))
        exc.append(np.array(ex if ex else [0], np.int32))
        offset.append(off)
    det_idx = np.zeros((2, ndet), np.int32)
    for d in range(ndet):
        for s in range(2):
            r, key = exc_of[s, d]
            det_idx[s, d] = 0 if r == 0 else offset[s][r] + uniq[s][r][key]
    refc = tuple(np.ascontiguousarray(np.asarray(refs[s], np.int32)) for s in range(2))
    return PhStruct(nel=(occa.shape[1], occb.shape[1]), refc=refc, rank=tuple(rank),
                    exc_start=tuple(exc_start), exc=tuple(exc), det_a=det_idx[0].copy(),
                    det_b=det_idx[1].copy(), ci=ci)


# ---------------------------------------------------------------------------------------------
# device trial (C ABI: afq_phmsd_*)
# ---------------------------------------------------------------------------------------------

import ctypes  # noqa: E402

from . import afq  # noqa: E402


class PhExcitations(ctypes.Structure):
    _fields_ = [("nel", ctypes.c_int * 2), ("nconf", ctypes.c_int * 2),
                ("rank", ctypes.c_void_p * 2), ("exc_start", ctypes.c_void_p * 2),
                ("exc", ctypes.c_void_p * 2), ("refc", ctypes.c_void_p * 2),
                ("ndet", ctypes.c_int), ("det_a", ctypes.c_void_p), ("det_b", ctypes.c_void_p),
                ("ci", ctypes.c_void_p)]


class PHMSD:
    """PHMSD<DEVICE_MEMORY> for COLLINEAR walkers (PHMSD.hpp/.cpp): `ham` must be a COLLINEAR
    HamiltonianOperations whose trial rows are the active orbitals of both spins
    (PhmsdProblem.PsiT, nup = ndown = nact), as Wavefunction.cpp:459-474 builds it."""

    def __init__(self, ham, ph: PhStruct):
        if ham.walker_type != afq.AFQ_COLLINEAR:
            raise afq.AfqError("PHMSD requires COLLINEAR walkers")
        self.ham = ham
        self.ctx = ham.ctx
        self.ph = ph
        self.nup, self.ndn = ph.nel
        keep = []

        def arr(a, t=np.int32):
            a = np.ascontiguousarray(a, dtype=t)
            keep.append(a)
            return a.ctypes.data_as(ctypes.c_void_p)

        ex = PhExcitations()
        for s in range(2):
            ex.nel[s] = ph.nel[s]
            ex.nconf[s] = len(ph.rank[s])
            ex.rank[s] = arr(ph.rank[s])
            ex.exc_start[s] = arr(ph.exc_start[s])
            ex.exc[s] = arr(ph.exc[s])
            ex.refc[s] = arr(ph.refc[s])
        ex.ndet = ph.ndet
        ex.det_a = arr(ph.det_a)
        ex.det_b = arr(ph.det_b)
        ex.ci = arr(ph.ci, np.complex128)
        h = ctypes.c_void_p()
        afq._ck(afq.lib().afq_phmsd_create(ham.h, ctypes.byref(ex), ctypes.byref(h)))
        self.h = h

    @property
    def nact(self):
        return (self.ham.nup, self.ham.ndown)

    def MixedDensityMatrix(self, wset, G, Ov):
        c = wset.cstruct()
        afq._ck(afq.lib().afq_phmsd_mixed_density_matrix(self.h, c.Phi, c.Phi_b, c.phi_stride,
                                                          wset.W, afq.ptr(G), afq.ptr(Ov)))

    def Log_Overlap(self, wset, Ov):
        c = wset.cstruct()
        afq._ck(afq.lib().afq_phmsd_log_overlap(self.h, c.Phi, c.Phi_b, c.phi_stride, wset.W,
                                                 afq.ptr(Ov)))

    def Energy(self, wset, E, Ov):
        c = wset.cstruct()
        afq._ck(afq.lib().afq_phmsd_energy(self.h, c.Phi, c.Phi_b, c.phi_stride, wset.W,
                                            afq.ptr(E), afq.ptr(Ov)))

    def vbias(self, wset, v, dt):
        """PHMSD::vbias (PHMSD.cpp:53-72): compact multi-determinant G, then the
        hamiltonian's vbias."""
        torch = afq._torch()
        N = self.ham.N
        G = torch.empty((wset.W, sum(self.nact), N), dtype=torch.complex128, device="cuda")
        Ov = torch.empty(wset.W, dtype=torch.complex128, device="cuda")
        self.MixedDensityMatrix(wset, G, Ov)
        self.ham.vbias(G, v, dt)

    def vMF(self, v, dt):
        afq._ck(afq.lib().afq_phmsd_vmf(self.h, dt, afq.ptr(v)))

    def close(self):
        if getattr(self, "h", None):
            afq.lib().afq_phmsd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
