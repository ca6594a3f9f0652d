"""The reference's HDF5 input files as device-path inputs (C ABI: include/afq_h5.h,
libafq_h5.so).

read_hamiltonian / read_wavefunction return plain arrays in the layouts the device
HamiltonianOperations and PHMSD trial take; load_system puts a Hamiltonian file and a trial
file together the way the reference's factories do (RealDenseHamiltonian.cpp:52-186,
Wavefunction.cpp:256-300 for NOMSD and 330-496 for PHMSD).  Walkers start as the trial
determinant, Phi = PsiT^H (execute.cpp initial walkers from the trial).
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libafq_h5.so")
NOMSD, PHMSD = 0, 1
CLOSED, COLLINEAR, NONCOLLINEAR = 0, 1, 2
_lib = None


class H5Error(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise H5Error(f"{LIB_PATH} is not built (make -C safire_amd, needs HDF5 headers)")
        L = ctypes.CDLL(LIB_PATH)
        I, P, D = ctypes.c_int, ctypes.c_void_p, ctypes.c_double
        IP = ctypes.POINTER(I)
        S = ctypes.c_char_p
        sigs = {
            "afq_h5_last_error": ([], ctypes.c_char_p),
            "afq_h5_hamiltonian_shape": ([S, IP, IP, IP, IP], I),
            "afq_h5_read_hamiltonian": ([S, I, P, P, ctypes.POINTER(D)], I),
            "afq_h5_wavefunction_info": ([S, IP, IP, IP, IP, IP, IP, IP], I),
            "afq_h5_read_psit": ([S, I, I, I, I, P], I),
            "afq_h5_read_ci": ([S, I, I, P], I),
            "afq_h5_read_occupations": ([S, I, I, I, P, P], I),
            "afq_h5_results_append": ([S, S, S, I, P, I, ctypes.c_long, P], I),
            "afq_h5_results_shape": ([S, S, S, ctypes.POINTER(ctypes.c_long), IP, P, IP], I),
            "afq_h5_results_read": ([S, S, S, P], I),
            "afq_h5_results_names": ([S, S, P, ctypes.c_long, ctypes.POINTER(ctypes.c_long)], I),
        }
        for name, (args, res) in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _ck(rc):
    if rc != 0:
        raise H5Error(lib().afq_h5_last_error().decode())


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclasses.dataclass
class Hamiltonian:
    h1: np.ndarray        # [nspin][N][N] real
    Likn: np.ndarray      # [N][N][M] real, or [2][N][N][M] with spin-dependent vectors
    E0: float

    @property
    def N(self):
        return self.Likn.shape[-2]

    @property
    def M(self):
        return self.Likn.shape[-1]


def read_hamiltonian(path: str, nelec: int) -> Hamiltonian:
    """/Hamiltonian of a dense factorized file; nelec (all spins) enters the Madelung term."""
    ns, N, M, nsl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    b = path.encode()
    _ck(lib().afq_h5_hamiltonian_shape(b, ctypes.byref(ns), ctypes.byref(N), ctypes.byref(M),
                                       ctypes.byref(nsl)))
    h1 = np.empty((ns.value, N.value, N.value))
    L = np.empty(((nsl.value,) if nsl.value > 1 else ()) + (N.value, N.value, M.value))
    E0 = ctypes.c_double()
    _ck(lib().afq_h5_read_hamiltonian(b, nelec, _p(h1), _p(L), ctypes.byref(E0)))
    return Hamiltonian(h1=h1, Likn=L, E0=E0.value)


@dataclasses.dataclass
class Wavefunction:
    kind: int              # NOMSD or PHMSD
    walker_type: int       # CLOSED / COLLINEAR / NONCOLLINEAR of spin_type
    N: int
    nup: int
    ndown: int
    ci: np.ndarray         # [ndets] complex
    psit: list             # PsiT_k dense complex matrices, in file order
    occa: np.ndarray | None = None   # PHMSD [ndets][nup]
    occb: np.ndarray | None = None   # PHMSD [ndets][ndown]


def read_wavefunction(path: str) -> Wavefunction:
    b = path.encode()
    v = [ctypes.c_int() for _ in range(7)]
    _ck(lib().afq_h5_wavefunction_info(b, *[ctypes.byref(x) for x in v]))
    kind, wt, N, nup, ndown, ndets, npsit = (x.value for x in v)
    npol = 2 if wt == NONCOLLINEAR else 1
    ci = np.empty(ndets, np.complex128)
    _ck(lib().afq_h5_read_ci(b, kind, ndets, _p(ci)))
    psit = []
    for k in range(npsit):
        if kind == NOMSD:
            rows = ndown if (wt == COLLINEAR and k % 2 == 1) else nup
            cols = npol * N
        else:
            rows = cols = npol * N   # PHMSD references rotate all orbitals (Wavefunction.cpp:124-129)
        m = np.empty((rows, cols), np.complex128)
        _ck(lib().afq_h5_read_psit(b, kind, k, rows, cols, _p(m)))
        psit.append(m)
    occa = occb = None
    if kind == PHMSD:
        occa = np.empty((ndets, nup), np.int32)
        occb = np.empty((ndets, ndown), np.int32)
        _ck(lib().afq_h5_read_occupations(b, ndets, nup, ndown, _p(occa), _p(occb)))
    return Wavefunction(kind=kind, walker_type=wt, N=N, nup=nup, ndown=ndown, ci=ci, psit=psit,
                        occa=occa, occb=occb)


@dataclasses.dataclass
class System:
    """Inputs of HamiltonianOperations (and of a PHMSD trial) from one pair of files."""
    h1: np.ndarray          # [N][N] real, or [2][N][N] (alpha, beta) when spin-dependent
    Likn: np.ndarray        # [N][N][M] real, or [2][N][N][M] when spin-dependent
    E0: float
    PsiT: np.ndarray        # [nup (+ ndown)][N] trial rows (COLLINEAR: up rows, then down)
    nup: int
    ndown: int
    walker_type: int        # afq.AFQ_CLOSED / afq.AFQ_COLLINEAR
    ph: object = None       # PHMSD: safire_amd.phmsd.PhStruct over the active orbitals
    nact: int = 0
    occa: np.ndarray | None = None   # PHMSD: sorted active occupations [ndet][nup] ...
    occb: np.ndarray | None = None   # ... [ndet][ndown]
    ci: np.ndarray | None = None     # PHMSD: coefficients of the sorted strings; NOMSD: ci
    dets: np.ndarray | None = None   # NOMSD with several determinants: [ndet][nup (+ ndown)][N]

    def walkers(self, W: int):
        """W copies of the trial determinant (Phi = PsiT^H per spin block); for several NOMSD
        determinants the one of largest |ci| (NOMSD::initial_guess, NOMSD.cpp:131-153)."""
        rows = self.PsiT if self.dets is None else self.dets[int(np.argmax(np.abs(self.ci)))]
        up = np.ascontiguousarray(rows[:self.nup].conj().T)
        Pa = np.broadcast_to(up, (W,) + up.shape).copy()
        if self.walker_type == CLOSED:
            return Pa, None
        dn = np.ascontiguousarray(rows[self.nup:self.nup + self.ndown].conj().T)
        return Pa, np.broadcast_to(dn, (W,) + dn.shape).copy()


def _one_body(h: Hamiltonian) -> np.ndarray:
    """hcore as [N][N] when its spin blocks are equal, else [2][N][N] (alpha, beta)."""
    if all(np.array_equal(h.h1[s], h.h1[0]) for s in range(1, h.h1.shape[0])):
        return np.ascontiguousarray(h.h1[0])
    return np.ascontiguousarray(h.h1)


def load_system(ham_path: str, wfn_path: str, walker_type: int | None = None) -> System:
    """A NOMSD (one or several determinants) or PHMSD trial with its Hamiltonian.  walker_type
    defaults to the trial's spin_type; a CLOSED trial may be run as COLLINEAR walkers
    (readWfn.h:129-140 duplicates the orbitals)."""
    wf = read_wavefunction(wfn_path)
    wt = wf.walker_type if walker_type is None else walker_type
    if wt == NONCOLLINEAR or wf.walker_type == NONCOLLINEAR:
        raise H5Error("NONCOLLINEAR walkers are not supported by the device path")
    nelec = wf.nup + (wf.ndown if wt == COLLINEAR else wf.nup)
    ham = read_hamiltonian(ham_path, nelec)
    if ham.N != wf.N:
        raise H5Error(f"Hamiltonian has {ham.N} orbitals, the trial {wf.N}")
    h1 = _one_body(ham)
    if wt == CLOSED and (h1.ndim == 3 or ham.Likn.ndim == 4):
        raise H5Error("spin-dependent integrals need COLLINEAR walkers")
    if wf.kind == NOMSD:
        # PsiT_k: determinant k (CLOSED), or k = 2 d + spin (COLLINEAR) (readWfn.h:90-140)
        def rows(d):
            if wf.walker_type == CLOSED and wt == CLOSED:
                return wf.psit[d]
            if wf.walker_type == CLOSED and wt == COLLINEAR:
                return np.concatenate([wf.psit[d], wf.psit[d]])
            if wf.walker_type == COLLINEAR and wt == COLLINEAR:
                return np.concatenate([wf.psit[2 * d], wf.psit[2 * d + 1]])
            raise H5Error("a COLLINEAR trial cannot drive CLOSED walkers")
        ndet = len(wf.ci)
        dets = np.ascontiguousarray(np.stack([rows(d) for d in range(ndet)])) if ndet > 1 else None
        return System(h1=h1, Likn=ham.Likn, E0=ham.E0, PsiT=np.ascontiguousarray(rows(0)),
                      nup=wf.nup, ndown=wf.ndown if wt == COLLINEAR else 0, walker_type=wt,
                      ci=wf.ci.copy(), dets=dets)
    return _phmsd_system(ham, h1, wf)


def _phmsd_system(ham: Hamiltonian, h1: np.ndarray, wf: Wavefunction) -> System:
    """PHMSD (COLLINEAR): occupations over a single orbital reference (RHF-like: PsiT_0 rotates
    all orbitals, or the identity without one), the active space of Wavefunction.cpp:366-395
    (orbitals occupied in any determinant of either spin, one shared numbering), and the trial
    rows of both spins = the active orbitals (Wavefunction.cpp:459-474)."""
    from .phmsd import build_ph_struct
    if wf.walker_type != COLLINEAR:
        raise H5Error("PHMSD trials need COLLINEAR walkers")
    if len(wf.psit) > 1:
        raise H5Error("PHMSD trials with one orbital reference per spin are not supported")
    N = wf.N
    ref = wf.psit[0] if wf.psit else np.eye(N, dtype=np.complex128)
    used = np.zeros(N, bool)
    used[wf.occa.ravel()] = True
    used[wf.occb.ravel()] = True
    act = np.nonzero(used)[0]
    if not (np.array_equal(np.sort(wf.occa[0]), act[:wf.nup]) and
            np.array_equal(np.sort(wf.occb[0]), act[:wf.ndown])):
        raise H5Error("the first determinant must occupy the lowest active orbitals of each spin")
    mo2act = -np.ones(N, np.int64)
    mo2act[act] = np.arange(len(act))
    occa = np.sort(mo2act[wf.occa], axis=1)
    occb = np.sort(mo2act[wf.occb], axis=1)
    ci = wf.ci.copy()
    # sorting each occupation string into ascending order takes the permutation's sign
    # (sort_with_sign, Wavefunction.cpp:108-120)
    for d in range(len(ci)):
        for o in (mo2act[wf.occa[d]], mo2act[wf.occb[d]]):
            inv = sum(1 for i in range(len(o)) for j in range(i + 1, len(o)) if o[j] < o[i])
            if inv % 2:
                ci[d] = -ci[d]
    ph = build_ph_struct(occa, occb, ci)
    orbs = np.ascontiguousarray(ref[act])
    return System(h1=h1, Likn=ham.Likn, E0=ham.E0, PsiT=np.concatenate([orbs, orbs]),
                  nup=wf.nup, ndown=wf.ndown, walker_type=COLLINEAR, ph=ph, nact=len(act),
                  occa=occa, occb=occb, ci=ci)
