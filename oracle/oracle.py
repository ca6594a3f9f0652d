"""ctypes binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker or the timed CPU baseline, never as a product path.
Arrays are numpy complex128 / float64, C-contiguous, in the layouts documented in
afq_oracle.cpp.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_inverse_logdet": [_P, _I, _I, _P, _P],
            "orc_log_overlap": [_P, _P, _I, _I, _I, _P],
            "orc_mixed_density_matrix": [_P, _P, _I, _I, _I, _P, _P, _I],
            "orc_propagate": [_P, _P, _P, _I, _I, _I, _I, ctypes.c_char],
            "orc_orthogonalize": [_P, _I, _I, _I, _P],
            "orc_vbias": [_P, _P, _I, _I, _I, _I, _D, _D, _P],
            "orc_vhs": [_P, _P, _I, _I, _I, _D, _P],
            "orc_energy": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _D, _P],
            "orc_probit": [_D],
            "orc_philox4x32_10": [_P, _P, _P],
            "orc_field_uniforms": [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _I, _I, _P],
            "orc_construct_x": [_P, _P, _D, _I, _I, _I, _I, _P, _P, _P],
            "orc_walker_update": [_I, _D, _D, _D, _P, _P, _P, _D, _D, _I, _I, _D, _D, _D,
                                  _P, _P, _P, _P, _P, _P],
            "orc_average_pseudo_energy": [_I, _P, _P, _D],
            "orc_half_rotate": [_P, _P, _P, _I, _I, _I, _P, _P],
            "orc_vexx": [_P, _I, _I, _P],
            "orc_exp_hermitian": [_P, _I, _P],
            "orc_generate_p1": [_P, _P, _P, _P, _I, _I, _D, _I, _P],
            "orc_vmf": [_P, _P, _I, _I, _I, _D, _D, _P],
            "orc_step": [_P, _I, _P, _D, _P, _P, _P, _P, _P, _P],
            "orc_num_threads": [],
            "orc_comb": [_I, _P, _D, _P, _P],
            "orc_population_control": [_I, _D, _I, _P, _P],
            "orc_ph_excited_1body": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P],
            "orc_ph_excited_2body": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _P],
            "orc_phmsd_mdm": [_P, _P, _P, _I, _P, _P],
            "orc_phmsd_log_overlap": [_P, _P, _P, _I, _P],
            "orc_phmsd_energy": [_P, _P, _P, _I, _D, _P, _P, _I, _P, _P],
            "orc_phmsd_gmf": [_P, _P],
            "orc_msd_brute": [_I, _I, _I, _I, _I, _P, _P, _I, _P, _P, _P, _P, _P, _I, _D, _P, _P,
                              _I, _P, _P, _P],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        L.orc_probit.restype = _D
        L.orc_average_pseudo_energy.restype = _D
        L.orc_num_threads.restype = _I
        L.orc_comb.restype = _I
        L.orc_population_control.restype = _I
        _lib = L
    return _lib


def _c(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# --- Slater determinant operations --------------------------------------------------------

def inverse_logdet(A):
    A = _c(A, np.complex128)
    nb, n, _ = A.shape
    ov = np.zeros(nb, np.complex128)
    Ainv = np.empty_like(A)
    lib().orc_inverse_logdet(_p(A), nb, n, _p(ov), _p(Ainv))
    return Ainv, ov


def log_overlap(PsiT, Phi, ovlp=None):
    PsiT = _c(PsiT, np.complex128)
    Phi = _c(Phi, np.complex128)
    W, N, n = Phi.shape
    ov = np.zeros(W, np.complex128) if ovlp is None else _c(ovlp, np.complex128).copy()
    lib().orc_log_overlap(_p(PsiT), _p(Phi), W, N, n, _p(ov))
    return ov


def mixed_density_matrix(PsiT, Phi, compact=True, ovlp=None):
    PsiT = _c(PsiT, np.complex128)
    Phi = _c(Phi, np.complex128)
    W, N, n = Phi.shape
    G = np.empty((W, n, N) if compact else (W, N, N), np.complex128)
    ov = np.zeros(W, np.complex128) if ovlp is None else _c(ovlp, np.complex128).copy()
    lib().orc_mixed_density_matrix(_p(PsiT), _p(Phi), W, N, n, _p(G), _p(ov), int(compact))
    return G, ov


def propagate(P1, V, Phi, order=6, TA="N"):
    P1 = _c(P1, np.complex128)
    V = _c(V, np.complex128)
    Phi = _c(Phi, np.complex128).copy()
    W, N, n = Phi.shape
    lib().orc_propagate(_p(P1), _p(V), _p(Phi), W, N, n, order, TA.encode())
    return Phi


def orthogonalize(Phi):
    Phi = _c(Phi, np.complex128).copy()
    W, N, n = Phi.shape
    ldet = np.zeros(W, np.complex128)
    lib().orc_orthogonalize(_p(Phi), W, N, n, _p(ldet))
    return Phi, ldet


# --- Cholesky hamiltonian operations -------------------------------------------------------

def vbias(Lnak, G, dt, c=2.0):
    Lnak = _c(Lnak, np.complex128)
    M, nel, N = Lnak.shape
    G = _c(G, np.complex128).reshape(-1, nel * N)
    W = G.shape[0]
    v = np.empty((W, M), np.complex128)
    lib().orc_vbias(_p(Lnak), _p(G), W, M, nel, N, dt, c, _p(v))
    return v


def vhs(Likn, X, dt):
    Likn = _c(Likn, np.float64)
    X = _c(X, np.complex128)
    N, _, M = Likn.shape
    W = X.shape[0]
    V = np.empty((W, N, N), np.complex128)
    lib().orc_vhs(_p(Likn), _p(X), W, N, M, dt, _p(V))
    return V


def energy(haj, Lnak, G, nup, ndown=0, closed=True, E0=0.0):
    haj = _c(haj, np.complex128)
    Lnak = _c(Lnak, np.complex128)
    G = _c(G, np.complex128)
    W = G.shape[0]
    N = haj.shape[-1]
    # CLOSED: Lnak [M][nup][N]; COLLINEAR: [M][nup][N] followed by [M][ndown][N]
    M = Lnak.size // ((nup if closed else nup + ndown) * N)
    E = np.empty((W, 3), np.complex128)
    lib().orc_energy(_p(haj), _p(Lnak), _p(G), W, N, nup, ndown, M, int(closed), E0, _p(E))
    return E


# --- fields, weights -----------------------------------------------------------------------

def probit(p):
    return lib().orc_probit(float(p))


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.empty(4, np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def field_uniforms(seed, stream, step, W, M):
    u = np.empty((W, M), np.float64)
    lib().orc_field_uniforms(seed, stream, step, W, M, _p(u))
    return u


def construct_x(u, vbias_, vMF, bound=50.0, project_vbias=False, free_projection=False):
    u = _c(u, np.float64)
    X = _c(vbias_, np.complex128).copy()
    vMF = _c(vMF, np.complex128)
    W, M = X.shape
    HW = np.empty(W, np.complex128)
    mf = np.empty(W, np.complex128)
    lib().orc_construct_x(_p(u), _p(vMF), bound, int(project_vbias), int(free_projection), W, M,
                          _p(X), _p(HW), _p(mf))
    return X, HW, mf


def walker_update(state, dt, Eshift, E0, new_ovlp, mf, HW, lower_cut=1.0, upper_cut=10.0,
                  free_projection=False, use_cp=False, weight_floor=100.0, weight_frac=0.1,
                  target_pop=None):
    """state: dict of complex128 arrays weight, phase, eloc, ovlp, theta (updated in place)."""
    W = state["weight"].shape[0]
    if target_pop is None:
        target_pop = W
    new_ovlp = _c(new_ovlp, np.complex128)
    mf = _c(mf, np.complex128)
    HW = _c(HW, np.complex128)
    lib().orc_walker_update(W, dt, Eshift, E0, _p(new_ovlp), _p(mf), _p(HW), lower_cut, upper_cut,
                            int(free_projection), int(use_cp), weight_floor, weight_frac,
                            float(target_pop), _p(state["weight"]), _p(state["phase"]),
                            _p(state["eloc"]), _p(state["ovlp"]), _p(state["theta"]), None)


def average_pseudo_energy(weight, eloc, dt):
    weight = _c(weight, np.complex128)
    eloc = _c(eloc, np.complex128)
    return lib().orc_average_pseudo_energy(weight.shape[0], _p(weight), _p(eloc), dt)


def comb(weight, r):
    weight = _c(weight, np.complex128)
    W = weight.shape[0]
    off = np.zeros(W, np.int32)
    src = np.zeros(W, np.int32)
    rc = lib().orc_comb(W, _p(weight), float(r), _p(off), _p(src))
    if rc != 0:
        raise ValueError("total walker weight collapsed")
    return off, src


def population_control(state, r):
    """Comb resampling of a walker-state dict (Phi, weight, phase, eloc, ovlp, theta) in place."""
    W = state["weight"].shape[0]
    keys = ["Phi", "phase", "eloc", "ovlp", "theta"]
    parts = [np.ascontiguousarray(state[k].reshape(W, -1)) for k in keys]
    rows = np.ascontiguousarray(np.concatenate(parts, axis=1))
    w = state["weight"]
    rc = lib().orc_population_control(W, float(r), rows.shape[1], _p(rows), _p(w))
    if rc != 0:
        raise ValueError("total walker weight collapsed")
    o = 0
    for k, part in zip(keys, parts):
        n = part.shape[1]
        state[k][...] = rows[:, o:o + n].reshape(state[k].shape)
        o += n


# --- setup ---------------------------------------------------------------------------------

def half_rotate(PsiT, Likn, h1):
    PsiT = _c(PsiT, np.complex128)
    Likn = _c(Likn, np.float64)
    h = _c(h1, np.complex128)
    nel, N = PsiT.shape
    M = Likn.shape[2]
    Lnak = np.empty((M, nel, N), np.complex128)
    haj = np.empty((nel, N), np.complex128)
    lib().orc_half_rotate(_p(PsiT), _p(Likn), _p(h), N, nel, M, _p(Lnak), _p(haj))
    return Lnak, haj


def vexx(Likn):
    Likn = _c(Likn, np.float64)
    N, _, M = Likn.shape
    out = np.empty((N, N), np.float64)
    lib().orc_vexx(_p(Likn), N, M, _p(out))
    return out


def exp_hermitian(A):
    A = _c(A, np.complex128)
    out = np.empty_like(A)
    lib().orc_exp_hermitian(_p(A), A.shape[0], _p(out))
    return out


def vmf(PsiT, Lnak, dt, c=2.0):
    PsiT = _c(PsiT, np.complex128)
    Lnak = _c(Lnak, np.complex128)
    M, nel, N = Lnak.shape
    out = np.empty(M, np.complex128)
    lib().orc_vmf(_p(PsiT), _p(Lnak), N, nel, M, dt, c, _p(out))
    return out


def generate_p1(h1, vexx_, Likn, vMF, dt, inverse=False):
    h = _c(h1, np.complex128)
    vx = _c(vexx_, np.float64)
    Likn = _c(Likn, np.float64)
    vMF = _c(vMF, np.complex128)
    N = h.shape[0]
    M = Likn.shape[2]
    P1 = np.empty((N, N), np.complex128)
    lib().orc_generate_p1(_p(h), _p(vx), _p(Likn), _p(vMF), N, M, dt, int(inverse), _p(P1))
    return P1


class _Problem(ctypes.Structure):
    _fields_ = [("N", _I), ("nel", _I), ("M", _I), ("order", _I),
                ("Likn", _P), ("Lnak", _P), ("PsiT", _P), ("P1", _P), ("vMF", _P),
                ("dt", _D), ("E0", _D), ("vbias_bound", _D), ("lower_cut", _D), ("upper_cut", _D),
                ("weight_floor", _D), ("weight_frac", _D), ("target_pop", _D),
                ("haj", _P), ("local_energy", _I)]


class OracleSystem:
    """Setup + stepping of a CLOSED single-determinant walker set, all in the oracle."""

    def __init__(self, problem, dt=0.005, order=6, vbias_bound=50.0, target_pop=None):
        self.problem = problem
        self.dt = dt
        self.order = order
        self.Likn = _c(problem.Likn, np.float64)
        self.PsiT = _c(problem.PsiT, np.complex128)
        self.Lnak, self.haj = half_rotate(self.PsiT, self.Likn, problem.h1)
        self.vexx = vexx(self.Likn)
        self.vMF = vmf(self.PsiT, self.Lnak, dt)
        self.P1 = generate_p1(problem.h1, self.vexx, self.Likn, self.vMF, dt)
        self.E0 = problem.E0
        self.vbias_bound = vbias_bound
        self.target_pop = target_pop
        self._keep = []

    def struct(self, W, local_energy=False):
        s = _Problem(self.problem.N, self.problem.nup, self.problem.M, self.order,
                     _p(self.Likn), _p(self.Lnak), _p(self.PsiT), _p(self.P1), _p(self.vMF),
                     self.dt, self.E0, self.vbias_bound, 1.0, 10.0, 100.0, 0.1,
                     float(self.target_pop if self.target_pop is not None else W),
                     _p(self.haj), int(local_energy))
        return s

    def init_walkers(self, Phi):
        """Initial properties (execute.cpp:170-183): OVLP and E from Energy, PSEUDO_ELOC_ = E
        (seedPseudoEnergy), WEIGHT = PHASE = 1; returns (state, Eshift)."""
        Phi = _c(Phi, np.complex128).copy()
        W = Phi.shape[0]
        G, ov = mixed_density_matrix(self.PsiT, Phi, compact=True)
        ov = 2.0 * ov
        E = energy(self.haj, self.Lnak, G, self.problem.nup, closed=True, E0=self.E0)
        eloc = E.sum(axis=1)
        state = dict(Phi=Phi, weight=np.ones(W, np.complex128), phase=np.ones(W, np.complex128),
                     eloc=eloc.copy(), ovlp=ov, theta=np.zeros(W, np.complex128))
        Eshift = float((np.sum(state["weight"] * eloc) / np.sum(state["weight"])).real)
        return state, Eshift

    def step(self, state, u, Eshift, local_energy=False):
        W = state["Phi"].shape[0]
        s = self.struct(W, local_energy)
        u = _c(u, np.float64)
        lib().orc_step(ctypes.byref(s), W, _p(u), Eshift, _p(state["Phi"]), _p(state["weight"]),
                       _p(state["phase"]), _p(state["eloc"]), _p(state["ovlp"]), _p(state["theta"]))

    def orthogonalize(self, state):
        Phi, _ = orthogonalize(state["Phi"])
        state["Phi"][...] = Phi
        state["ovlp"][...] = 2.0 * log_overlap(self.PsiT, Phi)


# --- density-matrix observables (numpy) -------------------------------------------------------

def rdm_measure(G, weights, closed, coeff=None):
    """One reference's contribution to OneRDM::measure (onerdm.hpp:53-58) and
    DiagonalTwoRDM::measure (diagonal_twordm.hpp:55-78), divided by sum_w weights
    (Measurements.hpp:98-121).  G [W][nspin][N][N]; returns (onerdm [nspin][N][N], the unfolded
    diagonal blocks [2 CLOSED / 3 COLLINEAR][N][N] = (aaaa), (aabb)[, (bbbb)]), walker by walker."""
    G = np.asarray(G, np.complex128)
    w = np.asarray(weights, np.complex128)
    c = w.copy() if coeff is None else w * np.asarray(coeff, np.complex128)
    W, nsp, N, _ = G.shape
    one = np.zeros((nsp, N, N), np.complex128)
    two = np.zeros((2 if closed else 3, N, N), np.complex128)
    for k in range(W):
        for s in range(nsp):
            g = G[k, s]
            one[s] += c[k] * g
            d = np.diag(g)
            two[2 * s] += c[k] * (np.outer(d, d) - g * g.T)
        dn = G[k, 0 if closed else 1]
        two[1] += c[k] * np.outer(np.diag(G[k, 0]), np.diag(dn))
    den = w.sum()
    return one / den, two / den


def fold_diag2rdm(full, closed):
    """DiagonalTwoRDM::compress (diagonal_twordm.hpp:107-140), element by element."""
    N = full.shape[1]
    out = []
    for i in range(N):
        for j in range(i + 1, N):
            out.append((full[0, i, j] + full[0, j, i]) / 2)
        for j in range(N):
            out.append(full[1, i, j])
    if not closed:
        for i in range(N):
            for j in range(i + 1, N):
                out.append((full[2, i, j] + full[2, j, i]) / 2)
    return np.array(out, np.complex128)


# --- back-propagation by time-evolved operators (numpy over the oracle's kernels) ---------------

def propagate_operators(P1, P1inv, Vs, X, Y, M, S, order=6):
    """AFQMCBasePropagator::PropagateOperators (AFQMCBasePropagator.cpp:538-684) for one spin
    block: for each V of Vs (oldest first) X <- P1 expM(V) P1 X ('N') and, with vHS scaled by
    -1, Y <- P1inv expM(-V)^H P1inv Y ('H' with P1x_inv, :585-588); then the regularization
    with the walker Slater matrices S [W][N][n] (:607-618): SY = S^H Y, Y <- S SY,
    M += Y^H X, SY = S^H X, X <- X - S SY.  X, Y, M [W][N][N]; returns new (X, Y, M)."""
    X = np.array(X, np.complex128)
    Y = np.array(Y, np.complex128)
    M = np.array(M, np.complex128)
    for V in Vs:
        X = propagate(P1, V, X, order, "N")
        Y = propagate(P1inv, -np.asarray(V), Y, order, "H")
    S = np.asarray(S, np.complex128)
    for w in range(X.shape[0]):
        SH = S[w].conj().T
        Y[w] = S[w] @ (SH @ Y[w])
        M[w] += Y[w].conj().T @ X[w]
        X[w] = X[w] - S[w] @ (SH @ X[w])
    return X, Y, M


def time_evolved_density_matrix(RefT, X, Y, M, Phi, transpose_m=False):
    """One spin block of constructTimeEvolvedMeasurementInputs (TimeEvolvedBPEstimator.hpp:
    104-131): G = M + (Ref^H X)^T (Gc conj(Y)), with Gc the compact mixed density matrix of the
    reference (RefT = Ref^H [n][N]) and the walker Phi [W][N][n].  transpose_m adds M^T."""
    Gc, _ = mixed_density_matrix(RefT, Phi)
    RefT = np.asarray(RefT, np.complex128)
    out = np.empty_like(np.asarray(M, np.complex128))
    for w in range(out.shape[0]):
        out[w] = (M[w].T if transpose_m else M[w]) + (RefT @ X[w]).T @ (Gc[w] @ Y[w].conj())
    return out


# --- particle-hole multi-determinant trial (phmsd_oracle.cpp) ------------------------------

class _Ph(ctypes.Structure):
    _fields_ = [("N", _I), ("nel", _I * 2), ("nact", _I * 2), ("orbs", _P * 2), ("nconf", _I * 2),
                ("rank", _P * 2), ("exc_start", _P * 2), ("exc", _P * 2), ("refc", _P * 2),
                ("ndet", _I), ("det_a", _P), ("det_b", _P), ("ci", _P)]


class PhTrial:
    """The oracle's view of a PHMSD trial: `ph` is a safire_amd.phmsd.PhStruct (the excitation
    numbering), `orbs` the trial orbitals [nact][N] shared by both spins."""

    def __init__(self, ph, orbs):
        self.ph = ph
        self.orbs = _c(orbs, np.complex128)
        self._keep = [self.orbs]
        k = self._keep
        arr = lambda a, t=np.int32: (k.append(_c(a, t)), k[-1])[1]
        s = _Ph()
        s.N = self.orbs.shape[1]
        for sp in range(2):
            s.nel[sp] = ph.nel[sp]
            s.nact[sp] = self.orbs.shape[0]
            s.orbs[sp] = _p(self.orbs)
            s.nconf[sp] = len(ph.rank[sp])
            s.rank[sp] = _p(arr(ph.rank[sp]))
            s.exc_start[sp] = _p(arr(ph.exc_start[sp]))
            s.exc[sp] = _p(arr(ph.exc[sp]))
            s.refc[sp] = _p(arr(ph.refc[sp]))
        s.ndet = ph.ndet
        s.det_a = _p(arr(ph.det_a))
        s.det_b = _p(arr(ph.det_b))
        s.ci = _p(arr(ph.ci, np.complex128))
        self.s = s

    @property
    def nact(self):
        return self.orbs.shape[0]

    def mixed_density_matrix(self, Phi_a, Phi_b):
        Pa, Pb = _c(Phi_a, np.complex128), _c(Phi_b, np.complex128)
        W, N = Pa.shape[:2]
        G = np.empty((W, 2 * self.nact, N), np.complex128)
        ov = np.empty(W, np.complex128)
        lib().orc_phmsd_mdm(ctypes.byref(self.s), _p(Pa), _p(Pb), W, _p(G), _p(ov))
        return G, ov

    def log_overlap(self, Phi_a, Phi_b):
        Pa, Pb = _c(Phi_a, np.complex128), _c(Phi_b, np.complex128)
        W = Pa.shape[0]
        ov = np.empty(W, np.complex128)
        lib().orc_phmsd_log_overlap(ctypes.byref(self.s), _p(Pa), _p(Pb), W, _p(ov))
        return ov

    def energy(self, haj, Lnak, Phi_a, Phi_b, E0=0.0):
        """haj [2 nact][N], Lnak [M][nact][N] (alpha) followed by the beta block."""
        haj = _c(haj, np.complex128)
        Lnak = _c(Lnak, np.complex128)
        M = Lnak.size // (2 * self.nact * self.orbs.shape[1])
        Pa, Pb = _c(Phi_a, np.complex128), _c(Phi_b, np.complex128)
        W = Pa.shape[0]
        E = np.empty((W, 3), np.complex128)
        ov = np.empty(W, np.complex128)
        lib().orc_phmsd_energy(ctypes.byref(self.s), _p(haj), _p(Lnak), M, E0, _p(Pa), _p(Pb), W,
                               _p(E), _p(ov))
        return E, ov

    def gmf(self):
        G = np.empty((2 * self.nact, self.orbs.shape[1]), np.complex128)
        lib().orc_phmsd_gmf(ctypes.byref(self.s), _p(G))
        return G


def ph_excited_1body(iexcit, refc, S, R, wgt):
    iexcit, refc = _c(iexcit, np.int32), _c(refc, np.int32)
    S, R, wgt = _c(S, np.complex128), _c(R, np.complex128), _c(wgt, np.complex128)
    nwalk, ndet, nex, nact = R.shape
    E = np.zeros(nwalk, np.complex128)
    lib().orc_ph_excited_1body(_p(iexcit), _p(refc), _p(S), _p(R), _p(wgt), nwalk, ndet, nex, nact,
                               S.shape[1], _p(E))
    return E


def ph_excited_2body(iexcit, refc, Twina, R, wgt):
    """Twina [nwalk][nelec][nchol][nact], R [nwalk][ndet][nex][nact], wgt [ndet][nwalk] ->
    (EX, EJ) [nwalk] and KE [ndet][nwalk][nchol]."""
    iexcit, refc = _c(iexcit, np.int32), _c(refc, np.int32)
    T, R, wgt = _c(Twina, np.complex128), _c(R, np.complex128), _c(wgt, np.complex128)
    nwalk, ndet, nex, nact = R.shape
    nelec, nchol = T.shape[1], T.shape[2]
    EX = np.zeros(nwalk, np.complex128)
    EJ = np.zeros(nwalk, np.complex128)
    KE = np.zeros((ndet, nwalk, nchol), np.complex128)
    lib().orc_ph_excited_2body(_p(iexcit), _p(refc), _p(T), _p(R), _p(wgt), nwalk, ndet, nex, nact,
                               nelec, nchol, _p(EX), _p(EJ), _p(KE))
    return EX, EJ, KE


def msd_brute(orbs, occa, occb, ci, haj, Lnak, Phi_a, Phi_b, E0=0.0):
    """Explicit sum over determinants: G [W][2 nact][N], E [W][3], log overlap [W]."""
    orbs = _c(orbs, np.complex128)
    occa, occb = _c(occa, np.int32), _c(occb, np.int32)
    ci = _c(ci, np.complex128)
    haj, Lnak = _c(haj, np.complex128), _c(Lnak, np.complex128)
    Pa, Pb = _c(Phi_a, np.complex128), _c(Phi_b, np.complex128)
    nact, N = orbs.shape
    W, _, nup = Pa.shape
    ndn = Pb.shape[2]
    M = Lnak.size // (2 * nact * N)
    G = np.empty((W, 2 * nact, N), np.complex128)
    E = np.empty((W, 3), np.complex128)
    ov = np.empty(W, np.complex128)
    lib().orc_msd_brute(N, nup, ndn, nact, nact, _p(orbs), _p(orbs), len(ci), _p(occa), _p(occb),
                        _p(ci), _p(haj), _p(Lnak), M, E0, _p(Pa), _p(Pb), W, _p(G), _p(E), _p(ov))
    return G, E, ov


# --- multi-determinant NOMSD trial (NOMSD.cpp:39-61, 159-333, 338-520) ------------------------
# A trial of ndet determinants, each given per spin block in PsiT form: dets[d] = (Pa [nup][N],
# Pb [ndown][N] or None for CLOSED).  Everything below combines the single-determinant oracle
# functions above, in the order the reference combines its determinants.

def _log_conj_c(c):
    """log(conj(c)), with |c| < 1e-12 replaced by 1e-12 (NOMSD.cpp:316)."""
    return np.log(1e-12 + 0j) if abs(c) < 1e-12 else np.log(np.conj(c))


def msd_det_log_overlaps(dets, ci, Phi_a, Phi_b=None):
    """[ndet][W] log <D_d|Phi> + log conj(c_d): CLOSED doubles the alpha block (NOMSD.cpp:305-322)."""
    W = Phi_a.shape[0]
    out = np.zeros((len(dets), W), np.complex128)
    for d, (Pa, Pb) in enumerate(dets):
        ov = log_overlap(Pa, Phi_a)
        ov = 2.0 * ov if Pb is None else ov + log_overlap(Pb, Phi_b)
        out[d] = ov + _log_conj_c(ci[d])
    return out


def msd_log_overlap(dets, ci, Phi_a, Phi_b=None):
    """NOMSD::Log_Overlap with ndet > 1: log sum_d exp(l_d) about the largest real part
    (NOMSD.cpp:324-331)."""
    lo = msd_det_log_overlaps(dets, ci, Phi_a, Phi_b)
    lm = lo[np.argmax(lo.real, axis=0), np.arange(lo.shape[1])]
    return lm + np.log(np.sum(np.exp(lo - lm), axis=0))


def msd_mixed_density_matrix(dets, ci, Phi_a, Phi_b, log_m):
    """NOMSD::MixedDensityMatrix with ndet > 1 (NOMSD.cpp:240-277): full G [W][nspin*N*N] =
    sum_d w_d G_d / sum_d w_d with w_d = conj(c_d) exp(ov_d - log_m), ov_d the determinant's
    log overlap (alpha doubled for CLOSED), log_m the walkers' OVLP; Ov = sum_d w_d + log_m, as
    the reference leaves it."""
    W, N, _ = Phi_a.shape
    nspin = 1 if dets[0][1] is None else 2
    G = np.zeros((W, nspin * N * N), np.complex128)
    Ov = np.zeros(W, np.complex128)
    for d, (Pa, Pb) in enumerate(dets):
        Ga, ov = mixed_density_matrix(Pa, Phi_a, compact=False)
        blocks = [Ga.reshape(W, N * N)]
        if Pb is None:
            ov = 2.0 * ov
        else:
            Gb, ovb = mixed_density_matrix(Pb, Phi_b, compact=False)
            ov = ov + ovb
            blocks.append(Gb.reshape(W, N * N))
        wd = np.conj(ci[d]) * np.exp(ov - log_m)
        Ov += wd
        G += wd[:, None] * np.concatenate(blocks, axis=1)
    G /= Ov[:, None]
    return G, Ov + log_m


def vbias_full(Likn, G, dt, c):
    """Real3IndexFactorization::vbias for a full G (ndet > 1, :395-423): v[w][m] =
    sqrt(dt) c sum_s sum_ik G[w][s][i][k] L[i][k][m]."""
    Likn = _c(Likn, np.float64)
    N, _, M = Likn.shape
    G = _c(G, np.complex128)
    W = G.shape[0]
    nspin = G.shape[1] // (N * N)
    Lm = np.ascontiguousarray(np.transpose(Likn, (2, 0, 1)).reshape(M, 1, N * N))
    Lm = np.ascontiguousarray(np.repeat(Lm, nspin, axis=1).reshape(M, nspin * N, N))
    return vbias(Lm.astype(np.complex128), G.reshape(W, nspin * N, N), dt, c)


def msd_half_rotated(dets, Likn, h1):
    """Per-determinant half-rotated tensors (RealDenseHamiltonian.cpp:141-163): for determinant
    d, Lnak_d (the alpha block [M][nup][N], then for COLLINEAR the beta block [M][ndown][N],
    the layout energy() takes) and haj_d [nel][N] (alpha rows then beta)."""
    out = []
    for Pa, Pb in dets:
        La, ha = half_rotate(Pa, Likn, h1)
        if Pb is None:
            out.append((La, ha))
        else:
            Lb, hb = half_rotate(Pb, Likn, h1)
            out.append((np.concatenate([La.ravel(), Lb.ravel()]), np.concatenate([ha, hb])))
    return out


def msd_energy(dets, ci, rot, Phi_a, Phi_b=None, E0=0.0):
    """NOMSD::Energy with ndet > 1 (NOMSD.cpp:185-216): each determinant's local energy from
    its compact G and its half-rotated tensors, combined with w_d = conj(c_d) exp(ov_d - max_d
    Re ov_d); E [W][3], Ov = log sum_d conj(c_d) exp(ov_d - log_m) + log_m."""
    W = Phi_a.shape[0]
    closed = dets[0][1] is None
    nup = dets[0][0].shape[0]
    ndn = 0 if closed else dets[0][1].shape[0]
    ovs, Es = [], []
    for (Pa, Pb), (Lnak, haj) in zip(dets, rot):
        Ga, ov = mixed_density_matrix(Pa, Phi_a, compact=True)
        if closed:
            ov, G = 2.0 * ov, Ga
        else:
            Gb, ovb = mixed_density_matrix(Pb, Phi_b, compact=True)
            ov, G = ov + ovb, np.concatenate([Ga, Gb], axis=1)
        ovs.append(ov)
        Es.append(energy(haj, Lnak, G, nup, ndn, closed=closed, E0=E0))
    ovs = np.array(ovs)
    lm = ovs[np.argmax(ovs.real, axis=0), np.arange(W)]
    wd = np.conj(np.asarray(ci))[:, None] * np.exp(ovs - lm)
    E = np.einsum("dw,dwk->wk", wd, np.array(Es)) / wd.sum(0)[:, None]
    return E, np.log(wd.sum(0)) + lm


def msd_vmf(dets, ci, Likn, dt):
    """NOMSD::vMF with ndet > 1 (NOMSD.cpp:438-516) and zero_imag: the trial density
    sum_{p <= q} Re(conj(c_q) c_p e^{ov_qp}) G_qp / sum of the same weights, ov_qp and G_qp the
    overlap (alpha block only doubled nowhere) and full density matrix of determinant q against
    determinant p as the walker; v = vbias of that full G, imaginary part dropped."""
    N = dets[0][0].shape[1]
    closed = dets[0][1] is None
    nspin = 1 if closed else 2
    Gs = np.zeros((1, nspin * N * N), np.complex128)
    Os = 0.0
    for p in range(len(dets)):
        for q in range(p, len(dets)):
            Ga, ov = mixed_density_matrix(dets[q][0], np.conj(dets[p][0].T)[None], compact=False)
            blocks = [Ga.reshape(1, N * N)]
            if not closed:
                Gb, ovb = mixed_density_matrix(dets[q][1], np.conj(dets[p][1].T)[None],
                                               compact=False)
                ov = ov + ovb
                blocks.append(Gb.reshape(1, N * N))
            scl = float((np.conj(ci[q]) * ci[p] * np.exp(ov[0])).real)
            Os += scl
            Gs += scl * np.concatenate(blocks, axis=1)
    Gs /= Os
    v = vbias_full(Likn, Gs, dt, 2.0 if closed else 1.0)[0]
    return v.real + 0j
