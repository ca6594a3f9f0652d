"""Python host mirror of the C ABI in include/afq.h (libafq_hip.so).

The classes follow SAFIRE's interfaces for the walker step so tests read like the reference's:
  HamiltonianOperations  ~ Real3IndexFactorization      vbias / vHS / energy / vMF / P1
  det_ops                ~ det_ops free functions       MixedDensityMatrix / Log_Overlap /
                                                        Propagate / orthogonalize / inverse_logdet
  Propagator             ~ AFQMCBasePropagator          Propagate / Orthogonalize / Energy
  WalkerSet              ~ WalkerSet (device-resident walker matrices and properties)

Device memory and the stream come from PyTorch (plumbing only); every computation runs in the
HIP library.  There is no CPU fallback: importing this module on a machine without the built
library raises, and every call requires a GPU.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libafq_hip.so")

AFQ_CLOSED, AFQ_COLLINEAR = 0, 1
AFQ_FP64, AFQ_MIXED = 0, 1

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


class AfqError(RuntimeError):
    pass


class UpdateParams(ctypes.Structure):
    _fields_ = [("dt", _D), ("E0", _D), ("lower_cut", _D), ("upper_cut", _D),
                ("weight_floor", _D), ("weight_frac", _D), ("target_pop", _D),
                ("free_projection", _I), ("use_cp", _I)]


class PropParams(ctypes.Structure):
    _fields_ = [("dt", _D), ("order", _I), ("vbias_bound", _D), ("lower_cut", _D),
                ("upper_cut", _D), ("weight_floor", _D), ("weight_frac", _D), ("target_pop", _D),
                ("free_projection", _I), ("use_cp", _I), ("project_force_bias", _I),
                ("seed", ctypes.c_uint64), ("stream", ctypes.c_uint32), ("local_energy", _I)]


class Walkers(ctypes.Structure):
    _fields_ = [("Phi", _P), ("Phi_b", _P), ("weight", _P), ("phase", _P), ("eloc", _P),
                ("ovlp", _P), ("theta", _P), ("W", _I), ("phi_stride", ctypes.c_int64),
                ("prop_stride", ctypes.c_int64), ("E1", _P), ("EXX", _P), ("EJ", _P)]


# name -> argtypes (restype int unless noted)
_SIGS = {
    "afq_abi_version": [],
    "afq_last_error": [],
    "afq_ctx_create": [_I, _P, ctypes.POINTER(_P)],
    "afq_ctx_destroy": [_P],
    "afq_ctx_synchronize": [_P],
    "afq_ctx_scratch_bytes": [_P, ctypes.POINTER(ctypes.c_uint64)],
    "afq_tuning_report": [ctypes.c_char_p, _I],
    "afq_tune_force": [ctypes.c_char_p, _I],
    "afq_ham_create": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _D, _I, ctypes.POINTER(_P)],
    "afq_ham_destroy": [_P],
    "afq_ham_tensors": [_P, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_P),
                        ctypes.POINTER(_I)],
    "afq_vbias": [_P, _P, _I, _D, _I, _P],
    "afq_vhs": [_P, _P, _I, _D, _I, _P],
    "afq_energy": [_P, _P, _I, _I, _I, _I, _I, _P],
    "afq_energy_spin": [_P, _I, _P, _I, _I, _I, _I, _I, _P, _P],
    "afq_vmf": [_P, _D, _P],
    "afq_one_body_propagator": [_P, _P, _D, _P],
    "afq_mixed_density_matrix": [_P, _P, _I, _I, _P, ctypes.c_int64, _I, _P, _P, _I],
    "afq_mixed_density_matrix_noalloc": [_P, _P, _I, _I, _P, ctypes.c_int64, _I, _P, _P, _I],
    "afq_log_overlap": [_P, _P, _I, _I, _P, ctypes.c_int64, _I, _P],
    "afq_inverse_logdet": [_P, _P, _I, _I, _P, _P],
    "afq_propagate": [_P, _P, _P, _P, ctypes.c_int64, _I, _I, _I, _I, ctypes.c_char, _I],
    "afq_orthogonalize": [_P, _P, ctypes.c_int64, _I, _I, _I, _P],
    "afq_field_uniforms": [_P, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _I, _I, _P],
    "afq_construct_x": [_P, _P, _P, _I, _I, _D, _I, _I, _P, _P, _P],
    "afq_walker_update": [_P, _I, ctypes.POINTER(UpdateParams), _P, _P, _P, _P, _P, _P, _P, _P,
                          _P],
    "afq_average_pseudo_energy": [_P, _I, _P, _P, _D, _D, _P, _P],
    "afq_prop_create": [_P, _I, ctypes.POINTER(PropParams), ctypes.POINTER(_P)],
    "afq_prop_destroy": [_P],
    "afq_prop_views": [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)],
    "afq_prop_step": [_P, ctypes.POINTER(Walkers), _P, _P],
    "afq_prop_orthogonalize": [_P, ctypes.POINTER(Walkers)],
    "afq_prop_set_step": [_P, ctypes.c_uint32],
    "afq_prop_energy": [_P, ctypes.POINTER(Walkers), _I, _P, _P],
    "afq_prop_step_graph": [_P, ctypes.POINTER(Walkers), _P, _I, _I, _D, _P],
    "afq_prop_population_control": [_P, ctypes.POINTER(Walkers), _D],
    "afq_prop_status": [_P, ctypes.POINTER(_I)],
    "afq_prop_bound_stats": [_P, ctypes.POINTER(ctypes.c_int64), _I],
    "afq_prop_enable_backprop": [_P, _I],
    "afq_prop_bp_anchor": [_P, ctypes.POINTER(Walkers)],
    "afq_prop_backpropagate": [_P, ctypes.POINTER(Walkers), _I, _I, _I, _P, _P],
    "afq_prop_bp_views": [_P, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_I)],
    "afq_rdm_measure": [_P, _I, _I, _I, _P, _P, _P, _I, _P, _P],
    "afq_prop_bp_measure": [_P, ctypes.POINTER(Walkers), _I, _I, _I, _P, _P],
    "afq_prop_vhs_timing": [_P, _I],
    "afq_prop_vhs_ms": [_P, ctypes.POINTER(ctypes.c_float)],
    "afq_prop_te_anchor": [_P, _I],
    "afq_prop_te_propagate": [_P, ctypes.POINTER(Walkers), _I],
    "afq_prop_te_measure": [_P, ctypes.POINTER(Walkers), _I, _I, _P, _P],
    "afq_prop_te_views": [_P, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_P),
                          ctypes.POINTER(_P)],
    "afq_prop_mixed_measure": [_P, ctypes.POINTER(Walkers), _P, _P],
    "afq_prop_number_of_references": [_P, ctypes.POINTER(_I)],
    "afq_prop_create_phmsd": [_P, _I, ctypes.POINTER(PropParams), ctypes.POINTER(_P)],
    "afq_phmsd_create": [_P, _P, ctypes.POINTER(_P)],
    "afq_phmsd_destroy": [_P],
    "afq_phmsd_mixed_density_matrix": [_P, _P, _P, ctypes.c_int64, _I, _P, _P],
    "afq_phmsd_log_overlap": [_P, _P, _P, ctypes.c_int64, _I, _P],
    "afq_phmsd_energy": [_P, _P, _P, ctypes.c_int64, _I, _P, _P],
    "afq_phmsd_vmf": [_P, _D, _P],
    "afq_prop_create_nomsd": [_P, _I, ctypes.POINTER(PropParams), ctypes.POINTER(_P)],
    "afq_nomsd_create": [_P, _I, _P, _P, ctypes.POINTER(_P)],
    "afq_nomsd_destroy": [_P],
    "afq_nomsd_log_overlap": [_P, _P, _P, ctypes.c_int64, _I, _P],
    "afq_nomsd_mixed_density_matrix": [_P, _P, _P, ctypes.c_int64, _I, _P, _P, _P],
    "afq_nomsd_vbias": [_P, _P, _P, ctypes.c_int64, _I, _P, _D, _I, _P],
    "afq_nomsd_energy": [_P, _P, _P, ctypes.c_int64, _I, _I, _P, _P],
    "afq_nomsd_vmf": [_P, _D, _P],
    "afq_vbias_full": [_P, _P, _I, _D, _I, _P],
    "afq_ham_create_spin": [_P, _I, _I, _I, _I, _I, _P, _I, _P, _I, _P, _D, _I, ctypes.POINTER(_P)],
}

_lib = None


ABI_VERSION = 302   # afq_abi_version() of the library these bindings describe
NOALLOC_SCRATCH_BYTES = 32 << 20   # AFQ_NOALLOC_SCRATCH_BYTES


def lib():
    """The loaded library; raises when it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AfqError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(make -C safire_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _I
        L.afq_last_error.restype = ctypes.c_char_p
        if L.afq_abi_version() != ABI_VERSION:
            raise AfqError(f"{LIB_PATH} has ABI {L.afq_abi_version()}, this mirror binds "
                           f"{ABI_VERSION}: rebuild it")
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def tuning_report() -> str:
    buf = ctypes.create_string_buffer(1 << 16)
    _ck(lib().afq_tuning_report(buf, len(buf)))
    return buf.value.decode()


def tune_force(site, candidate: int = -1):
    """Pin the autotuner's choice at a product site ("name" or "name:f32"/"name:f64"); a
    negative candidate unpins the site, site None unpins every site."""
    _ck(lib().afq_tune_force(None if site is None else site.encode(), int(candidate)))


def _ck(rc):
    if rc != 0:
        raise AfqError(f"afq error {rc}: {lib().afq_last_error().decode()}")


# ---------------------------------------------------------------------------------------------
# device arrays (torch is plumbing: allocation, host<->device copies, the stream)
# ---------------------------------------------------------------------------------------------

def _torch():
    import torch
    return torch


def device_tensor(a, dtype=None):
    """Copy a numpy array to the GPU (complex stays complex)."""
    torch = _torch()
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(a).to("cuda")


def empty(shape, dtype):
    torch = _torch()
    return torch.empty(shape, dtype=dtype, device="cuda")


def zeros(shape, dtype):
    torch = _torch()
    return torch.zeros(shape, dtype=dtype, device="cuda")


def ptr(t):
    if t is None:
        return None
    assert t.is_contiguous(), "arrays must be contiguous"
    return ctypes.c_void_p(t.data_ptr())


def host(t):
    _torch().cuda.synchronize()
    return t.detach().cpu().numpy()


# ---------------------------------------------------------------------------------------------
# context
# ---------------------------------------------------------------------------------------------

_STREAM = {}


def shared_stream(device: int = 0):
    """One non-default stream per device, made torch's current stream, on which every context
    and every torch operation of this process run: library calls and torch copies are then in
    program order, and the stream can be captured into HIP graphs."""
    torch = _torch()
    if device not in _STREAM:
        _STREAM[device] = torch.cuda.Stream(device=device)
    torch.cuda.set_stream(_STREAM[device])
    return _STREAM[device]


class Context:
    """afq_ctx on the process-wide stream of `device` (see shared_stream), or on an explicit
    torch stream (the serving loop gives every job its own); torch work on that context's
    buffers must then be issued on the same stream."""

    def __init__(self, device: int = 0, stream=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise AfqError("no GPU available: the AFQMC step runs only on the device")
        torch.cuda.set_device(device)
        self.stream = shared_stream(device) if stream is None else stream
        h = _P()
        _ck(lib().afq_ctx_create(device, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h)))
        self.h = h

    def synchronize(self):
        _ck(lib().afq_ctx_synchronize(self.h))

    def scratch_bytes(self):
        v = ctypes.c_uint64()
        _ck(lib().afq_ctx_scratch_bytes(self.h, ctypes.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            lib().afq_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------------
# det_ops
# ---------------------------------------------------------------------------------------------

def _blocks(Phi):
    """(pointer, walker stride) of a [W][N][n] walker array whose [N][n] blocks are contiguous
    and whose walkers may be further apart (a SlaterMatrices view of a walker buffer)."""
    W, N, n = Phi.shape
    st = Phi.stride()
    if W > 1:
        assert st[1] == n and st[2] == 1, "each walker's [N][n] block must be contiguous"
    return ctypes.c_void_p(Phi.data_ptr()), (st[0] if W > 1 else N * n)


class det_ops:  # noqa: N801  (namespace mirroring sfqmc::afqmc::det_ops)
    @staticmethod
    def MixedDensityMatrix(ctx, PsiT, Phi, G, ovlp, compact=True):
        n, N = PsiT.shape
        W = Phi.shape[0]
        pp, bs = _blocks(Phi)
        _ck(lib().afq_mixed_density_matrix(ctx.h, ptr(PsiT), n, N, pp, bs, W, ptr(G), ptr(ovlp),
                                           int(compact)))

    @staticmethod
    def MixedDensityMatrixNoAlloc(ctx, PsiT, Phi, G, ovlp, compact=True):
        """MixedDensityMatrix through the allocation-free entry point (fixed per-GPU workspace
        of NOALLOC_SCRATCH_BYTES; larger calls raise)."""
        n, N = PsiT.shape
        W = Phi.shape[0]
        pp, bs = _blocks(Phi)
        _ck(lib().afq_mixed_density_matrix_noalloc(ctx.h, ptr(PsiT), n, N, pp, bs, W, ptr(G),
                                                   ptr(ovlp), int(compact)))

    @staticmethod
    def Log_Overlap(ctx, PsiT, Phi, ovlp):
        n, N = PsiT.shape
        pp, bs = _blocks(Phi)
        _ck(lib().afq_log_overlap(ctx.h, ptr(PsiT), n, N, pp, bs, Phi.shape[0], ptr(ovlp)))

    @staticmethod
    def inverse_logdet(ctx, A, ovlp, Ainv):
        nb, n, _ = A.shape
        _ck(lib().afq_inverse_logdet(ctx.h, ptr(A), nb, n, ptr(ovlp), ptr(Ainv)))

    @staticmethod
    def Propagate(ctx, P1, V, Phi, order=6, TA="N", precision=AFQ_FP64):
        W, N, n = Phi.shape
        pp, bs = _blocks(Phi)
        _ck(lib().afq_propagate(ctx.h, ptr(P1), ptr(V), pp, bs, W, N, n, order, TA.encode(),
                                precision))

    @staticmethod
    def orthogonalize(ctx, Phi, ldet):
        W, N, n = Phi.shape
        pp, bs = _blocks(Phi)
        _ck(lib().afq_orthogonalize(ctx.h, pp, bs, W, N, n, ptr(ldet)))


def field_uniforms(ctx, seed, stream, step, u):
    W, M = u.shape
    _ck(lib().afq_field_uniforms(ctx.h, seed, stream, step, W, M, ptr(u)))


def construct_x(ctx, u, vMF, X, HW, mf, bound=50.0, project_vbias=False, free_projection=False):
    W, M = X.shape
    _ck(lib().afq_construct_x(ctx.h, ptr(u), ptr(vMF), W, M, bound, int(project_vbias),
                              int(free_projection), ptr(X), ptr(HW), ptr(mf)))


def walker_update(ctx, params: UpdateParams, Eshift, new_ovlp, mf, HW, weight, phase, eloc, ovlp,
                  theta):
    _ck(lib().afq_walker_update(ctx.h, weight.shape[0], ctypes.byref(params), ptr(Eshift),
                                ptr(new_ovlp), ptr(mf), ptr(HW), ptr(weight), ptr(phase),
                                ptr(eloc), ptr(ovlp), ptr(theta)))


def average_pseudo_energy(ctx, weight, eloc, dt, out, relax=0.0, Eshift=None):
    _ck(lib().afq_average_pseudo_energy(ctx.h, weight.shape[0], ptr(weight), ptr(eloc), dt, relax,
                                        ptr(out), ptr(Eshift)))


# ---------------------------------------------------------------------------------------------
# HamiltonianOperations (Real3IndexFactorization)
# ---------------------------------------------------------------------------------------------

class HamiltonianOperations:
    """Real3IndexFactorization over h1 [N][N] and Likn [N][N][M], or spin-dependent integrals
    h1 [2][N][N] and/or Likn [2][N][N][M] (alpha, beta; COLLINEAR walkers)."""

    def __init__(self, ctx, h1, Likn, PsiT, nup, ndown=0, walker_type=AFQ_CLOSED, E0=0.0,
                 precision=AFQ_FP64):
        torch = _torch()
        self.ctx = ctx
        self.nspin_h1 = h1.shape[0] if h1.ndim == 3 else 1
        self.nspin_L = Likn.shape[0] if Likn.ndim == 4 else 1
        N = h1.shape[-1]
        M = Likn.shape[-1]
        self.N, self.M, self.nup, self.ndown = N, M, nup, (ndown if walker_type == AFQ_COLLINEAR else 0)
        self.walker_type = walker_type
        self.precision = precision
        self.E0 = E0
        dL = device_tensor(Likn, np.float64)
        dh = device_tensor(h1, np.float64)
        dP = device_tensor(PsiT, np.complex128)
        h = _P()
        if self.nspin_h1 == 1 and self.nspin_L == 1:
            _ck(lib().afq_ham_create(ctx.h, walker_type, N, nup, ndown, M, ptr(dL), ptr(dh),
                                     ptr(dP), E0, precision, ctypes.byref(h)))
        else:
            _ck(lib().afq_ham_create_spin(ctx.h, walker_type, N, nup, ndown, M, ptr(dL),
                                          self.nspin_L, ptr(dh), self.nspin_h1, ptr(dP), E0,
                                          precision, ctypes.byref(h)))
        ctx.synchronize()
        del dL, dh, dP
        self.h = h
        lp, hp, vp, rt = _P(), _P(), _P(), _I()
        _ck(lib().afq_ham_tensors(h, ctypes.byref(lp), ctypes.byref(hp), ctypes.byref(vp),
                                  ctypes.byref(rt)))
        self.real_trial = bool(rt.value)
        self._views = (lp.value, hp.value, vp.value)
        self._torch = torch

    @property
    def nel(self):
        return self.nup + (self.ndown if self.walker_type == AFQ_COLLINEAR else 0)

    def number_of_cholesky_vectors(self):
        return self.M

    def energy_offset(self):
        return self.E0

    def tensors(self):
        """Host copies of (Lnak [M][n][N] per spin block, haj [nel][N], vexx [N][N])."""
        torch = self._torch
        self.ctx.synchronize()
        L = torch.empty(self.M * self.nel * self.N, dtype=torch.complex128, device="cuda")
        H = torch.empty(self.nel * self.N, dtype=torch.complex128, device="cuda")
        V = torch.empty(self.N * self.N, dtype=torch.float64, device="cuda")
        for t, p in ((L, self._views[0]), (H, self._views[1]), (V, self._views[2])):
            _copy_d2d(t.data_ptr(), p, t.numel() * t.element_size())
        return host(L), host(H).reshape(self.nel, self.N), host(V).reshape(self.N, self.N)

    def vbias(self, G, v, dt, precision=None):
        W = G.shape[0]
        _ck(lib().afq_vbias(self.h, ptr(G), W, dt, self.precision if precision is None else precision,
                            ptr(v)))

    def vHS(self, X, V, dt, precision=None):
        W = X.shape[0]
        _ck(lib().afq_vhs(self.h, ptr(X), W, dt, self.precision if precision is None else precision,
                          ptr(V)))

    def vbias_full(self, G, v, dt, precision=None):
        """vbias of full density matrices G [W][nspin N N] (the ndet > 1 branch)."""
        W = G.shape[0]
        _ck(lib().afq_vbias_full(self.h, ptr(G), W, dt,
                                 self.precision if precision is None else precision, ptr(v)))

    def energy(self, E, G, addH1=True, addEJ=True, addEXX=True, precision=None):
        W = G.shape[0]
        _ck(lib().afq_energy(self.h, ptr(G), W, int(addH1), int(addEJ), int(addEXX),
                             self.precision if precision is None else precision, ptr(E)))

    def energy_spin(self, spin, E, G, EJn=None, addH1=True, addEJ=True, addEXX=True,
                    precision=None):
        """Per-spin energy of one block's compact G [W][n_spin][N]; EJn [W][M] receives the
        Coulomb vectors when addEJ."""
        W = G.shape[0]
        _ck(lib().afq_energy_spin(self.h, int(spin), ptr(G), W, int(addH1), int(addEJ),
                                  int(addEXX), self.precision if precision is None else precision,
                                  ptr(E), ptr(EJn)))

    def vMF(self, dt, v):
        _ck(lib().afq_vmf(self.h, dt, ptr(v)))

    def getOneBodyPropagator(self, vMF, dt, P1):
        _ck(lib().afq_one_body_propagator(self.h, ptr(vMF), dt, ptr(P1)))

    def close(self):
        if getattr(self, "h", None):
            lib().afq_ham_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _copy_d2d(dst, src, nbytes):
    """Device-to-device copy of a library-owned view into a torch buffer."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [_P, _P, ctypes.c_size_t, _I]
    hip.hipMemcpy.restype = _I
    rc = hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, 3)
    if rc != 0:
        raise AfqError(f"hipMemcpy failed ({rc})")


def rdm_measure(ctx, G, weights, walker_type, coeff=None, onerdm=True, diag2rdm=True,
                out=None):
    """One OneRDM / DiagonalTwoRDM measurement from per-walker density matrices G
    [W][nspin][N][N]: (onerdm [nspin][N][N] or None, diag2rdm [2 or 3][N][N] or None), each
    sum_w weight_w coeff_w (...) / sum_w weight_w.  The diagonal two-body blocks are (aaaa),
    (aabb)[, (bbbb)], unfolded (fold_diag2rdm gives the reference's layout).  With out =
    (onerdm, diag2rdm) from an earlier call the measurement is added to those tensors."""
    torch = _torch()
    W, nsp, N = G.shape[0], G.shape[1], G.shape[2]
    if nsp != (2 if walker_type == AFQ_COLLINEAR else 1):
        raise AfqError("G's spin dimension does not match the walker type")
    if out is not None:
        one, two = out
    else:
        one = torch.empty((nsp, N, N), dtype=torch.complex128, device="cuda") if onerdm else None
        two = (torch.empty((nsp + 1, N, N), dtype=torch.complex128, device="cuda") if diag2rdm
               else None)
    _ck(lib().afq_rdm_measure(ctx.h, int(walker_type), N, W, ptr(G), ptr(weights), ptr(coeff),
                              int(out is not None), ptr(one), ptr(two)))
    return one, two


def fold_diag2rdm(full, walker_type):
    """DiagonalTwoRDM::compress (diagonal_twordm.hpp:103-140): for each row i the symmetrized
    (aaaa) entries j > i, then the whole (aabb) row; for COLLINEAR the symmetrized (bbbb) upper
    triangle follows.  `full` is a host array [2 or 3][N][N]."""
    full = np.asarray(full)
    N = full.shape[1]
    iu = np.triu_indices(N, 1)
    out = []
    for i in range(N):
        out.append((full[0, i, i + 1:] + full[0, i + 1:, i]) / 2)
        out.append(full[1, i, :])
    if walker_type == AFQ_COLLINEAR:
        out.append(((full[2] + full[2].T) / 2)[iu])
    return np.concatenate(out)


# ---------------------------------------------------------------------------------------------
# walker set + propagator
# ---------------------------------------------------------------------------------------------

class WalkerSet:
    """Device-resident walkers: Slater matrices and the per-walker properties of
    WalkerConfig.hpp:28-48 (WEIGHT, PHASE, PSEUDO_ELOC_, OVLP, THETA)."""

    def __init__(self, Phi, Phi_b=None):
        torch = _torch()
        self.Phi = device_tensor(Phi, np.complex128)
        self.Phi_b = None if Phi_b is None else device_tensor(Phi_b, np.complex128)
        W = self.Phi.shape[0]
        self.W = W
        self.weight = torch.ones(W, dtype=torch.complex128, device="cuda")
        self.phase = torch.ones(W, dtype=torch.complex128, device="cuda")
        self.eloc = torch.zeros(W, dtype=torch.complex128, device="cuda")
        self.ovlp = torch.zeros(W, dtype=torch.complex128, device="cuda")
        self.theta = torch.zeros(W, dtype=torch.complex128, device="cuda")
        self._c = None

    def cstruct(self):
        self._c = Walkers(ptr(self.Phi), ptr(self.Phi_b), ptr(self.weight), ptr(self.phase),
                          ptr(self.eloc), ptr(self.ovlp), ptr(self.theta), self.W, 0, 0,
                          ptr(self.E1), ptr(self.EXX), ptr(self.EJ))
        return self._c

    # the energy-component columns are optional (the local-energy step fills them when given)
    E1 = EXX = EJ = None

    def keep_energy_components(self):
        """Allocate the E1_, EXX_, EJ_ columns that a local-energy step fills."""
        torch = _torch()
        for k in ("E1", "EXX", "EJ"):
            setattr(self, k, torch.zeros(self.W, dtype=torch.complex128, device="cuda"))

    def phi_stride(self):
        return 0


# column offsets of SAFIRE's walker row at zero temperature (WalkerSetBase.cpp:75-121): the
# Slater matrices, then WEIGHT, PHASE, PSEUDO_ELOC_, E1_, EXX_, EJ_, OVLP, LOGSCL_UP,
# LOGSCL_DN, IS_UNITARY, THETA
_ROW_PROPS = ("weight", "phase", "eloc", "E1", "EXX", "EJ", "ovlp", "logscl_up", "logscl_dn",
              "is_unitary", "theta")


class WalkerBuffer:
    """Walkers in SAFIRE's own buffer layout: one complex128 row per walker,
    [Phi_a (N x nup) | Phi_b (N x ndown) | WEIGHT, PHASE, PSEUDO_ELOC_, E1_, EXX_, EJ_, OVLP,
    LOGSCL_UP, LOGSCL_DN, IS_UNITARY, THETA] (WalkerSetBase.cpp:75-121).  Phi / Phi_b are the
    strided SlaterMatrices views and the properties the strided getProperty columns
    (WalkerSetBase.h:256-276, 463-481); the library reads and writes them in place through the
    walker strides of afq_walkers.  `pad` extra complex elements per row test that nothing
    assumes the row length."""

    def __init__(self, Phi, Phi_b=None, pad=0):
        torch = _torch()
        W, N, nup = Phi.shape
        ndn = 0 if Phi_b is None else Phi_b.shape[2]
        self.W, self.N, self.nup, self.ndn = W, N, nup, ndn
        self.row = N * (nup + ndn) + len(_ROW_PROPS) + pad
        self.buf = torch.zeros((W, self.row), dtype=torch.complex128, device="cuda")
        self.Phi = self.buf[:, :N * nup].view(W, N, nup)
        self.Phi.copy_(torch.from_numpy(np.ascontiguousarray(Phi, dtype=np.complex128)))
        self.Phi_b = None
        if ndn:
            self.Phi_b = self.buf[:, N * nup:N * (nup + ndn)].view(W, N, ndn)
            self.Phi_b.copy_(torch.from_numpy(np.ascontiguousarray(Phi_b, dtype=np.complex128)))
        base = N * (nup + ndn)
        for i, name in enumerate(_ROW_PROPS):
            setattr(self, name, self.buf[:, base + i])
        self.weight.fill_(1.0)
        self.phase.fill_(1.0)
        self._c = None

    def phi_stride(self):
        return self.row

    def cstruct(self):
        def a(t):
            return ctypes.c_void_p(t.data_ptr()) if t is not None else None
        self._c = Walkers(a(self.Phi), a(self.Phi_b), a(self.weight), a(self.phase), a(self.eloc),
                          a(self.ovlp), a(self.theta), self.W, self.row, self.row, a(self.E1),
                          a(self.EXX), a(self.EJ))
        return self._c


class Propagator:
    """AFQMCBasePropagator over the hamiltonian's single-determinant trial, or over a
    multi-determinant trial built on the same hamiltonian: PHMSD (safire_amd.phmsd.PHMSD) or
    NOMSD (safire_amd.nomsd.NOMSD)."""

    def __init__(self, ham: HamiltonianOperations, W, dt=0.005, order=6, vbias_bound=50.0,
                 seed=0, stream=1, target_pop=0.0, free_projection=False, trial=None,
                 local_energy=False):
        self.ham = ham
        self.ctx = ham.ctx
        self.dt = dt
        self.W = int(W)
        self.trial = trial
        p = PropParams(dt, order, vbias_bound, 1.0, 10.0, 100.0, 0.1, target_pop,
                       int(free_projection), 0, 0, seed, stream, int(bool(local_energy)))
        h = _P()
        if trial is None:
            _ck(lib().afq_prop_create(ham.h, W, ctypes.byref(p), ctypes.byref(h)))
        else:
            if trial.ham is not ham:
                raise AfqError("the trial was built over another hamiltonian")
            create = (lib().afq_prop_create_nomsd if getattr(trial, "kind", "") == "nomsd"
                      else lib().afq_prop_create_phmsd)
            _ck(create(trial.h, W, ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        v, P = _P(), _P()
        _ck(lib().afq_prop_views(h, ctypes.byref(v), ctypes.byref(P)))
        self._vmf, self._p1 = v.value, P.value

    def vMF(self):
        torch = _torch()
        out = torch.empty(self.ham.M, dtype=torch.complex128, device="cuda")
        self.ctx.synchronize()
        _copy_d2d(out.data_ptr(), self._vmf, out.numel() * 16)
        return out

    def P1(self):
        """P1 [N][N], or [2][N][N] (alpha, beta) for spin-dependent integrals."""
        torch = _torch()
        np1 = 2 if max(self.ham.nspin_L, self.ham.nspin_h1) > 1 else 1
        shape = (self.ham.N, self.ham.N) if np1 == 1 else (2, self.ham.N, self.ham.N)
        out = torch.empty(shape, dtype=torch.complex128, device="cuda")
        self.ctx.synchronize()
        _copy_d2d(out.data_ptr(), self._p1, out.numel() * 16)
        return out

    def Propagate(self, wset: WalkerSet, Eshift, u=None):
        _ck(lib().afq_prop_step(self.h, ctypes.byref(wset.cstruct()), ptr(u), ptr(Eshift)))

    def Orthogonalize(self, wset: WalkerSet):
        _ck(lib().afq_prop_orthogonalize(self.h, ctypes.byref(wset.cstruct())))

    def Energy(self, wset: WalkerSet, E, ovlp, precision=AFQ_FP64):
        _ck(lib().afq_prop_energy(self.h, ctypes.byref(wset.cstruct()), precision, ptr(E), ptr(ovlp)))

    def set_step(self, step):
        _ck(lib().afq_prop_set_step(self.h, step))

    def step_graph(self, wset: WalkerSet, Eshift, ortho=False, relax=0.0, energy_out=None,
                   popctl=False):
        _ck(lib().afq_prop_step_graph(self.h, ctypes.byref(wset.cstruct()), ptr(Eshift), int(ortho),
                                      int(popctl), relax, ptr(energy_out)))

    def population_control(self, wset: WalkerSet, r):
        _ck(lib().afq_prop_population_control(self.h, ctypes.byref(wset.cstruct()), float(r)))

    def bound_stats(self, reset=False):
        """Bound-trigger counters: {'vbias'|'eloc'|'weight': (total, upper, lower)}."""
        c = (ctypes.c_int64 * 9)()
        _ck(lib().afq_prop_bound_stats(self.h, c, int(bool(reset))))
        return {"vbias": tuple(c[0:3]), "eloc": tuple(c[3:6]), "weight": tuple(c[6:9])}

    def collapsed(self) -> bool:
        v = _I()
        _ck(lib().afq_prop_status(self.h, ctypes.byref(v)))
        return bool(v.value)

    # ---- back-propagation (BackPropEstimator, single-determinant trial) ----
    def enable_backprop(self, nback: int):
        """Every later step stores its fields and path-restoration factor (rings of nback and
        3 nback slots per walker)."""
        _ck(lib().afq_prop_enable_backprop(self.h, int(nback)))
        self.nback = int(nback)

    def bp_anchor(self, wset: WalkerSet):
        _ck(lib().afq_prop_bp_anchor(self.h, ctypes.byref(wset.cstruct())))

    def backpropagate(self, wset: WalkerSet, nbp: int, ortho: int = 10, restore: int = 1):
        """(G [W][nspin][N][N], weights [W]) of the trial back-propagated over the last nbp
        steps against the anchors; the weights carry the path-restoration factors of the last
        restore * nbp steps (0 none, 1 path_restoration, 2 extra_path_restoration)."""
        torch = _torch()
        nspin = 2 if self.ham.walker_type == AFQ_COLLINEAR else 1
        N = self.ham.N
        G = torch.empty((wset.W, nspin, N, N), dtype=torch.complex128, device="cuda")
        w = torch.empty(wset.W, dtype=torch.complex128, device="cuda")
        _ck(lib().afq_prop_backpropagate(self.h, ctypes.byref(wset.cstruct()), int(nbp), int(ortho),
                                         int(restore), ptr(G), ptr(w)))
        return G, w

    def _rdm_out(self, onerdm, diag2rdm):
        torch = _torch()
        nsp = 2 if self.ham.walker_type == AFQ_COLLINEAR else 1
        N = self.ham.N
        one = torch.empty((nsp, N, N), dtype=torch.complex128, device="cuda") if onerdm else None
        two = (torch.empty((nsp + 1, N, N), dtype=torch.complex128, device="cuda") if diag2rdm
               else None)
        return one, two

    def bp_measure(self, wset: WalkerSet, nbp: int, ortho: int = 10, restore: int = 1,
                   onerdm=True, diag2rdm=True):
        """BackPropEstimator's OneRDM [nspin][N][N] and unfolded DiagonalTwoRDM [nspin+1][N][N]
        of the trial's references back-propagated over the last nbp steps (None where not
        asked for)."""
        one, two = self._rdm_out(onerdm, diag2rdm)
        _ck(lib().afq_prop_bp_measure(self.h, ctypes.byref(wset.cstruct()), int(nbp), int(ortho),
                                      int(restore), ptr(one), ptr(two)))
        return one, two

    def mixed_measure(self, wset: WalkerSet, onerdm=True, diag2rdm=True):
        """MixedEstimator's OneRDM and unfolded DiagonalTwoRDM of the current walkers."""
        one, two = self._rdm_out(onerdm, diag2rdm)
        _ck(lib().afq_prop_mixed_measure(self.h, ctypes.byref(wset.cstruct()), ptr(one), ptr(two)))
        return one, two

    def number_of_references(self) -> int:
        v = _I()
        _ck(lib().afq_prop_number_of_references(self.h, ctypes.byref(v)))
        return v.value

    def vhs_timing(self, on: bool):
        """Record HIP events around the vHS launch of every later step (inside the step graph)."""
        _ck(lib().afq_prop_vhs_timing(self.h, int(bool(on))))

    def vhs_ms(self) -> float:
        """Duration of the last step's vHS launch in ms."""
        v = ctypes.c_float()
        _ck(lib().afq_prop_vhs_ms(self.h, ctypes.byref(v)))
        return float(v.value)

    # ---- back-propagation by time-evolved operators (TimeEvolvedBPEstimator) ----
    def te_anchor(self, follow_walkers: bool = False):
        """X = Y = I and M = 0 for every walker slot (setAnchor); follow_walkers makes
        population control copy the operators with their walkers (the reference leaves them
        with the slots)."""
        _ck(lib().afq_prop_te_anchor(self.h, int(bool(follow_walkers))))

    def te_propagate(self, wset: WalkerSet, nsteps: int):
        """PropagateOperators over the nsteps most recent stored fields, then the
        regularization with the current walker matrices."""
        _ck(lib().afq_prop_te_propagate(self.h, ctypes.byref(wset.cstruct()), int(nsteps)))

    def te_measure(self, wset: WalkerSet, restore_steps: int, transpose_m: bool = False,
                   onerdm=True, diag2rdm=True):
        """OneRDM [nspin][N][N] and unfolded DiagonalTwoRDM [nspin+1][N][N] of the Green
        functions dressed by the time-evolved operators, with WEIGHT times the factors of the
        last restore_steps steps.  transpose_m=False is the reference's G_d = M + ...; True
        adds M^T, which equals the back-propagated estimator."""
        one, two = self._rdm_out(onerdm, diag2rdm)
        _ck(lib().afq_prop_te_measure(self.h, ctypes.byref(wset.cstruct()), int(restore_steps),
                                      int(bool(transpose_m)), ptr(one), ptr(two)))
        return one, two

    def te_operators(self):
        """Host copies (X, Y, M) [W][nspin][N][N] and P1inv [N][N]."""
        torch = _torch()
        x, y, m, q = _P(), _P(), _P(), _P()
        _ck(lib().afq_prop_te_views(self.h, ctypes.byref(x), ctypes.byref(y), ctypes.byref(m),
                                    ctypes.byref(q)))
        self.ctx.synchronize()
        N = self.ham.N
        nsp = 2 if self.ham.walker_type == AFQ_COLLINEAR else 1
        out = []
        for v in (x, y, m):
            if not v.value:
                out.append(None)
                continue
            t = torch.empty((self.W, nsp, N, N), dtype=torch.complex128, device="cuda")
            _copy_d2d(t.data_ptr(), v.value, t.numel() * 16)
            out.append(host(t))
        p1 = torch.empty((N, N), dtype=torch.complex128, device="cuda")
        _copy_d2d(p1.data_ptr(), q.value, p1.numel() * 16)
        return out[0], out[1], out[2], host(p1)

    def bp_rings(self, W: int):
        """Copies of the field ring [W][nback][M] and the factor ring [W][3 nback]."""
        torch = _torch()
        f, q, n = _P(), _P(), _I()
        _ck(lib().afq_prop_bp_views(self.h, ctypes.byref(f), ctypes.byref(q), ctypes.byref(n)))
        self.ctx.synchronize()
        nb, M = n.value, self.ham.M
        fields = torch.empty((W, nb, M), dtype=torch.complex128, device="cuda")
        facs = torch.empty((W, 3 * nb), dtype=torch.complex128, device="cuda")
        _copy_d2d(fields.data_ptr(), f.value, fields.numel() * 16)
        _copy_d2d(facs.data_ptr(), q.value, facs.numel() * 16)
        return host(fields), host(facs)

    def close(self):
        if getattr(self, "h", None):
            lib().afq_prop_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
