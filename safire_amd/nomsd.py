"""Multi-determinant NOMSD trial (NOMSD.hpp/.cpp with ndet > 1): a non-orthogonal expansion
sum_d c_d |D_d> over determinants of the hamiltonian's basis.

Determinant 0 is the trial the hamiltonian was built with (its half-rotated tensors are
shared); each further determinant is half-rotated on the device when the trial is created.  The
walkers' stored log overlap (WalkerSet OVLP) is the scale of the determinant weights, as in
NOMSD::MixedDensityMatrix (NOMSD.cpp:219-278).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import afq


class NOMSD:
    """NOMSD<DEVICE_MEMORY> with ndet >= 1 determinants.  PsiT [ndet][nel][N] complex (device
    tensor, the conjugate-transposed orbitals of each determinant, alpha rows first), ci [ndet]
    (the expansion coefficients).  Determinant 0 must be the hamiltonian's own PsiT."""

    kind = "nomsd"

    def __init__(self, ham, PsiT, ci):
        torch = afq._torch()
        self.ham = ham
        self.ctx = ham.ctx
        ci = np.ascontiguousarray(np.asarray(ci, dtype=np.complex128).ravel())
        self.ndet = len(ci)
        nel = ham.nup + (ham.ndown if ham.walker_type == afq.AFQ_COLLINEAR else 0)
        if tuple(PsiT.shape) != (self.ndet, nel, ham.N) or PsiT.dtype != torch.complex128:
            raise afq.AfqError(f"PsiT must be complex128 [{self.ndet}][{nel}][{ham.N}]")
        self.PsiT = PsiT.contiguous()
        self.ci = ci
        h = ctypes.c_void_p()
        afq._ck(afq.lib().afq_nomsd_create(ham.h, self.ndet, afq.ptr(self.PsiT),
                                            ci.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)))
        self.h = h

    def _full_shape(self, W):
        nspin = 2 if self.ham.walker_type == afq.AFQ_COLLINEAR else 1
        return (W, nspin * self.ham.N * self.ham.N)

    def MixedDensityMatrix(self, wset, G, Ov, log_m=None):
        """Full G [W][nspin N N] and Ov [W] (sum of the weights plus the scale); the scale is
        log_m, the walkers' stored log overlap when not given."""
        c = wset.cstruct()
        lm = afq.ptr(log_m) if log_m is not None else c.ovlp
        afq._ck(afq.lib().afq_nomsd_mixed_density_matrix(self.h, c.Phi, c.Phi_b, c.phi_stride,
                                                          wset.W, lm, afq.ptr(G), afq.ptr(Ov)))

    def Log_Overlap(self, wset, Ov):
        c = wset.cstruct()
        afq._ck(afq.lib().afq_nomsd_log_overlap(self.h, c.Phi, c.Phi_b, c.phi_stride, wset.W,
                                                 afq.ptr(Ov)))

    def Energy(self, wset, E, Ov, precision=None):
        c = wset.cstruct()
        prec = self.ham.precision if precision is None else precision
        afq._ck(afq.lib().afq_nomsd_energy(self.h, c.Phi, c.Phi_b, c.phi_stride, wset.W, prec,
                                            afq.ptr(E), afq.ptr(Ov)))

    def vbias(self, wset, v, dt, precision=None):
        """NOMSD::vbias (NOMSD.cpp:39-65): full mixed density matrix scaled by the walkers'
        stored log overlap, then the full-G contraction."""
        c = wset.cstruct()
        prec = self.ham.precision if precision is None else precision
        afq._ck(afq.lib().afq_nomsd_vbias(self.h, c.Phi, c.Phi_b, c.phi_stride, wset.W, c.ovlp,
                                           dt, prec, afq.ptr(v)))

    def vMF(self, v, dt):
        afq._ck(afq.lib().afq_nomsd_vmf(self.h, dt, afq.ptr(v)))

    def close(self):
        if getattr(self, "h", None):
            afq.lib().afq_nomsd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
