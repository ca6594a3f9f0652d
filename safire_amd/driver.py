"""The AFQMC driver loop on the device and the energy estimator.

run_afqmc mirrors run_afqmc.cpp:69-100 for one GPU: every step is one captured HIP graph
(Propagate, Orthogonalize every `ortho_interval` steps, comb population control every
`popctl_interval` steps, the growth estimator, and the Eshift relaxation during equilibration).
After equilibration the energy estimator (EnergyEstimator.h:51-90) evaluates the local energies
every `measure_interval` steps and records the weighted average sum_w w E_w / sum_w w
(MeasurementOutput, Measurements.hpp:92-121) together with the weight statistics.

blocking_error is the Flyvbjerg-Petersen blocking analysis of a correlated series (Flyvbjerg &
Petersen, J. Chem. Phys. 91, 461 (1989)), the method behind safiretools' rebinning analysis.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import afq


@dataclasses.dataclass
class Trace:
    steps: list
    energy: list          # complex weighted-average local energy per measurement
    components: list      # (E1, EXX, EJ) weighted averages
    total_weight: list
    eff_walkers: list
    growth: float = float("nan")   # growth-estimator energy after the last step
    eshift: float = float("nan")   # Eshift after the last step
    onerdm_mixed: list = dataclasses.field(default_factory=list)   # [nspin][N][N] per measurement
    onerdm_bp: list = dataclasses.field(default_factory=list)      # [nspin][N][N] per window

    def mean_error(self, skip=0):
        e = np.real(np.asarray(self.energy[skip:]))
        return float(e.mean()), blocking_error(e)


def blocking_error(x) -> float:
    """Standard error of the mean of a correlated series: the plateau of the blocked estimates
    (the largest estimate over blocking levels with at least 8 blocks)."""
    x = np.asarray(x, dtype=np.float64)
    best = 0.0
    while x.size >= 8:
        n = x.size
        err = np.sqrt(np.var(x, ddof=1) / n) if n > 1 else 0.0
        best = max(best, err)
        m = n // 2
        x = 0.5 * (x[0:2 * m:2] + x[1:2 * m:2])
    return float(best)


def initialize(prop: "afq.Propagator", ws: "afq.WalkerSet", precision=afq.AFQ_FP64):
    """execute.cpp:170-183: Energy of the initial walkers -> OVLP and PSEUDO_ELOC_, Eshift =
    weighted average energy (averageEnergy, average_energy.hpp:68-78).  Returns Eshift (device)."""
    import torch
    W = ws.W
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    prop.Energy(ws, E, ov, precision)
    et = E.sum(1)
    ws.eloc.copy_(et)
    ws.ovlp.copy_(ov)
    w = ws.weight
    esh = ((w * et).sum() / w.sum()).real.reshape(1).to(torch.float64).contiguous()
    return esh


def mixed_onerdm(prop, ws, psit_dev, nup, ndown):
    """MixedEstimator's OneRDM (MixedEstimator.hpp:122-140, onerdm.hpp:40-60) of one
    measurement: sum_w w G_w / sum_w w with G_w the full mixed density matrix of each spin."""
    import torch
    W, N = ws.W, prop.ham.N
    coll = ws.Phi_b is not None
    G = torch.empty((W, 2 if coll else 1, N, N), dtype=torch.complex128, device="cuda")
    ov = torch.zeros(W, dtype=torch.complex128, device="cuda")
    Gs = afq.empty((W, N, N), torch.complex128)
    afq.det_ops.MixedDensityMatrix(prop.ctx, psit_dev[:nup], ws.Phi, Gs, ov, compact=False)
    G[:, 0] = Gs
    if coll:
        afq.det_ops.MixedDensityMatrix(prop.ctx, psit_dev[nup:nup + ndown], ws.Phi_b, Gs, ov,
                                       compact=False)
        G[:, 1] = Gs
    w = ws.weight
    return (w[:, None, None, None] * G).sum(0) / w.sum()


def run_afqmc(prop: "afq.Propagator", ws: "afq.WalkerSet", steps: int, equilibration: int = 100,
              ortho_interval: int = 5, popctl_interval: int = 10, measure_interval: int = 10,
              relax: float | None = None, energy_precision=afq.AFQ_FP64, Eshift=None,
              progress=None, progress_interval: int = 1000, onerdm_interval: int = 0,
              psit=None, backprop: dict | None = None) -> Trace:
    """progress(step, trace), when given, is called every progress_interval steps.

    onerdm_interval > 0 (with psit, the trial rows [nup (+ ndown)][N]) records the mixed
    estimator's OneRDM every onerdm_interval steps after equilibration.  backprop =
    {"length": L, "ortho": k, "restore": True} records BackPropEstimator's OneRDM
    (BackPropEstimator.hpp:150-260): the anchor is taken after the last equilibration step (or
    before the first step without equilibration); L steps later the trial is back-propagated
    over them and the weighted average of the back-propagated density matrices recorded, and the
    anchor moves to that step."""
    import torch
    W = ws.W
    if relax is None:  # parameter_defaults.cpp:310-318
        relax = 1.0 - np.exp(-10.0 / equilibration) if equilibration > 1 else 1.0
    if Eshift is None:
        Eshift = initialize(prop, ws)
    e_out = afq.empty(1, torch.float64)
    E = afq.empty((W, 3), torch.complex128)
    ov = afq.empty(W, torch.complex128)
    tr = Trace([], [], [], [], [])
    psit_dev = None
    if onerdm_interval > 0:
        if psit is None:
            raise ValueError("the mixed OneRDM needs the trial rows (psit)")
        psit_dev = afq.device_tensor(np.asarray(psit), np.complex128)
    nup = ws.Phi.shape[2]
    ndown = 0 if ws.Phi_b is None else ws.Phi_b.shape[2]
    bp_len = 0
    bp_pos = -1
    if backprop:
        bp_len = int(backprop["length"])
        if getattr(prop, "nback", 0) < bp_len:
            prop.enable_backprop(bp_len)
        prop.bp_anchor(ws)   # the anchor before the first step
    for step in range(steps):
        prop.step_graph(ws, Eshift, ortho=(step % ortho_interval == 0),
                        popctl=(popctl_interval > 0 and step % popctl_interval == 0),
                        relax=(relax if step < equilibration else 0.0), energy_out=e_out)
        if step >= equilibration and step % measure_interval == 0:
            prop.Energy(ws, E, ov, energy_precision)
            w = ws.weight
            sw = w.sum()
            comps = (w[:, None] * E).sum(0) / sw
            aw = w.abs()
            neff = (aw.sum() ** 2 / (aw ** 2).sum())
            vals = afq.host(torch.cat([comps, sw.reshape(1), neff.reshape(1).to(torch.complex128)]))
            tr.steps.append(step)
            tr.components.append(vals[:3])
            tr.energy.append(complex(vals[:3].sum()))
            tr.total_weight.append(complex(vals[3]))
            tr.eff_walkers.append(float(vals[4].real))
        if psit_dev is not None and step >= equilibration and step % onerdm_interval == 0:
            tr.onerdm_mixed.append(afq.host(mixed_onerdm(prop, ws, psit_dev, nup, ndown)))
        if bp_len:
            if step == equilibration - 1:
                prop.bp_anchor(ws)
                bp_pos = step
            elif step >= equilibration and step - bp_pos >= bp_len:
                G, wr = prop.backpropagate(ws, bp_len, backprop.get("ortho", ortho_interval),
                                           backprop.get("restore", True))
                tr.onerdm_bp.append(afq.host((wr[:, None, None, None] * G).sum(0) / wr.sum()))
                prop.bp_anchor(ws)
                bp_pos = step
        if progress is not None and (step + 1) % progress_interval == 0:
            progress(step + 1, tr)
    torch.cuda.synchronize()
    tr.growth = float(afq.host(e_out)[0])
    tr.eshift = float(afq.host(Eshift)[0])
    return tr
