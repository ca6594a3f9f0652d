"""The AFQMC driver loop on the device and the energy estimator.

run_afqmc mirrors run_afqmc.cpp:69-100 for one GPU: every step is one captured HIP graph
(Propagate, Orthogonalize every `ortho_interval` steps, comb population control every
`popctl_interval` steps, the growth estimator, and the Eshift relaxation during equilibration).
After equilibration the energy estimator (EnergyEstimator.h:51-90) evaluates the local energies
every `measure_interval` steps and records the weighted average sum_w w E_w / sum_w w
(MeasurementOutput, Measurements.hpp:92-121) together with the weight statistics.

Error bars (Trace.mean_error, mean_error) are the standard error of bin means, the quantity the
reference's result tools report: the bins span twelve integrated autocorrelation times of the
series (safire_amd.stats).  blocking_error, the Flyvbjerg-Petersen blocking analysis
(J. Chem. Phys. 91, 461 (1989)), is kept as an independent estimate.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import afq, stats


@dataclasses.dataclass
class Trace:
    steps: list
    energy: list          # complex weighted-average local energy per measurement
    components: list      # (E1, EXX, EJ) weighted averages
    total_weight: list
    eff_walkers: list
    overlap: list = dataclasses.field(default_factory=list)  # sum_w w e^{log ovlp_w} / sum_w w
    phase: list = dataclasses.field(default_factory=list)    # sum_w w / sum_w |w|
    growth: float = float("nan")   # growth-estimator energy after the last step
    eshift: float = float("nan")   # Eshift after the last step
    # MixedEstimator: observable name -> one entry per measurement
    mixed: dict = dataclasses.field(default_factory=dict)
    # BackPropEstimator: propagation steps -> observable name -> one entry per window
    backprop: dict = dataclasses.field(default_factory=dict)
    # TimeEvolvedBPEstimator: the same layout
    time_evolved_bp: dict = dataclasses.field(default_factory=dict)

    def mean_error(self, skip=0):
        """(mean, error) of the real energy after `skip` measurements, rebinned."""
        m, e = stats.average(np.real(np.asarray(self.energy)), skip=skip)
        return float(m), float(e)


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


def _observables(one, two, walker_type) -> dict:
    """{"OneRDM": [nspin][N][N], "DiagonalTwoRDM": the reference's folded vector} (host
    copies) of one measurement."""
    out = {}
    if one is not None:
        out["OneRDM"] = afq.host(one)
    if two is not None:
        out["DiagonalTwoRDM"] = afq.fold_diag2rdm(afq.host(two), walker_type)
    return out


def _record(store: dict, obs: dict):
    for k, v in obs.items():
        store.setdefault(k, []).append(v)


def mean_error(series, skip=0) -> tuple:
    """(mean, error) of a series of complex observable measurements, element by element, as
    the reference's Results.average gives them: one bin size for the whole observable (the
    one its slowest component needs), the mean of the bins and their standard error, whose
    complex form is error(re) + i error(im)."""
    return stats.average(np.asarray(series, np.complex128), skip=skip)


def run_afqmc(prop: "afq.Propagator", ws: "afq.WalkerSet", steps: int, equilibration: int = 100,
              ortho_interval: int = 5, popctl_interval: int = 10, measure_interval: int = 10,
              relax: float | None = None, energy_precision=afq.AFQ_FP64, Eshift=None,
              progress=None, progress_interval: int = 1000,
              estimators: dict | None = None) -> Trace:
    """progress(step, trace), when given, is called every progress_interval steps.

    `estimators` takes the reference's input block (parameters.hpp estimator parameters):
      {"mixed": {"measure_interval": k, "onerdm": {}, "diag_twordm": {}},
       "backprop": {"propagation_steps": [L1, L2, ...], "path_restoration": True,
                    "extra_path_restoration": False, "walker_ortho_interval": k,
                    "onerdm": {}, "diag_twordm": {}},
       "time_evolved_bp": {"propagation_steps": [L1, L2, ...], "path_restoration": True,
                           "extra_path_restoration": False, "transpose_m": False,
                           "follow_walkers": False, "onerdm": {}, "diag_twordm": {}}}
    The mixed estimator (MixedEstimator.hpp:122-140) measures its observables on the current
    walkers every measure_interval steps after equilibration.  Both estimators sum over the
    trial's reference determinants (one for NOMSD, the expansion for PHMSD).  Back-propagation (BackPropEstimator.hpp:150-260) anchors after
    the last equilibration step (before the first step without equilibration); when the steps
    since the anchor equal one of the propagation lengths the trial is back-propagated over
    them and the observables measured with the (path-restored) weights, and at the longest
    length the anchor moves to the current step.  The time-evolved estimator
    (TimeEvolvedBPEstimator.hpp:199-262) keeps the same anchors and windows, but evolves the
    operators forward over each segment between consecutive lengths and dresses the current
    walkers' density matrices with them; its regularization assumes orthonormal walkers, so it
    is exact when the walkers are orthogonalized every step (ortho_interval = 1).  Both
    back-propagation estimators share one field ring."""
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
    est = estimators or {}
    mixed = est.get("mixed")
    bp = est.get("backprop")
    wt = prop.ham.walker_type
    if mixed is not None:
        mixed_interval = int(mixed.get("measure_interval", measure_interval))
    te = est.get("time_evolved_bp")

    def lengths(block):
        lens = sorted(int(x) for x in block["propagation_steps"])
        if not lens or len(set(lens)) != len(lens) or lens[0] <= 0:
            raise ValueError("propagation_steps must be distinct and positive")
        return lens

    def restoration(block):
        return ((2 if block.get("extra_path_restoration", False) else 1)
                if block.get("path_restoration", True) else 0)

    bp_lens, te_lens = [], []
    bp_pos = te_pos = -1
    if bp is not None:
        bp_lens = lengths(bp)
        restore = restoration(bp)
        bp_ortho = int(bp.get("walker_ortho_interval", ortho_interval))
        for L in bp_lens:
            tr.backprop[L] = {}
    if te is not None:
        te_lens = lengths(te)
        te_restore = restoration(te)
        te_transpose = bool(te.get("transpose_m", False))
        te_follow = bool(te.get("follow_walkers", False))
        for L in te_lens:
            tr.time_evolved_bp[L] = {}
    nback = max(bp_lens[-1:] + te_lens[-1:] + [0])
    if nback and getattr(prop, "nback", 0) < nback:
        prop.enable_backprop(nback)
    # the anchors before the first step
    if bp_lens:
        prop.bp_anchor(ws)
    if te_lens:
        prop.te_anchor(te_follow)
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
            # EnergyEstimator.h:86-87: the weighted average of the walkers' overlaps
            ovl = (w * torch.exp(ov)).sum() / sw
            # weight_statistics (Measurements.hpp:130-146): phase = sum w / sum |w|
            phase = sw / aw.sum()
            vals = afq.host(torch.cat([comps, sw.reshape(1), neff.reshape(1).to(torch.complex128),
                                       ovl.reshape(1), phase.reshape(1)]))
            tr.steps.append(step)
            tr.components.append(vals[:3])
            tr.energy.append(complex(vals[:3].sum()))
            tr.total_weight.append(complex(vals[3]))
            tr.eff_walkers.append(float(vals[4].real))
            tr.overlap.append(complex(vals[5]))
            tr.phase.append(complex(vals[6]))
        if mixed is not None and step >= equilibration and step % mixed_interval == 0:
            one, two = prop.mixed_measure(ws, "onerdm" in mixed, "diag_twordm" in mixed)
            _record(tr.mixed, _observables(one, two, wt))
        if bp_lens:
            if step == equilibration - 1:
                prop.bp_anchor(ws)
                bp_pos = step
            elif step >= equilibration:
                n = step - bp_pos
                if n in tr.backprop:
                    one, two = prop.bp_measure(ws, n, bp_ortho, restore, "onerdm" in bp,
                                               "diag_twordm" in bp)
                    _record(tr.backprop[n], _observables(one, two, wt))
                if n >= bp_lens[-1]:
                    prop.bp_anchor(ws)
                    bp_pos = step
        if te_lens:
            if step == equilibration - 1:
                prop.te_anchor(te_follow)
                te_pos = step
            elif step >= equilibration:
                n = step - te_pos
                if n in tr.time_evolved_bp:
                    # the operators already cover the previous (shorter) length
                    i = te_lens.index(n)
                    prop.te_propagate(ws, n - (te_lens[i - 1] if i > 0 else 0))
                    one, two = prop.te_measure(ws, te_restore * n, te_transpose,
                                               "onerdm" in te, "diag_twordm" in te)
                    _record(tr.time_evolved_bp[n], _observables(one, two, wt))
                if n >= te_lens[-1]:
                    prop.te_anchor(te_follow)
                    te_pos = step
        if progress is not None and (step + 1) % progress_interval == 0:
            progress(step + 1, tr)
    torch.cuda.synchronize()
    tr.growth = float(afq.host(e_out)[0])
    tr.eshift = float(afq.host(Eshift)[0])
    return tr
