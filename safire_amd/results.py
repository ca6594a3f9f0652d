"""The results file of a run: the binned observables of the estimators in the reference's
layout (Estimators::write, Estimators.hpp:115-137; Measurements::write, Measurements.hpp:69-80;
Accumulator, Accumulator.hpp:34-143), and their analysis as the reference's result tools give
it (safiretools Results.average / Results.evaluate: rebinning with the optimal bin size, the
standard error of the bins, the jackknife for functions of several observables).

write_results appends the complete bins of binsize consecutive measurements of every observable
a driver Trace recorded; the partial bin at the end is returned, as the reference keeps its
unfinished bin for the next write.  Observable names follow the reference's estimators:
Energy, OnebodyEnergy, ExchangeEnergy, CoulombEnergy, Overlap, ProjectionTime,
EffectiveNumWalkers, Phase, TotalWeight (EnergyEstimator.h:80-98), MixedEstimator/<obs>,
BackPropEstimator/Steps=<L>/<obs> and TimeEvolvedBP/Steps=<L>/<obs> with <obs> OneRDM or
DiagonalTwoRDM.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import stats
from .h5inputs import _ck, lib


def _append(path, stage, name, bins):
    bins = np.asarray(bins)
    cplx = np.iscomplexobj(bins)
    data = np.ascontiguousarray(bins, np.complex128 if cplx else np.float64)
    dims = np.array(bins.shape[1:], dtype=np.int64)
    _ck(lib().afq_h5_results_append(str(path).encode(), stage.encode(), name.encode(),
                                    len(dims), dims.ctypes.data, int(cplx), bins.shape[0],
                                    data.ctypes.data))


def _series(trace, dt):
    """{name: list of measurements} of a driver Trace."""
    out = {}
    if trace.energy:
        comps = np.asarray(trace.components)
        out["Energy"] = list(trace.energy)
        out["OnebodyEnergy"] = list(comps[:, 0])
        out["ExchangeEnergy"] = list(comps[:, 1])
        out["CoulombEnergy"] = list(comps[:, 2])
        out["Overlap"] = list(trace.overlap)
        out["ProjectionTime"] = [(s + 1) * dt for s in trace.steps]
        out["EffectiveNumWalkers"] = list(trace.eff_walkers)
        if trace.phase:
            out["Phase"] = list(trace.phase)
        out["TotalWeight"] = list(trace.total_weight)
    for obs, v in trace.mixed.items():
        out[f"MixedEstimator/{obs}"] = v
    for prefix, store in (("BackPropEstimator", trace.backprop),
                          ("TimeEvolvedBP", trace.time_evolved_bp)):
        for L, group in store.items():
            for obs, v in group.items():
                out[f"{prefix}/Steps={L}/{obs}"] = v
    return out


def write_results(path, trace, dt: float, stage: int = 0, binsize: int = 1) -> dict:
    """Append the complete bins of every observable of `trace` to Measurements/Stage<stage>
    of `path`; returns {name: measurements left over in the unfinished bin}."""
    if binsize <= 0:
        raise ValueError("binsize has to be positive")
    left = {}
    for name, v in _series(trace, dt).items():
        x = np.asarray(v)
        nb = x.shape[0] // binsize
        if nb > 0:
            _append(path, f"Stage{stage}", name,
                    x[: nb * binsize].reshape(nb, binsize, *x.shape[1:]).mean(axis=1))
        left[name] = x[nb * binsize:]
    return left


class Results:
    """The observables of one stage of a results file."""

    def __init__(self, path, stage: int | str = 0):
        self.path = str(path)
        self.stage = stage if isinstance(stage, str) else f"Stage{stage}"

    def observable_names(self) -> list:
        n = ctypes.c_long()
        _ck(lib().afq_h5_results_names(self.path.encode(), self.stage.encode(), None, 0,
                                       ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _ck(lib().afq_h5_results_names(self.path.encode(), self.stage.encode(), buf, n.value,
                                       ctypes.byref(n)))
        txt = buf.value.decode()
        return txt.split("\n") if txt else []

    def bins(self, name: str) -> np.ndarray:
        nb, rank, cplx = ctypes.c_long(), ctypes.c_int(), ctypes.c_int()
        dims = np.zeros(8, np.int64)
        _ck(lib().afq_h5_results_shape(self.path.encode(), self.stage.encode(), name.encode(),
                                       ctypes.byref(nb), ctypes.byref(rank), dims.ctypes.data,
                                       ctypes.byref(cplx)))
        shape = (nb.value,) + tuple(int(d) for d in dims[: rank.value])
        out = np.empty(shape, np.complex128 if cplx.value else np.float64)
        _ck(lib().afq_h5_results_read(self.path.encode(), self.stage.encode(), name.encode(),
                                      out.ctypes.data))
        return out

    def average(self, name: str, skip: int = 0, rebinsize: int | None = None):
        """(mean, standard error) of an observable, its bins rebinned."""
        return stats.average(self.bins(name), skip=skip, rebinsize=rebinsize)

    def evaluate(self, func, *names, skip: int = 0, rebinsize: int | None = None):
        """Jackknife value and error of func over the means of several observables, rebinned
        with one bin size (the largest their series need)."""
        series = [self.bins(n)[skip:] for n in names]
        if rebinsize is None:
            rebinsize = max(stats.optimal_rebinsize(s) for s in series)
        binned = [stats.rebinning_analysis(s, rebinsize=rebinsize)[0] for s in series]
        return stats.jackknife(func, *binned)
