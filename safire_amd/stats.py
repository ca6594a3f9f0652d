"""Error analysis of the estimators' time series, as the reference's result tools do it
(python/safiretools/results/stats.py:23-95, 126-180, 241-280).

- statistical_inefficiency: kappa = 1 + 2 sum_t rho(t) from the normalized autocorrelation
  function, summed with Geyer's initial positive sequence (Geyer, Stat. Sci. 7, 473 (1992)):
  rho(2j) + rho(2j + 1) pairs are added up to the first pair that is not positive; kappa is
  floored at 1 and a series constant to roundoff gets 1.  The autocorrelation comes from a
  zero-padded real FFT of the deviations from the mean (the mean taken in double precision,
  the transform in single precision, as the reference does).
- optimal_rebinsize: the bin size B minimizing the squared relative bias of the variance of
  the mean, ((kappa^2 - 1) / (2 kappa B))^2 for an exponentially decaying autocorrelation,
  plus the relative variance 2 / (m - 1) of an error estimate from m bins; kappa is the one of
  the slowest component, chosen on one half of the series and estimated on the other, both
  ways round and averaged.
- rebinning_analysis: bin means of consecutive samples and the autocorrelation time
  tau = (kappa - 1) / 2 with kappa = B (std of the bins / std of the samples)^2, the largest
  over the components.
- standard_error: std(ddof=1) / sqrt(m) of the bins, real and imaginary parts separately.
- jackknife: value and error of a function of the means of binned series.
"""
from __future__ import annotations

import numpy as np

_ROUNDOFF = 1e-9            # a spread up to this is roundoff: nothing to correlate
_FFT_CHUNK = 1 << 24        # complex elements of one batch of transforms


def _fft_len(n: int) -> int:
    from scipy.fft import next_fast_len
    return next_fast_len(n, real=True)


def _geyer(columns: np.ndarray, nfft: int) -> np.ndarray:
    """kappa of every (non-constant) column of a [samples][columns] array."""
    from scipy.fft import irfft, rfft
    n, ncol = columns.shape
    dev = np.empty((ncol, n), np.float32)
    np.subtract(columns.T, columns.mean(axis=0)[:, None], out=dev, casting="same_kind")
    power = np.abs(rfft(dev, n=nfft, axis=1)) ** 2
    acov = irfft(power, n=nfft, axis=1)[:, :n] / n      # divided by n: positive semidefinite
    rho = acov / acov[:, :1]
    pairs = rho[:, : n // 2 * 2].reshape(ncol, -1, 2).sum(axis=2)
    bad = pairs <= 0.0
    npos = np.where(bad.any(axis=1), bad.argmax(axis=1), pairs.shape[1])
    keep = np.arange(pairs.shape[1])[None, :] < npos[:, None]
    s = np.where(keep, pairs, 0.0).sum(axis=1)
    # the pairs start at t = 0 and so count the leading 1 twice
    return np.maximum(2.0 * s.astype(np.float64) - 1.0, 1.0)


def statistical_inefficiency(samples) -> float | np.ndarray:
    """kappa of a real series, or of every column of a [samples][columns] array."""
    samples = np.asarray(samples, np.float64)
    cols = samples[:, None] if samples.ndim == 1 else samples
    n = cols.shape[0]
    kappa = np.ones(cols.shape[1])
    if n >= 2:
        moving = np.flatnonzero(np.ptp(cols, axis=0) > _ROUNDOFF)
        nfft = _fft_len(2 * n - 1)
        step = max(1, _FFT_CHUNK // nfft)
        for i in range(0, moving.size, step):
            sel = moving[i:i + step]
            kappa[sel] = _geyer(np.take(cols, sel, axis=1), nfft)
    return float(kappa[0]) if samples.ndim == 1 else kappa


def optimal_rebinsize(samples, skip: int = 0) -> int:
    """Bin size for rebinning_analysis chosen from the series itself (1 for an uncorrelated
    series)."""
    x = np.asarray(samples)[skip:]
    n = x.shape[0]
    if n < 2:
        return 1
    comp = x.reshape(n, -1)
    if np.iscomplexobj(comp):
        comp = np.concatenate([comp.real, comp.imag], axis=1)
    h0, h1 = (np.atleast_1d(statistical_inefficiency(h)) for h in np.array_split(comp, 2))
    kappa = 0.5 * float(h1[np.argmax(h0)] + h0[np.argmax(h1)])
    sizes = np.arange(1, n // 2 + 1)
    counts = n // sizes
    cost = ((kappa ** 2 - 1.0) / (2.0 * kappa * sizes)) ** 2 + 2.0 / (counts - 1)
    return int(sizes[np.argmin(cost)])


def rebinning_analysis(samples, *, rebinsize: int, skip: int = 0):
    """(bin means [m][...], autocorrelation time) of a series whose leading axis runs over
    samples; the samples past the last whole bin are dropped."""
    x = np.asarray(samples)[skip:]
    B = int(rebinsize)
    m = x.shape[0] // B
    if m < 2:
        raise ValueError(f"{x.shape[0]} samples in bins of {B} leave {m} bins, too few for an "
                         "error estimate")
    x = x[: m * B]
    naive = np.std(x, axis=0, ddof=1)
    bins = x.reshape(m, B, *x.shape[1:]).mean(axis=1)
    binstd = np.std(bins, axis=0, ddof=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        kappa = B * (binstd / naive) ** 2
    kappa = np.where(naive > 0.0, kappa, 1.0)
    return bins, max(0.5 * (float(np.max(kappa)) - 1.0), 0.0)


def standard_error(bins):
    """Standard error of the mean over the leading axis; complex bins get re + i im errors."""
    bins = np.asarray(bins)
    if np.iscomplexobj(bins):
        s = bins.real.std(axis=0, ddof=1) + 1j * bins.imag.std(axis=0, ddof=1)
    else:
        s = bins.std(axis=0, ddof=1)
    return s / np.sqrt(bins.shape[0])


def average(samples, skip: int = 0, rebinsize: int | None = None):
    """(mean, standard error) of an observable's series after rebinning (the optimal bin size
    unless one is given), the analysis behind the reference's Results.average."""
    if rebinsize is None:
        rebinsize = optimal_rebinsize(samples, skip)
    bins, _ = rebinning_analysis(samples, rebinsize=rebinsize, skip=skip)
    return bins.mean(axis=0), standard_error(bins)


def jackknife(func, *binned):
    """Jackknife value and error of func over the means of one or more binned series with the
    same number of bins (leave-one-bin-out pseudovalues)."""
    counts = {np.asarray(b).shape[0] for b in binned}
    if len(counts) != 1:
        raise ValueError("the binned series must agree in their number of bins")
    m = counts.pop()
    if m < 2:
        raise ValueError("a jackknife needs at least 2 bins")
    binned = [np.asarray(b) for b in binned]
    tot = [b.sum(axis=0) for b in binned]
    loo = np.array([func(*((t - b[i]) / (m - 1) for t, b in zip(tot, binned)))
                    for i in range(m)])
    full = func(*(b.mean(axis=0) for b in binned))
    pseudo = m * full - (m - 1) * loo
    return pseudo.mean(axis=0), standard_error(pseudo)
