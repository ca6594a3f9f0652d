"""Error bars of correlated Monte Carlo series (the estimators' time series).

The reference's result tools report an observable as the mean of its series with the standard
error of bin means (python/safiretools/results/stats.py); this module provides the same
quantities from the textbook methods, in its own formulation:

- statistical_inefficiency: kappa = 1 + 2 sum_{t>=1} rho(t), the factor by which correlation
  inflates the variance of the mean.  The autocovariance comes from an FFT of the centred
  series (float64, zero padded to a power of two, so no wrap-around), and the sum is truncated
  with Geyer's initial monotone sequence estimator (C. J. Geyer, Stat. Sci. 7, 473 (1992),
  section 3.3): with Gamma_k = gamma(2k) + gamma(2k+1), keep Gamma_k while it is positive and
  replace each by the smallest of those before it; kappa = 2 sum_k Gamma_k / gamma(0) - 1,
  at least 1.  A series whose spread is at rounding level has kappa = 1.
- optimal_rebinsize: bins of 12 integrated autocorrelation times, tau_int = (kappa - 1) / 2 of
  the slowest component (every column, real and imaginary parts separately), so that the
  variance of a bin mean is within a few percent of its asymptotic value; at least
  MIN_BINS bins are kept, and an uncorrelated series (kappa ~ 1) is not rebinned.
- rebinning_analysis: means of consecutive bins (the tail past the last whole bin dropped)
  and the autocorrelation time implied by them, tau = (B var(bins) / var(samples) - 1) / 2,
  the largest over the components (H. Flyvbjerg, H. G. Petersen, J. Chem. Phys. 91, 461
  (1989)).
- standard_error: sample standard deviation of the bins / sqrt(number of bins); the real and
  imaginary parts of complex bins separately.
- jackknife: value and error of a function of the means of one or more binned series from the
  leave-one-bin-out estimates (B. Efron, The Jackknife, the Bootstrap and Other Resampling
  Plans, SIAM 1982).
"""
from __future__ import annotations

import numpy as np

MIN_BINS = 8          # fewest bins an error estimate is made from when the series allows it
BIN_TAUS = 12.0       # bin length in integrated autocorrelation times
_COLUMN_BLOCK = 1 << 22   # float64 elements of one FFT batch


def _flat_columns(x: np.ndarray) -> np.ndarray:
    """[samples][components] float64 view of a series; complex parts become separate columns."""
    c = x.reshape(x.shape[0], -1)
    if np.iscomplexobj(c):
        c = np.concatenate([c.real, c.imag], axis=1)
    return np.asarray(c, np.float64)


def _flat(series: np.ndarray) -> bool:
    """True for the columns whose spread is rounding noise of their magnitude."""
    spread = np.ptp(series, axis=0)
    scale = np.maximum(np.max(np.abs(series), axis=0), 1.0)
    return spread <= 1e-12 * scale


def _kappa_block(cols: np.ndarray) -> np.ndarray:
    """Statistical inefficiency of each column of a [n][k] float64 block (n >= 2)."""
    n, k = cols.shape
    centred = cols - cols.mean(axis=0)
    size = 1 << int(np.ceil(np.log2(2 * n)))
    spec = np.fft.rfft(centred, n=size, axis=0)
    gamma = np.fft.irfft(spec.real ** 2 + spec.imag ** 2, n=size, axis=0)[:n] / n
    npair = n // 2
    Gam = gamma[0:2 * npair:2] + gamma[1:2 * npair:2]            # [npair][k]
    positive = np.cumprod(Gam > 0.0, axis=0).astype(bool)       # the initial positive run
    mono = np.minimum.accumulate(np.where(positive, Gam, np.inf), axis=0)
    total = np.where(positive, mono, 0.0).sum(axis=0)
    g0 = gamma[0]
    out = np.ones(k)
    ok = g0 > 0.0
    out[ok] = 2.0 * total[ok] / g0[ok] - 1.0
    return np.maximum(out, 1.0)


def statistical_inefficiency(samples) -> float | np.ndarray:
    """kappa of a real series, or of every column of a [samples][columns] array."""
    x = np.asarray(samples, np.float64)
    cols = x.reshape(x.shape[0], -1) if x.ndim > 1 else x[:, None]
    n, k = cols.shape
    kappa = np.ones(k)
    if n >= 2:
        live = np.flatnonzero(~_flat(cols))
        step = max(1, _COLUMN_BLOCK // (2 * n))
        for start in range(0, live.size, step):
            idx = live[start:start + step]
            kappa[idx] = _kappa_block(cols[:, idx])
    return float(kappa[0]) if x.ndim == 1 else kappa


def optimal_rebinsize(samples, skip: int = 0) -> int:
    """Bin size for rebinning_analysis chosen from the series itself (1 when uncorrelated)."""
    x = np.asarray(samples)[skip:]
    n = x.shape[0]
    if n < 2 * MIN_BINS:
        return 1
    kappa = float(np.max(np.atleast_1d(statistical_inefficiency(_flat_columns(x)))))
    tau_int = 0.5 * (kappa - 1.0)
    B = int(np.ceil(BIN_TAUS * tau_int))
    return int(min(max(B, 1), n // MIN_BINS))


def rebinning_analysis(samples, *, rebinsize: int, skip: int = 0):
    """(bin means [m][...], autocorrelation time) of a series whose leading axis runs over
    samples; the samples past the last whole bin are dropped."""
    x = np.asarray(samples)[skip:]
    B = max(int(rebinsize), 1)
    m = x.shape[0] // B
    if m < 2:
        raise ValueError(f"{x.shape[0]} samples in bins of {B} make {m} bin(s); an error "
                         "estimate needs at least 2")
    used = x[: m * B]
    bins = used.reshape((m, B) + x.shape[1:]).mean(axis=1)
    v_samples = np.var(_flat_columns(used), axis=0, ddof=1)
    v_bins = np.var(_flat_columns(bins), axis=0, ddof=1)
    live = v_samples > 0.0
    tau = 0.0
    if live.any():
        tau = float(np.max(0.5 * (B * v_bins[live] / v_samples[live] - 1.0)))
    return bins, max(tau, 0.0)


def standard_error(bins):
    """Standard error of the mean over the leading axis; complex bins give err(re) + i err(im)."""
    b = np.asarray(bins)
    root_m = np.sqrt(b.shape[0])
    if np.iscomplexobj(b):
        return (np.std(b.real, axis=0, ddof=1) + 1j * np.std(b.imag, axis=0, ddof=1)) / root_m
    return np.std(b, axis=0, ddof=1) / root_m


def average(samples, skip: int = 0, rebinsize: int | None = None):
    """(mean, standard error) of an observable's series after rebinning (optimal_rebinsize
    unless a bin size is given)."""
    B = optimal_rebinsize(samples, skip) if rebinsize is None else rebinsize
    bins, _ = rebinning_analysis(samples, rebinsize=B, skip=skip)
    return bins.mean(axis=0), standard_error(bins)


def jackknife(func, *binned):
    """Jackknife value and error of func(mean_1, mean_2, ...) over binned series that share
    their number of bins m: theta_i = func of the means without bin i, value = m theta_all -
    (m - 1) mean_i theta_i, error = sqrt((m - 1) / m sum_i (theta_i - mean theta)^2)."""
    arrays = [np.asarray(b) for b in binned]
    m = arrays[0].shape[0]
    if any(a.shape[0] != m for a in arrays):
        raise ValueError("the binned series must have the same number of bins")
    if m < 2:
        raise ValueError("a jackknife needs at least 2 bins")
    means = [a.mean(axis=0) for a in arrays]
    theta_all = np.asarray(func(*means))
    # mean without bin i: (m mean - b_i) / (m - 1)
    theta = np.stack([np.asarray(func(*[(m * mu - a[i]) / (m - 1) for mu, a in zip(means, arrays)]))
                      for i in range(m)])
    theta_bar = theta.mean(axis=0)
    value = m * theta_all - (m - 1) * theta_bar
    scale = np.sqrt((m - 1) / m)
    if np.iscomplexobj(theta):
        err = scale * (np.sqrt(((theta.real - theta_bar.real) ** 2).sum(axis=0)) +
                       1j * np.sqrt(((theta.imag - theta_bar.imag) ** 2).sum(axis=0)))
    else:
        err = scale * np.sqrt(((theta - theta_bar) ** 2).sum(axis=0))
    return value, err
