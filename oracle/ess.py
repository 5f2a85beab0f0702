"""Rank-normalised split bulk effective sample size — test infrastructure only.

The reference has no ESS code (only an autocorrelation helper,
hmclab/Helpers/Processing.py:6-8).  This restates the published estimator
of Vehtari, Gelman, Simpson, Carpenter & Buerkner (2021), "Rank-normalization,
folding, and localization: an improved R-hat", Bayesian Analysis 16(2):
split chains in half, rank-normalise the pooled draws with the Blom offset
(r - 3/8)/(S + 1/4), estimate the multi-chain autocorrelation from FFT
autocovariances, truncate with Geyer's initial positive sequence and enforce
the initial monotone sequence.  One parameter at a time, plain loops.
"""
from __future__ import annotations

import numpy as np
from scipy.special import ndtri
from scipy.stats import rankdata


def _autocov(x):
    """Biased autocovariance of a 1-D array via zero-padded FFT."""
    n = x.size
    x = x - x.mean()
    m = 1
    while m < 2 * n:
        m *= 2
    f = np.fft.rfft(x, m)
    ac = np.fft.irfft(f * np.conj(f), m)[:n]
    return ac / n


def ess_raw(chains):
    """ESS of a (M, N) array of draws (no splitting, no ranking)."""
    chains = np.asarray(chains, dtype=np.float64)
    M, N = chains.shape
    if N < 4:
        return float("nan")
    acov = np.stack([_autocov(c) for c in chains])
    chain_mean = chains.mean(axis=1)
    mean_var = acov[:, 0].mean() * N / (N - 1.0)
    var_plus = mean_var * (N - 1.0) / N
    if M > 1:
        var_plus += chain_mean.var(ddof=1)
    if not var_plus > 0:
        return float("nan")
    rho = np.zeros(N)
    rho_even = 1.0
    rho[0] = rho_even
    rho_odd = 1.0 - (mean_var - acov[:, 1].mean()) / var_plus
    rho[1] = rho_odd
    t = 1
    while t < N - 3 and (rho_even + rho_odd) > 0.0:
        rho_even = 1.0 - (mean_var - acov[:, t + 1].mean()) / var_plus
        rho_odd = 1.0 - (mean_var - acov[:, t + 2].mean()) / var_plus
        if rho_even + rho_odd >= 0.0:
            rho[t + 1] = rho_even
            rho[t + 2] = rho_odd
        t += 2
    max_t = t - 2
    if rho_even > 0.0:
        rho[max_t + 1] = rho_even
    t = 1
    while t <= max_t - 2:
        if rho[t + 1] + rho[t + 2] > rho[t - 1] + rho[t]:
            rho[t + 1] = (rho[t - 1] + rho[t]) / 2.0
            rho[t + 2] = rho[t + 1]
        t += 2
    ess = float(M * N)
    tau = -1.0 + 2.0 * rho[: max_t + 1].sum() + rho[max_t + 1]
    tau = max(tau, 1.0 / np.log10(ess))
    return ess / tau


def split_chains(chains):
    chains = np.asarray(chains)
    M, N = chains.shape
    h = N // 2
    return np.concatenate([chains[:, :h], chains[:, N - h:]], axis=0)


def rank_normalize(chains):
    shape = chains.shape
    r = rankdata(np.asarray(chains).reshape(-1), method="average")
    S = r.size
    return ndtri((r - 0.375) / (S + 0.25)).reshape(shape)


def ess_bulk(chains):
    """Bulk ESS of one parameter, (M chains, N draws)."""
    return ess_raw(rank_normalize(split_chains(chains)))


def min_ess_bulk(draws):
    """Min over parameters; ``draws`` is (M chains, N draws, P params)."""
    draws = np.asarray(draws)
    return min(ess_bulk(draws[:, :, k]) for k in range(draws.shape[2]))
