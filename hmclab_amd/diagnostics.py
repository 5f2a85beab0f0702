"""Rank-normalised split bulk ESS, vectorised over parameters.

The reference has no ESS (only hmclab/Helpers/Processing.py:6-8).  This
implements the estimator of Vehtari et al. (2021, Bayesian Analysis 16(2)):
split chains in half, rank-normalise pooled draws with (r - 3/8)/(S + 1/4)
(average ranks for ties, which rejections produce), FFT autocovariances,
Geyer's initial positive sequence with the initial monotone correction.
``min_ess_bulk`` accepts NumPy arrays or torch tensors (the heavy ranking and
FFT steps then run where the tensor lives); the truncation loop runs on the
host over (N, P) autocorrelations.
"""
from __future__ import annotations

import numpy as np
from scipy.special import ndtri


def _rho_to_ess(rho: np.ndarray, M: int, N: int) -> np.ndarray:
    """Geyer truncation per parameter. rho: (N, P) autocorrelations, rho[0]=1."""
    P = rho.shape[1]
    ess = np.empty(P)
    for p in range(P):
        r = rho[:, p]
        out = np.zeros(N)
        rho_even, rho_odd = 1.0, r[1]
        out[0], out[1] = rho_even, rho_odd
        t = 1
        while t < N - 3 and (rho_even + rho_odd) > 0.0:
            rho_even, rho_odd = r[t + 1], r[t + 2]
            if rho_even + rho_odd >= 0.0:
                out[t + 1], out[t + 2] = rho_even, rho_odd
            t += 2
        max_t = t - 2
        if rho_even > 0.0:
            out[max_t + 1] = rho_even
        t = 1
        while t <= max_t - 2:
            if out[t + 1] + out[t + 2] > out[t - 1] + out[t]:
                out[t + 1] = (out[t - 1] + out[t]) / 2.0
                out[t + 2] = out[t + 1]
            t += 2
        total = float(M * N)
        tau = -1.0 + 2.0 * out[: max_t + 1].sum() + out[max_t + 1]
        tau = max(tau, 1.0 / np.log10(total))
        ess[p] = total / tau
    return ess


def _autocorr_numpy(z: np.ndarray):
    """z: (M, N, P) -> rho (N, P)."""
    M, N, P = z.shape
    zc = z - z.mean(axis=1, keepdims=True)
    m = 1
    while m < 2 * N:
        m *= 2
    f = np.fft.rfft(zc, n=m, axis=1)
    acov = np.fft.irfft(f * np.conj(f), n=m, axis=1)[:, :N, :] / N
    chain_mean = z.mean(axis=1)
    mean_var = acov[:, 0, :].mean(axis=0) * N / (N - 1.0)
    var_plus = mean_var * (N - 1.0) / N
    if M > 1:
        var_plus = var_plus + chain_mean.var(axis=0, ddof=1)
    rho = 1.0 - (mean_var[None, :] - acov.mean(axis=0)) / var_plus[None, :]
    rho[0] = 1.0
    return rho


def _rank_normalize_numpy(x: np.ndarray) -> np.ndarray:
    """x: (S, P) -> normal scores with average ranks for ties."""
    S, P = x.shape
    srt = np.sort(x, axis=0)
    out = np.empty_like(x, dtype=np.float64)
    for p in range(P):
        lo = np.searchsorted(srt[:, p], x[:, p], side="left")
        hi = np.searchsorted(srt[:, p], x[:, p], side="right")
        r = 0.5 * (lo + 1 + hi)
        out[:, p] = ndtri((r - 0.375) / (S + 0.25))
    return out


def _split(draws):
    M, N = draws.shape[0], draws.shape[1]
    h = N // 2
    if hasattr(draws, "narrow"):   # torch
        import torch

        return torch.cat([draws[:, :h], draws[:, N - h:]], dim=0)
    return np.concatenate([draws[:, :h], draws[:, N - h:]], axis=0)


def ess_bulk(draws) -> np.ndarray:
    """Bulk ESS per parameter of draws (M chains, N draws, P params); N >= 4
    (split halves need at least two draws each)."""
    if draws.shape[1] < 4:
        raise ValueError("bulk ESS needs at least 4 draws per chain")
    if hasattr(draws, "narrow"):
        return _ess_bulk_torch(draws)
    x = np.asarray(draws, dtype=np.float64)
    if x.ndim == 2:
        x = x[:, :, None]
    s = _split(x)
    M, N, P = s.shape
    z = _rank_normalize_numpy(s.reshape(M * N, P)).reshape(M, N, P)
    rho = _autocorr_numpy(z)
    return _rho_to_ess(rho, M, N)


def _ess_bulk_torch(draws) -> np.ndarray:
    import torch

    x = draws.to(torch.float64)
    if x.dim() == 2:
        x = x[:, :, None]
    s = _split(x)
    M, N, P = s.shape
    flat = s.reshape(M * N, P).transpose(0, 1).contiguous()   # (P, S)
    srt, _ = torch.sort(flat, dim=1)
    lo = torch.searchsorted(srt, flat, right=False).to(torch.float64)
    hi = torch.searchsorted(srt, flat, right=True).to(torch.float64)
    r = 0.5 * (lo + 1.0 + hi)
    q = (r - 0.375) / (M * N + 0.25)
    z = torch.special.ndtri(q).transpose(0, 1).reshape(M, N, P)
    zc = z - z.mean(dim=1, keepdim=True)
    m = 1
    while m < 2 * N:
        m *= 2
    f = torch.fft.rfft(zc, n=m, dim=1)
    acov = torch.fft.irfft(f * torch.conj(f), n=m, dim=1)[:, :N, :] / N
    chain_mean = z.mean(dim=1)
    mean_var = acov[:, 0, :].mean(dim=0) * N / (N - 1.0)
    var_plus = mean_var * (N - 1.0) / N
    if M > 1:
        var_plus = var_plus + chain_mean.var(dim=0, unbiased=True)
    rho = 1.0 - (mean_var[None, :] - acov.mean(dim=0)) / var_plus[None, :]
    rho[0] = 1.0
    return _rho_to_ess(rho.cpu().numpy(), M, N)


def min_ess_bulk(draws) -> float:
    return float(np.min(ess_bulk(draws)))
