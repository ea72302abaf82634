"""Sample diagnostics for ACHR runs: moments, KS, bulk ESS and split-R-hat.

The reference reports no diagnostics (it writes chain end points only); these
follow the standard definitions used for MCMC output:
* split-R-hat and bulk ESS on rank-normalised draws (Vehtari, Gelman,
  Simpson, Carpenter, Buerkner, "Rank-normalization, folding, and
  localization: an improved R-hat", Bayesian Analysis 2021), with the
  autocorrelation sum truncated by Geyer's initial monotone sequence;
* the two-sample Kolmogorov-Smirnov statistic sup |F1 - F2|.
All functions take traces shaped (chains, draws, reactions) and work on every
reaction at once.
"""
from __future__ import annotations

import numpy as np


def _norm_ppf(p: np.ndarray) -> np.ndarray:
    from scipy.special import ndtri
    return ndtri(p)


def rank_normalize(x: np.ndarray) -> np.ndarray:
    """(chains, draws, k) -> normal scores of the pooled fractional ranks ((r - 3/8) / (S + 1/4))."""
    from scipy.stats import rankdata
    c, d, k = x.shape
    flat = x.reshape(c * d, k)
    r = rankdata(flat, axis=0, method="average")
    z = _norm_ppf((r - 0.375) / (c * d + 0.25))
    return z.reshape(c, d, k)


def split_chains(x: np.ndarray) -> np.ndarray:
    """(chains, draws, k) -> (2*chains, draws//2, k): first and second halves as separate chains."""
    c, d, k = x.shape
    h = d // 2
    return np.concatenate([x[:, :h], x[:, d - h:]], axis=0)


def autocov(x: np.ndarray) -> np.ndarray:
    """Autocovariance along axis 1 of (chains, draws, k), biased (divide by draws), via FFT."""
    c, d, k = x.shape
    xc = x - x.mean(axis=1, keepdims=True)
    nfft = 1 << int(np.ceil(np.log2(2 * d)))
    f = np.fft.rfft(xc, n=nfft, axis=1)
    ac = np.fft.irfft(f * np.conj(f), n=nfft, axis=1)[:, :d]
    return ac / d


def ess(x: np.ndarray) -> np.ndarray:
    """Effective sample size per reaction of (chains, draws, k), Geyer initial monotone sequence."""
    c, d, k = x.shape
    if d < 4:
        return np.full(k, np.nan)
    acov = autocov(x)
    chain_mean = x.mean(axis=1)
    chain_var = acov[:, 0, :] * d / (d - 1.0)
    mean_var = chain_var.mean(axis=0)
    var_plus = mean_var * (d - 1.0) / d
    if c > 1:
        var_plus = var_plus + chain_mean.var(axis=0, ddof=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        rho = 1.0 - (mean_var[None, :] - acov.mean(axis=0)) / var_plus[None, :]  # (d, k)
    rho[0] = 1.0
    out = np.empty(k)
    for j in range(k):
        if not np.isfinite(var_plus[j]) or var_plus[j] <= 0:
            out[j] = np.nan
            continue
        r = rho[:, j]
        # pair sums P_t = rho_{2t} + rho_{2t+1}, keep while positive, make monotone
        npairs = d // 2
        P = r[0:2 * npairs:2] + r[1:2 * npairs:2]
        tau = -1.0
        prev = np.inf
        for t in range(npairs):
            p = P[t]
            if p <= 0:
                break
            p = min(p, prev)
            prev = p
            tau += 2.0 * p
        tau = max(tau, 1.0 / np.log10(c * d) if c * d > 1 else 1.0)
        out[j] = c * d / tau
    return out


def bulk_ess(x: np.ndarray) -> np.ndarray:
    """Bulk ESS: ESS of the rank-normalised split chains."""
    return ess(rank_normalize(split_chains(x)))


def rhat(x: np.ndarray) -> np.ndarray:
    """Split-R-hat on rank-normalised draws (per reaction)."""
    z = rank_normalize(split_chains(x))
    c, d, k = z.shape
    means = z.mean(axis=1)
    B = d * means.var(axis=0, ddof=1)
    Wv = z.var(axis=1, ddof=1).mean(axis=0)
    var_plus = (d - 1.0) / d * Wv + B / d
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.sqrt(var_plus / Wv)


def ks_2samp(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """sup |F_a - F_b| per column of (draws, k) samples."""
    a = np.sort(np.asarray(a), axis=0)
    b = np.sort(np.asarray(b), axis=0)
    out = np.empty(a.shape[1])
    for j in range(a.shape[1]):
        allv = np.concatenate([a[:, j], b[:, j]])
        fa = np.searchsorted(a[:, j], allv, side="right") / a.shape[0]
        fb = np.searchsorted(b[:, j], allv, side="right") / b.shape[0]
        out[j] = np.max(np.abs(fa - fb))
    return out


def moment_error(samples: np.ndarray, ref_mean: np.ndarray, ref_std: np.ndarray, lb: np.ndarray,
                 ub: np.ndarray):
    """Max over reactions of |mean - ref| and |std - ref| as fractions of the range ub - lb.

    samples: (k, draws). Fixed reactions (zero range) are skipped.
    """
    rng = ub - lb
    live = rng > 0
    mu = samples.mean(axis=1)
    sd = samples.std(axis=1, ddof=1)
    em = np.abs(mu - ref_mean)[live] / rng[live]
    es = np.abs(sd - ref_std)[live] / rng[live]
    return float(em.max(initial=0.0)), float(es.max(initial=0.0))
