"""Convergence diagnostics used to report ESS/s at R-hat < 1.01.

Restates the reference's R diagnostics so the numbers are comparable:
  ess      R/diagnostics.R:36-99   per-chain Geyer initial positive sequence
                                   on the sample ACF (lag <= 2000), summed
  rhat     R/diagnostics.R:127-159 split R-hat (call with warmup=0 on kept
                                   draws, as R/diagnose.R:68-69 does)
  mcse     R/diagnose.R:70         sd / sqrt(ESS)
"""
from __future__ import annotations

import numpy as np


def _acf(x: np.ndarray, lag_max: int) -> np.ndarray:
    """Sample autocorrelations 1..lag_max of each column (stats::acf with
    demean=TRUE: autocovariances divided by n, normalised by lag 0)."""
    n = x.shape[0]
    xc = x - x.mean(axis=0, keepdims=True)
    nfft = 1
    while nfft < 2 * n:
        nfft *= 2
    f = np.fft.rfft(xc, n=nfft, axis=0)
    acov = np.fft.irfft(f * np.conj(f), n=nfft, axis=0)[: lag_max + 1] / n
    return acov[1:] / acov[0:1]


def ess_chains(draws: np.ndarray) -> np.ndarray:
    """Per-chain ESS of an (n_iter, n_chains) array (R/diagnostics.R:36-59)."""
    x = np.asarray(draws, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    n, m = x.shape
    out = np.full(m, float(n))
    if n < 8:
        return out
    var = x.var(axis=0, ddof=1)
    ok = var > 0
    if not ok.any():
        return out
    lag_max = min(n - 1, 2000)
    rho = _acf(x[:, ok], lag_max)
    npair = rho.shape[0] // 2
    if npair < 1:
        return out
    pair = rho[0:2 * npair:2] + rho[1:2 * npair:2]
    # Pairs are kept up to (not including) the first non-positive one.
    nonpos = pair <= 0
    last = np.where(nonpos.any(axis=0), nonpos.argmax(axis=0), npair)
    keep = np.arange(npair)[:, None] < last[None, :]
    tau = np.maximum(1.0 + 2.0 * np.where(keep, pair, 0.0).sum(axis=0), 1.0)
    out[ok] = np.where(last < 1, float(n), np.minimum(n / tau, n))
    return out


def ess(draws: np.ndarray) -> float:
    """Summed per-chain ESS (R/diagnostics.R:96-99) with warmup = 0."""
    return float(ess_chains(draws).sum())


def rhat(draws: np.ndarray) -> float:
    """Split R-hat of an (n_iter, n_chains) array, warmup = 0
    (R/diagnostics.R:127-159)."""
    x = np.asarray(draws, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    n = x.shape[0]
    half = n // 2
    if half < 2:
        raise ValueError("need at least 4 post-warmup iterations per chain "
                         "for split R-hat.")
    splits = np.concatenate([x[:half], x[n - half:]], axis=1)
    sub_means = splits.mean(axis=0)
    sub_vars = splits.var(axis=0, ddof=1)
    between = half * sub_means.var(ddof=1)
    within = sub_vars.mean()
    if within <= 0:
        return float("nan")
    var_hat = ((half - 1) * within + between) / half
    return float(np.sqrt(var_hat / within))


def summarize(draws: np.ndarray) -> dict:
    """Per-parameter mean, sd, R-hat, ESS and MCSE (R/diagnose.R:64-84) of an
    (n_iter, n_chains, n_params) array."""
    d = np.asarray(draws, dtype=np.float64)
    if d.ndim == 2:
        d = d[:, :, None]
    rows = []
    for p in range(d.shape[2]):
        M = d[:, :, p]
        v = M.ravel()
        e = ess(M)
        rows.append({"mean": float(v.mean()), "sd": float(v.std(ddof=1)),
                     "rhat": rhat(M), "ess": e,
                     "mcse": float(v.std(ddof=1) / np.sqrt(max(e, 1.0)))})
    return {"params": rows,
            "max_rhat": max(r["rhat"] for r in rows),
            "min_ess": min(r["ess"] for r in rows)}
