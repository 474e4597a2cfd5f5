"""Convergence diagnostics used to report ESS/s at R-hat < 1.01.

Restates the reference's R diagnostics so the numbers are comparable:
  ess      R/diagnostics.R:36-99   per-chain Geyer initial positive sequence
                                   on the sample ACF (lag <= 2000), summed
  rhat     R/diagnostics.R:127-159 split R-hat (call with warmup=0 on kept
                                   draws, as R/diagnose.R:68-69 does)
  mcse     R/diagnose.R:70         sd / sqrt(ESS)
"""
from __future__ import annotations

import os

import numpy as np
import scipy.fft as _fft

# FFT worker threads: the cores this process may use, capped by
# OMP_NUM_THREADS where the host sets it.
_WORKERS = len(os.sched_getaffinity(0))
if os.environ.get("OMP_NUM_THREADS", "").isdigit():
    _WORKERS = max(1, min(_WORKERS, int(os.environ["OMP_NUM_THREADS"])))


def _acf(x: np.ndarray, lag_max: int) -> np.ndarray:
    """Sample autocorrelations 1..lag_max of each column (stats::acf with
    demean=TRUE: autocovariances divided by n, normalised by lag 0).  The
    transform is zero-padded to at least n + lag_max points, so the circular
    products at lags 0..lag_max are the linear ones."""
    n = x.shape[0]
    xc = x - x.mean(axis=0, keepdims=True)
    nfft = _fft.next_fast_len(n + lag_max + 1, real=True)
    f = _fft.rfft(xc, n=nfft, axis=0, workers=_WORKERS)
    pw = f.real * f.real + f.imag * f.imag
    acov = _fft.irfft(pw, n=nfft, axis=0, workers=_WORKERS)[: lag_max + 1] / n
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


def _rhat_params(d: np.ndarray) -> np.ndarray:
    """rhat() of every parameter of an (n_iter, n_chains, n_params) array."""
    n = d.shape[0]
    half = n // 2
    if half < 2:
        raise ValueError("need at least 4 post-warmup iterations per chain "
                         "for split R-hat.")
    splits = np.concatenate([d[:half], d[n - half:]], axis=1)
    sub_means = splits.mean(axis=0)                  # (2C, P)
    sub_vars = splits.var(axis=0, ddof=1)
    between = half * sub_means.var(axis=0, ddof=1)
    within = sub_vars.mean(axis=0)
    var_hat = ((half - 1) * within + between) / half
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.sqrt(var_hat / within)
    return np.where(within > 0, r, np.nan)


def summarize(draws: np.ndarray) -> dict:
    """Per-parameter mean, sd, R-hat, ESS and MCSE (R/diagnose.R:64-84) of an
    (n_iter, n_chains, n_params) array; every parameter's chains go through
    one autocorrelation transform."""
    d = np.asarray(draws, dtype=np.float64)
    if d.ndim == 2:
        d = d[:, :, None]
    n, C, P = d.shape
    e = ess_chains(d.reshape(n, C * P)).reshape(C, P).sum(axis=0)
    r = _rhat_params(d)
    flat = d.reshape(n * C, P)
    mean = flat.mean(axis=0)
    sd = flat.std(axis=0, ddof=1)
    rows = [{"mean": float(mean[p]), "sd": float(sd[p]), "rhat": float(r[p]),
             "ess": float(e[p]),
             "mcse": float(sd[p] / np.sqrt(max(e[p], 1.0)))} for p in range(P)]
    return {"params": rows,
            "max_rhat": max(x["rhat"] for x in rows),
            "min_ess": min(x["ess"] for x in rows)}


def ess_ar(x: np.ndarray) -> float:
    """Effective sample size of one chain from its spectral density at
    frequency zero, estimated by an autoregression: the estimator of
    coda::effectiveSize, which the reference's M1 benchmark harness sums
    over chains (benchmark/harness_util.R:9-12).

    The series is centred; autocovariances (divisor n) up to lag
    p_max = min(n - 1, floor(10 log10 n)) give the Yule-Walker fits of
    every order by the Levinson-Durbin recursion; the order p minimises
    n log(v_p) + 2 p (Akaike); the innovation variance is rescaled by
    n / (n - p - 1), and S(0) = v / (1 - sum(phi))^2.  ESS = n var(x) / S(0)
    (var with divisor n - 1), 0 for a constant chain."""
    x = np.asarray(x, dtype=np.float64).ravel()
    n = x.size
    if n < 2:
        return float(n)
    xc = x - x.mean()
    r0 = float(np.dot(xc, xc)) / n
    if r0 == 0.0:
        return 0.0
    pmax = int(min(n - 1, np.floor(10.0 * np.log10(n))))
    r = np.array([np.dot(xc[:n - k], xc[k:]) / n for k in range(pmax + 1)])
    phi = np.zeros(0)
    v = r0
    fits, var = [phi], [v]
    for k in range(1, pmax + 1):
        kappa = (r[k] - np.dot(phi, r[k - 1:0:-1])) / v
        phi = np.concatenate([phi - kappa * phi[::-1], [kappa]])
        v = v * (1.0 - kappa * kappa)
        if not v > 0.0:
            break
        fits.append(phi)
        var.append(v)
    var = np.asarray(var)
    aic = n * np.log(var) + 2.0 * np.arange(var.size)
    p = int(np.argmin(aic))
    v_pred = var[p] * n / (n - (p + 1))
    spec = v_pred / (1.0 - float(np.sum(fits[p]))) ** 2
    return 0.0 if spec == 0.0 else n * float(np.var(x, ddof=1)) / spec
