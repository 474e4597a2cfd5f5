"""Adaptation primitives and batch orchestrators of the sampler.

Restates the reference's host-side orchestration (R/adaptive.R), vectorised
over chains:
  WelfordState            .am_init_state / .am_welford_update(_batch) /
                          .am_covariance           R/adaptive.R:15-61
  target_accept           .am_target_accept        R/adaptive.R:69-72
  robbins_monro_scale     .am_robbins_monro_scale  R/adaptive.R:78-85
  regularised_cholesky    .am_cholesky             R/adaptive.R:548-558
  warmup_batches          batch plan of .am_orchestrate_warmup :112-118
  default_ladder          .pt_default_ladder       R/adaptive.R:261-265

The per-chain Welford update runs sequentially over a batch's draws, as the
reference does; the pooled accumulator merges each chain's batch with the
parallel-combination identity of Chan, Golub and LeVeque (equal in exact
arithmetic to folding the draws in one by one).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

_LOW_D = [0.44, 0.35, 0.32, 0.30, 0.28, 0.27, 0.26, 0.25]


def target_accept(d: int) -> float:
    """R/adaptive.R:69-72."""
    return _LOW_D[d - 1] if 1 <= d <= 8 else 0.234


def robbins_monro_scale(scale, accept_rate, batch_idx, d, target=None):
    """R/adaptive.R:78-85: exp(log s + b^(-2/3) (acc - target))."""
    if target is None:
        target = target_accept(d)
    g = float(batch_idx) ** (-2.0 / 3.0)
    return np.exp(np.log(scale) + g * (np.asarray(accept_rate) - target))


def regularised_cholesky(sigma: np.ndarray, eps: float = 1e-8) -> np.ndarray:
    """R/adaptive.R:548-558: lower L with L L' = Sigma + eps*mean(diag)*I;
    the diagonal fallback when the perturbed matrix is still indefinite."""
    d = sigma.shape[0]
    reg = eps * float(np.mean(np.diag(sigma)))
    if not math.isfinite(reg) or reg <= 0:
        reg = eps
    s = sigma + reg * np.eye(d)
    try:
        return np.linalg.cholesky(s)
    except np.linalg.LinAlgError:
        return np.diag(np.sqrt(np.maximum(np.diag(s), reg)))


@dataclass
class WelfordState:
    """Per-chain running mean and cross-product (R/adaptive.R:15-26)."""
    n: np.ndarray      # (C,)
    mean: np.ndarray   # (C, d)
    M2: np.ndarray     # (C, d, d)
    sigma: np.ndarray  # (C, d, d) initial Sigma (used while n < 2)

    @staticmethod
    def init(n_chains: int, d: int, sigma_diag: np.ndarray) -> "WelfordState":
        sd = np.broadcast_to(np.asarray(sigma_diag, float), (n_chains, d))
        sig = np.zeros((n_chains, d, d))
        idx = np.arange(d)
        sig[:, idx, idx] = sd
        return WelfordState(np.zeros(n_chains, np.int64), np.zeros((n_chains, d)),
                            np.zeros((n_chains, d, d)), sig)

    def update_batch(self, X: np.ndarray) -> None:
        """X: (T, C, d) draws; sequential Welford per chain
        (R/adaptive.R:31-55)."""
        for t in range(X.shape[0]):
            x = X[t]
            self.n = self.n + 1
            delta = x - self.mean
            self.mean = self.mean + delta / self.n[:, None]
            delta2 = x - self.mean
            self.M2 = self.M2 + delta[:, :, None] * delta2[:, None, :]

    def covariance(self) -> np.ndarray:
        """R/adaptive.R:59-61."""
        out = self.sigma.copy()
        ok = self.n >= 2
        out[ok] = self.M2[ok] / (self.n[ok] - 1)[:, None, None]
        return out


@dataclass
class PooledWelford:
    """One accumulator over all chains' draws (R/adaptive.R:140-144)."""
    n: int
    mean: np.ndarray
    M2: np.ndarray
    sigma: np.ndarray

    @staticmethod
    def init(d: int, sigma_diag) -> "PooledWelford":
        return PooledWelford(0, np.zeros(d), np.zeros((d, d)),
                             np.diag(np.broadcast_to(np.asarray(sigma_diag, float), (d,))))

    def update_batch(self, X: np.ndarray) -> None:
        """Fold in every chain's batch, chain by chain (R/adaptive.R:170-177),
        merging each chain's batch statistics exactly."""
        T, C, d = X.shape
        for c in range(C):
            xs = X[:, c, :]
            mb = xs.mean(axis=0)
            dev = xs - mb
            Mb = dev.T @ dev
            n_new = self.n + T
            delta = mb - self.mean
            self.mean = self.mean + delta * (T / n_new)
            self.M2 = self.M2 + Mb + np.outer(delta, delta) * (self.n * T / n_new)
            self.n = n_new

    def covariance(self) -> np.ndarray:
        return self.sigma.copy() if self.n < 2 else self.M2 / (self.n - 1)


def warmup_batches(warmup: int) -> List[int]:
    """R/adaptive.R:112-118: 10..40 batches of ~50, the last absorbs the
    remainder."""
    nb = min(40, max(10, warmup // 50))
    nb = min(nb, warmup)
    k = warmup // nb
    sizes = [k] * nb
    sizes[-1] += warmup - nb * k
    return sizes


def default_ladder(n_chains: int, t_max: float = 10.0) -> np.ndarray:
    """R/adaptive.R:261-265."""
    if n_chains <= 1:
        return np.ones(n_chains)
    beta = t_max ** (1.0 / (n_chains - 1))
    return beta ** np.arange(n_chains)


def split_batches(total: int, every: int) -> Tuple[List[int], int]:
    """R/adaptive.R:284-289 (PT) and :436-441 (DE)."""
    nb = max(1, total // every)
    sizes = [every] * nb
    rem = total - nb * every
    if rem > 0:
        sizes[-1] += rem
    return sizes, nb
