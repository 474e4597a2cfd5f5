"""Online surrogate of a model's summed log-likelihood for delayed acceptance.

The sampler's own warmup proposals, each evaluated exactly by the chain
kernel, are the training pairs (theta, loglik).  The network
(n_params-256-256-1, SiLU) trains on the GPU; its inputs are whitened with
the training proposals' mean and covariance, its target standardised.  A
held-out tenth of the pairs measures the error that delayed acceptance
pays for: the standard deviation of (prediction - loglik), a constant offset
being irrelevant to Metropolis ratios.  Above the bound the surrogate is not
used and the sampler runs plain Metropolis.

No counterpart in the reference: this is the learned first stage that the
stepwise sampler (gpum_hip_model_run_da) runs in front of the exact
bytecode evaluation.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _native

# Held-out error bound (log-likelihood units).  Stage two accepts a move
# that stage one let through with probability E min(1, exp(X)), X the
# difference of the surrogate's errors at the two points; for independent
# N(0, e^2) errors that is 1/2 + exp(e^2) Phi(-sqrt(2) e): 0.81 at e = 0.5,
# 0.71 at e = 1, 0.63 at e = 2.  Past e = 1 the exact evaluations that
# stage one saves are increasingly spent again on stage-two rejections.
DEFAULT_BOUND = 1.0


@dataclass
class SurrogateFit:
    net: Optional[_native.Surrogate]
    mean: np.ndarray           # input whitening: x = whiten (theta - mean)
    whiten: np.ndarray
    usable: bool
    holdout_err: float         # sd of (prediction - loglik) on held-out pairs
    holdout_bias: float
    train_err: float
    bound: float
    n_train: int
    n_holdout: int
    train_ms: float
    loss_trace: np.ndarray = field(default_factory=lambda: np.zeros(0))

    def summary(self) -> dict:
        return {"usable": self.usable, "holdout_err": self.holdout_err,
                "holdout_bias": self.holdout_bias, "train_err": self.train_err,
                "bound": self.bound, "n_train": self.n_train,
                "n_holdout": self.n_holdout, "train_ms": self.train_ms}


def whitening(theta: np.ndarray):
    """Mean and W with W cov W^T = I (inverse Cholesky factor)."""
    mean = theta.mean(axis=0)
    cov = np.atleast_2d(np.cov(theta, rowvar=False))
    d = cov.shape[0]
    jitter = 1e-10 * max(float(np.trace(cov)) / d, 1e-300)
    Lc = np.linalg.cholesky(cov + jitter * np.eye(d))
    return mean, np.linalg.inv(Lc)


def fit_surrogate(theta, loglik, bound: float = DEFAULT_BOUND,
                  holdout: float = 0.1, steps: int = 4000, batch: int = 1024,
                  lr: float = 2e-3, seed: int = 0,
                  max_pairs: int = 400_000) -> SurrogateFit:
    theta = np.asarray(theta, dtype=np.float64)
    loglik = np.asarray(loglik, dtype=np.float64).ravel()
    theta = theta.reshape(loglik.size, -1)
    ok = np.isfinite(loglik) & np.all(np.isfinite(theta), axis=1)
    theta, loglik = theta[ok], loglik[ok]
    n = loglik.size
    if n < 64:
        d = theta.shape[1] if theta.ndim == 2 else 1
        return SurrogateFit(None, np.zeros(d), np.eye(d), False, np.inf,
                            np.nan, np.inf, bound, 0, 0, 0.0)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)[:max_pairs]
    n_hold = max(1, int(round(holdout * perm.size)))
    hold, train = perm[:n_hold], perm[n_hold:]
    th_tr, ll_tr = theta[train], loglik[train]
    mean, W = whitening(th_tr)
    t_mean, t_sd = float(ll_tr.mean()), float(ll_tr.std())
    if not t_sd > 0:
        t_sd = 1.0
    net = _native.Surrogate(theta.shape[1], seed)
    net.set_scaling(mean, W, t_mean, t_sd)
    trace, ms = net.train(th_tr, ll_tr, steps=steps, batch=batch, seed=seed,
                          lr=lr)
    err_h = net.predict(theta[hold]) - loglik[hold]
    k = min(train.size, 20000)
    err_t = net.predict(th_tr[:k]) - ll_tr[:k]
    herr = float(np.std(err_h))
    fit = SurrogateFit(net, mean, W, bool(herr <= bound), herr,
                       float(np.mean(err_h)),
                       float(np.std(err_t)), bound, int(train.size),
                       int(n_hold), ms, trace)
    return fit


# Proposal scale factors tried by the pilot (multiples of the warmup's
# Metropolis-tuned proposal).  Delayed acceptance moves the optimum up: a
# proposal the surrogate rejects costs no exact evaluation, so longer steps
# buy more distance per exact evaluation until the chains grow sticky
# (Sherlock, Thiery and Golightly, "Efficiency of delayed-acceptance random
# walk Metropolis algorithms", Ann. Statist. 49(5), 2021).
PILOT_FACTORS = (1.0, 1.3, 1.6, 2.0)
# First stream iteration of the pilots: far from the sampling run's.
PILOT_ITER_OFFSET = 1 << 30
# Least overall acceptance of a factor the pilot may pick.  The jump
# distance per millisecond keeps rising as acceptance falls (a rejected
# proposal costs almost nothing), but below a few percent each chain moves
# so rarely that the run needs many more iterations before its split R-hat
# settles: logistic-16 over 4e6 rows with 1024 chains needed over 16000
# iterations at 3.2 % acceptance against 9000 at 7.8 %.
PILOT_MIN_ACCEPT = 0.05


def tune_scale(dm, net, init, whiten, seed, proposal_l=None, proposal_sd=None,
               factors=PILOT_FACTORS, iters=60, keep=0.95,
               min_accept=PILOT_MIN_ACCEPT):
    """Scale factor of the delayed-acceptance proposal: short pilot runs from
    `init` at each factor, scored by the expected squared jump distance
    (whitened) per millisecond of device time; among the factors whose
    acceptance is at least `min_accept` (the first factor if none is), the
    smallest within `keep` of the best wins (the larger ones only add
    stickiness)."""
    n_chains, d = np.asarray(init).reshape(-1, net.n_in).shape
    W = np.asarray(whiten, dtype=np.float64).reshape(d, d)
    rows = []
    for k, f in enumerate(factors):
        r = dm.run_da(net, init, np.ones(n_chains), iters, seed,
                      proposal_l=None if proposal_l is None else np.asarray(proposal_l) * f,
                      proposal_sd=None if proposal_sd is None else np.asarray(proposal_sd) * f,
                      iter_offset=PILOT_ITER_OFFSET + k * iters)
        x = r.draws.astype(np.float64)
        z = np.diff(x, axis=0) @ W.T
        esjd = float(np.mean(np.sum(z * z, axis=2)))
        ms = r.kernel_ms / iters
        rows.append({"factor": f, "esjd": esjd, "ms_per_iter": ms,
                     "esjd_per_ms": esjd / ms,
                     "stage1_rate": float(np.mean(r.stage1_rate)),
                     "accept": float(np.mean(r.accept_rate))})
    cand = [r for r in rows if r["accept"] >= min_accept] or rows[:1]
    best = max(r["esjd_per_ms"] for r in cand)
    pick = next(r for r in cand if r["esjd_per_ms"] >= keep * best)
    return pick["factor"], rows
