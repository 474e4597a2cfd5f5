"""Seeded synthetic workloads for the sampler.

Each workload is a model written in the DSL plus a deterministic data
generator (numpy PCG64 from a fixed seed).  Stored data are float32 where
the workload says FP32; the sampler receives them promoted to f64 exactly
as the reference's R layer hands data to its backends.

  gaussian_mean   y ~ N(3.4, sd 2), n = 20 000; loglik -((y - mu)^2) / 8,
                  flat prior; 8 chains x 3000 iterations; closed-form
                  posterior N(mean(y), 4 / n).
  gamma           y ~ Gamma(shape 2.5, rate 0.7), n = 1 000 000 (FP32);
                  params (log_a, log_b), N(0, 10) priors on the logs;
                  4096 chains x 5000 iterations.
  mixture         y from 0.5 N(-3, 1) + 0.5 N(3, 1), n = 200 000 (FP32);
                  params mu1, mu2; 16 temperatures x 256 replicas.
  logistic16      16-dim Bayesian logistic regression, n = 4 000 000 (FP32),
                  equicorrelated (0.3) N(0, 1) covariates, N(0, 5) priors.
  regressions     512 independent linear regressions, n_m log-uniform in
                  [1e3, 1e6] (FP32), x1..x8 ~ N(0, 1), beta ~ N(0, 1),
                  sigma = 0.5; params b0..b8, log_sigma, flat priors;
                  DE-MC with 64 chains per model.  Closed-form posterior:
                  b ~ t_{n-p}(beta_hat, s^2 (X'X)^-1), sigma^2 ~
                  Inv-Gamma((n-p)/2, (n-p) s^2 / 2), p = 9.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .dsl import Model, gpum_model


@dataclass
class Workload:
    name: str
    model: Model
    data: np.ndarray  # (n_obs, n_cols) float64 (FP32-rounded when the spec says FP32)
    n_chains: int
    n_iter: int
    meta: Dict = field(default_factory=dict)

    @property
    def n_obs(self) -> int:
        return self.data.shape[0]

    @property
    def n_cols(self) -> int:
        return self.data.shape[1]


GAUSSIAN_MEAN = "~ -((y - mu)^2) / 8"
GAMMA = ("~ exp(log_a) * log_b - lgamma(exp(log_a)) + (exp(log_a) - 1) * log(y)"
         " - exp(log_b) * y")
GAMMA_PRIOR = "~ -0.5 * (log_a / 10)^2 - 0.5 * (log_b / 10)^2"
MIXTURE = ("~ log(0.5 * exp(-((y - mu1)^2) / 2) + 0.5 * exp(-((y - mu2)^2) / 2))")


def _logistic_formula(d: int) -> str:
    eta = " + ".join(f"b{j} * x{j}" for j in range(1, d + 1))
    return f"~ y * ({eta}) - log(1 + exp({eta}))"


def _logistic_prior(d: int) -> str:
    return "~ " + " + ".join(f"-0.5 * (b{j} / 5)^2" for j in range(1, d + 1))


def gaussian_mean(n: int = 20000, seed: int = 11, n_chains: int = 8,
                  n_iter: int = 3000) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    y = rng.normal(3.4, 2.0, size=n)
    m = gpum_model(GAUSSIAN_MEAN, params=["mu"], data=["y"])
    return Workload("gaussian_mean", m, y.reshape(-1, 1), n_chains, n_iter,
                    meta={"post_mean": float(np.mean(y)),
                          "post_sd": float(np.sqrt(4.0 / n))})


def gamma(n: int = 1_000_000, seed: int = 12, n_chains: int = 4096,
          n_iter: int = 5000) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    y = rng.gamma(shape=2.5, scale=1.0 / 0.7, size=n).astype(np.float32)
    m = gpum_model(GAMMA, params=["log_a", "log_b"], data=["y"],
                   prior=GAMMA_PRIOR)
    return Workload("gamma", m, y.astype(np.float64).reshape(-1, 1), n_chains,
                    n_iter, meta={"true": [np.log(2.5), np.log(0.7)]})


def gamma_posterior(y, prior_sd: float = 10.0, grid: int = 201,
                    span: float = 9.0) -> Dict:
    """Posterior of (log_a, log_b) for y ~ Gamma(shape a, rate b) with
    N(0, prior_sd) priors on the logs, from the sufficient statistics
    n, sum log y, sum y.

    The mode comes from Newton's method on the closed-form gradient and
    Hessian (digamma, trigamma); the Laplace approximation is N(mode, -H^-1).
    The exact posterior moments are then integrated on a grid of
    `grid` x `grid` points spanning +-`span` Laplace sds in the whitened
    coordinates u (theta = mode + L u, L L' = -H^-1): the log density is
    smooth and negligible at the edges, so the uniform rule converges
    geometrically."""
    from scipy.special import digamma, gammaln, polygamma
    y = np.asarray(y, np.float64).ravel()
    n, s1, s2 = float(y.size), float(np.log(y).sum()), float(y.sum())
    r = 1.0 / (prior_sd * prior_sd)

    def logpost(la, lb):
        a, b = np.exp(la), np.exp(lb)
        return (n * a * lb - n * gammaln(a) + (a - 1.0) * s1 - b * s2
                - 0.5 * r * (la * la + lb * lb))

    def grad_hess(th):
        la, lb = th
        a, b = math.exp(la), math.exp(lb)
        ga = a * (n * lb - n * digamma(a) + s1)
        g = np.array([ga - r * la, n * a - b * s2 - r * lb])
        H = np.array([[ga - n * a * a * polygamma(1, a) - r, n * a],
                      [n * a, -b * s2 - r]])
        return g, H

    m, v = y.mean(), y.var()
    th = np.array([math.log(m * m / v), math.log(m / v)])
    for _ in range(100):
        g, H = grad_hess(th)
        step = np.linalg.solve(H, g)
        th = th - step
        if np.max(np.abs(step)) < 1e-14:
            break
    _, H = grad_hess(th)
    cov = np.linalg.inv(-H)
    L = np.linalg.cholesky(cov)
    u = np.linspace(-span, span, grid)
    U = np.stack(np.meshgrid(u, u, indexing="ij"), axis=-1).reshape(-1, 2)
    pts = th + U @ L.T
    lp = logpost(pts[:, 0], pts[:, 1]) - logpost(th[0], th[1])
    wgt = np.exp(lp - lp.max())
    wgt /= wgt.sum()
    mean = wgt @ pts
    dev = pts - mean
    pcov = (dev * wgt[:, None]).T @ dev
    return {"mode": th, "laplace_sd": np.sqrt(np.diag(cov)),
            "mean": mean, "sd": np.sqrt(np.diag(pcov)),
            "corr": float(pcov[0, 1] / math.sqrt(pcov[0, 0] * pcov[1, 1]))}


def mixture(n: int = 200_000, seed: int = 13, n_temps: int = 16,
            n_replicas: int = 256, n_iter: int = 4000) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    comp = rng.random(n) < 0.5
    y = np.where(comp, rng.normal(-3.0, 1.0, n), rng.normal(3.0, 1.0, n))
    y = y.astype(np.float32)
    m = gpum_model(MIXTURE, params=["mu1", "mu2"], data=["y"])
    betas = np.geomspace(1.0, 0.01, n_temps)
    return Workload("mixture", m, y.astype(np.float64).reshape(-1, 1),
                    n_temps * n_replicas, n_iter,
                    meta={"temperatures": 1.0 / betas, "n_temps": n_temps,
                          "n_replicas": n_replicas, "swap_every": 10})


def logistic(n: int = 4_000_000, d: int = 16, seed: int = 15,
             n_chains: int = 1024, n_iter: int = 2000) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    z = rng.standard_normal((n, d), dtype=np.float32)
    w = rng.standard_normal((n, 1), dtype=np.float32)
    X = (np.sqrt(0.7) * z + np.sqrt(0.3) * w).astype(np.float32)
    beta = rng.normal(0.0, 0.5, size=d)
    eta = X.astype(np.float64) @ beta
    y = (rng.random(n) < 1.0 / (1.0 + np.exp(-eta))).astype(np.float32)
    data = np.empty((n, d + 1), np.float64)
    data[:, 0] = y
    data[:, 1:] = X
    m = gpum_model(_logistic_formula(d), params=[f"b{j}" for j in range(1, d + 1)],
                   data=["y"] + [f"x{j}" for j in range(1, d + 1)],
                   prior=_logistic_prior(d))
    return Workload("logistic16", m, data, n_chains, n_iter,
                    meta={"true": beta})


REGRESSION = ("~ -((y - b0 - b1 * x1 - b2 * x2 - b3 * x3 - b4 * x4 - b5 * x5"
              " - b6 * x6 - b7 * x7 - b8 * x8)^2) / (2 * exp(2 * log_sigma))"
              " - log_sigma")
REGRESSION_PARAMS = [f"b{j}" for j in range(9)] + ["log_sigma"]


def regression_posterior(data: np.ndarray) -> Dict:
    """Closed-form posterior moments of one regression (flat priors on b
    and log sigma): b ~ t_{n-p}(beta_hat, s^2 (X'X)^-1) with variance
    s^2 (X'X)^-1 (n-p)/(n-p-2); sigma^2 ~ Inv-Gamma(a, b), a = (n-p)/2,
    b = (n-p) s^2 / 2, so E log sigma = (log b - digamma(a)) / 2 and
    Var log sigma = trigamma(a) / 4."""
    from scipy.special import digamma, polygamma
    y = data[:, 0]
    X = np.column_stack([np.ones(len(y)), data[:, 1:]])
    n, p = X.shape
    XtX = X.T @ X
    beta = np.linalg.solve(XtX, X.T @ y)
    rss = float(np.sum((y - X @ beta) ** 2))
    s2 = rss / (n - p)
    cov = s2 * np.linalg.inv(XtX) * (n - p) / (n - p - 2)
    a, b = (n - p) / 2.0, rss / 2.0
    mean = np.append(beta, 0.5 * (np.log(b) - digamma(a)))
    sd = np.append(np.sqrt(np.diag(cov)), 0.5 * np.sqrt(polygamma(1, a)))
    return {"mean": mean, "sd": sd}


def regressions(n_models: int = 512, seed: int = 16, n_min: int = 1000,
                n_max: int = 1_000_000, n_chains: int = 64,
                n_iter: int = 4000) -> List[Workload]:
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = np.floor(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_models)))
    m = gpum_model(REGRESSION, params=REGRESSION_PARAMS,
                   data=["y"] + [f"x{j}" for j in range(1, 9)])
    out = []
    for k, n in enumerate(ns.astype(int)):
        X = rng.standard_normal((n, 8), dtype=np.float32)
        beta = rng.standard_normal(9)
        y = (beta[0] + X.astype(np.float64) @ beta[1:] +
             0.5 * rng.standard_normal(n)).astype(np.float32)
        data = np.empty((n, 9), np.float64)
        data[:, 0] = y
        data[:, 1:] = X
        out.append(Workload(f"regression{k}", m, data, n_chains, n_iter,
                            meta={"true": np.append(beta, np.log(0.5))}))
    return out


def by_name(name: str, **kw) -> Workload:
    return {"gaussian_mean": gaussian_mean, "gamma": gamma, "mixture": mixture,
            "logistic16": logistic}[name](**kw)
