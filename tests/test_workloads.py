"""The posterior references the benchmark checks draws against
(gpumetropolis_amd/workloads.py), each against an independent computation.

gamma_posterior: Newton mode, Laplace covariance and whitened-grid moments
vs a brute-force grid in the raw (log_a, log_b) coordinates, where the log
density is summed observation by observation (no sufficient statistics).

regression_posterior: the Student-t / inverse-gamma closed form vs exact
draws from the conjugate factorisation sigma^2 | y, then b | sigma^2, y.
"""
import math

import numpy as np
import pytest
from scipy.special import gammaln

from gpumetropolis_amd import workloads


def test_gamma_posterior_matches_brute_force_grid():
    rng = np.random.default_rng(3)
    y = rng.gamma(2.5, 1 / 0.7, 60).astype(np.float32).astype(np.float64)
    p = workloads.gamma_posterior(y)
    la = p["mode"][0] + np.linspace(-9, 9, 700) * p["laplace_sd"][0]
    lb = p["mode"][1] + np.linspace(-9, 9, 700) * p["laplace_sd"][1]
    A, B = np.meshgrid(la, lb, indexing="ij")
    a, b = np.exp(A), np.exp(B)
    lp = np.zeros_like(A)
    for yi in y:            # the gamma log density, term by term
        lp += a * np.log(b) - gammaln(a) + (a - 1) * np.log(yi) - b * yi
    lp += -0.5 * (A / 10) ** 2 - 0.5 * (B / 10) ** 2
    w = np.exp(lp - lp.max())
    w /= w.sum()
    mean = np.array([(w * A).sum(), (w * B).sum()])
    sd = np.sqrt([(w * (A - mean[0]) ** 2).sum(), (w * (B - mean[1]) ** 2).sum()])
    np.testing.assert_allclose(p["mean"], mean, rtol=0, atol=2e-4 * sd.max())
    np.testing.assert_allclose(p["sd"], sd, rtol=2e-3)
    # the mode is where the brute-force density peaks
    i, j = np.unravel_index(np.argmax(lp), lp.shape)
    assert abs(A[i, j] - p["mode"][0]) < 2 * (la[1] - la[0])
    assert abs(B[i, j] - p["mode"][1]) < 2 * (lb[1] - lb[0])
    # small n: the posterior is skewed, so the mean is off the mode
    assert np.max(np.abs(p["mean"] - p["mode"]) / p["sd"]) > 0.01


def test_gamma_posterior_concentrates_on_the_truth():
    y = workloads.gamma(n=200_000).data
    p = workloads.gamma_posterior(y)
    z = (p["mean"] - np.log([2.5, 0.7])) / p["sd"]
    assert np.all(np.abs(z) < 4)
    np.testing.assert_allclose(p["sd"], p["laplace_sd"], rtol=1e-3)


def test_regression_posterior_matches_conjugate_draws():
    w = workloads.regressions(n_models=1, n_min=40, n_max=41, seed=5)[0]
    post = workloads.regression_posterior(w.data)
    y = w.data[:, 0]
    X = np.column_stack([np.ones(len(y)), w.data[:, 1:]])
    n, p = X.shape
    XtX_inv = np.linalg.inv(X.T @ X)
    bhat = XtX_inv @ X.T @ y
    rss = float(np.sum((y - X @ bhat) ** 2))
    rng = np.random.default_rng(0)
    S = 400_000
    # p(sigma) ∝ 1/sigma: sigma^2 | y ~ Inv-Gamma((n-p)/2, rss/2)
    s2 = (rss / 2) / rng.gamma((n - p) / 2, 1.0, S)
    Lc = np.linalg.cholesky(XtX_inv)
    bs = bhat + (rng.standard_normal((S, p)) @ Lc.T) * np.sqrt(s2)[:, None]
    draws = np.column_stack([bs, 0.5 * np.log(s2)])
    m, s = draws.mean(axis=0), draws.std(axis=0)
    se = s / math.sqrt(S)
    assert np.all(np.abs(m - post["mean"]) < 4.5 * se)
    np.testing.assert_allclose(s, post["sd"], rtol=0.01)
