"""ESS, split R-hat and MCSE (gpumetropolis_amd/diagnostics.py) against the
reference's known-answer tests and against closed forms.

  ESS of iid draws in (3000, 4000]     tests/testthat/test-diagnostics.R:60-66
  R-hat > 1.5 on stuck chains          test-diagnostics.R:12-18
  R-hat refuses < 4 kept draws         test-diagnostics.R:20-22
  constant chain -> ESS = n            R/diagnostics.R:38-40

Closed forms: the AR(1) chain x_t = rho x_{t-1} + e_t has integrated
autocorrelation time (1 + rho) / (1 - rho), so ESS = n (1 - rho) / (1 + rho);
the sample autocorrelation is checked against its defining sum.
"""
import math

import numpy as np
import pytest

from gpumetropolis_amd import diagnostics as dg


def _ar1(n, rho, chains, seed):
    rng = np.random.default_rng(seed)
    e = rng.standard_normal((n, chains))
    x = np.empty((n, chains))
    x[0] = e[0] / math.sqrt(1 - rho * rho)
    for t in range(1, n):
        x[t] = rho * x[t - 1] + e[t]
    return x


def test_acf_matches_defining_sum():
    """stats::acf(demean=TRUE): c_k = (1/n) sum_{t<n-k} (x_t - m)(x_{t+k} - m),
    rho_k = c_k / c_0; the module uses an FFT, the check a direct sum."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((257, 2)).cumsum(axis=0)
    r = dg._acf(x, 40)
    n = x.shape[0]
    for j in range(2):
        v = x[:, j] - x[:, j].mean()
        c0 = np.dot(v, v) / n
        for k in range(1, 41):
            ck = np.dot(v[:n - k], v[k:]) / n
            assert r[k - 1, j] == pytest.approx(ck / c0, rel=1e-10, abs=1e-12)


def test_ess_iid_close_to_count():
    rng = np.random.default_rng(8)
    v = rng.standard_normal(4000)
    e = dg.ess(v)
    assert 3000 < e <= 4000


@pytest.mark.parametrize("rho", [0.5, 0.9])
def test_ess_ar1_closed_form(rho):
    n = 200_000
    x = _ar1(n, rho, 1, seed=int(rho * 10))
    want = n * (1 - rho) / (1 + rho)
    got = dg.ess(x)
    assert abs(got / want - 1) < 0.10, (got, want)


def test_ess_sums_over_chains():
    x = _ar1(20_000, 0.7, 4, seed=3)
    per = dg.ess_chains(x)
    assert dg.ess(x) == pytest.approx(per.sum())
    want = 20_000 * 0.3 / 1.7
    assert np.all(np.abs(per / want - 1) < 0.15)


def test_ess_antithetic_capped_at_n():
    """Negative lag-1 correlation: the first pair sum is <= 0, so Geyer's
    sequence is empty and the reference returns n (R/diagnostics.R:50-53)."""
    x = _ar1(10_000, -0.5, 1, seed=4)
    assert dg.ess(x) == 10_000.0


def test_ess_constant_and_short_chains():
    assert dg.ess(np.full((50, 1), 3.0)) == 50.0
    assert dg.ess(np.arange(7.0)) == 7.0      # n < 8 returns n


def test_rhat_near_one_for_iid_chains():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4000, 4))
    assert dg.rhat(x) < 1.01


def test_rhat_flags_stuck_chains():
    rng = np.random.default_rng(2)
    stuck = np.column_stack([rng.normal(0, 0.1, 400), rng.normal(10, 0.1, 400)])
    assert dg.rhat(stuck) > 1.5


def test_rhat_flags_drift_within_a_chain():
    """Split R-hat sees a chain whose halves disagree (the reason to split)."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2000, 1))
    x[1000:] += 3.0
    assert dg.rhat(x) > 1.5


def test_rhat_needs_four_draws():
    with pytest.raises(ValueError):
        dg.rhat(np.arange(6.0).reshape(3, 2))
    dg.rhat(np.random.default_rng(0).standard_normal((4, 2)))


def test_rhat_equal_half_means_closed_form():
    """With every half-chain mean equal, B = 0 and R-hat = sqrt((h-1)/h)."""
    rng = np.random.default_rng(6)
    h = 50
    base = rng.standard_normal(h)
    cols = []
    for _ in range(3):
        a = rng.permutation(base)
        b = rng.permutation(base)
        cols.append(np.concatenate([a, b]))
    x = np.column_stack(cols)
    assert dg.rhat(x) == pytest.approx(math.sqrt((h - 1) / h), rel=1e-12)


def test_rhat_affine_invariant():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((300, 3)) + np.array([0.0, 0.2, -0.1])
    assert dg.rhat(5.0 * x - 2.0) == pytest.approx(dg.rhat(x), rel=1e-12)


def test_rhat_constant_is_nan():
    assert math.isnan(dg.rhat(np.ones((10, 2))))


def test_summarize_mcse_of_iid_mean():
    """MCSE = sd / sqrt(ESS) (R/diagnose.R:70): for iid N(0, 1) draws it is
    about 1/sqrt(n), and the mean lies within a few MCSE of 0."""
    rng = np.random.default_rng(9)
    d = rng.standard_normal((5000, 4, 2))
    s = dg.summarize(d)
    for p in s["params"]:
        assert p["mcse"] == pytest.approx(1 / math.sqrt(20000), rel=0.1)
        assert abs(p["mean"]) < 4 * p["mcse"]
    assert s["max_rhat"] < 1.01
    assert s["min_ess"] > 15000


def test_summarize_equals_the_per_parameter_functions():
    """summarize() transforms all parameters at once; each parameter's ESS and
    R-hat equal ess() and rhat() of its (n_iter, n_chains) matrix."""
    x = _ar1(3000, 0.8, 12, seed=11).reshape(3000, 4, 3)
    x[:, :, 2] = x[:, :, 2] * 3.0 + 1.0
    x[:, 1, 1] += 0.5            # one chain off: R-hat above 1
    s = dg.summarize(x)
    for p in range(3):
        M = x[:, :, p]
        row = s["params"][p]
        assert row["ess"] == pytest.approx(dg.ess(M), rel=1e-12)
        assert row["rhat"] == pytest.approx(dg.rhat(M), rel=1e-12)
        assert row["mean"] == pytest.approx(M.mean(), rel=1e-12)
        assert row["sd"] == pytest.approx(M.std(ddof=1), rel=1e-12)
    assert s["max_rhat"] == pytest.approx(dg.rhat(x[:, :, 1]), rel=1e-12)


# The AR spectral ESS (coda::effectiveSize's estimator), pinned by what the
# mathematics fixes: independent draws give about n, an AR(1) chain with
# coefficient rho gives n (1 - rho) / (1 + rho), a constant chain 0.
def test_ess_ar_iid():
    x = np.random.default_rng(3).standard_normal(20000)
    assert abs(dg.ess_ar(x) / 20000 - 1.0) < 0.05


@pytest.mark.parametrize("rho", [0.5, 0.9, -0.3])
def test_ess_ar_ar1(rho):
    rng = np.random.default_rng(7)
    n = 200000
    e = rng.standard_normal(n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - rho * rho)
    for t in range(1, n):
        x[t] = rho * x[t - 1] + e[t]
    want = n * (1 - rho) / (1 + rho)
    assert abs(dg.ess_ar(x) / want - 1.0) < 0.06


def test_ess_ar_constant_chain():
    assert dg.ess_ar(np.full(500, 2.5)) == 0.0
