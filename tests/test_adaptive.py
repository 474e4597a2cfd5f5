"""Adaptation primitives (gpumetropolis_amd/adaptive.py) against the
reference's own known-answer tests and the mathematics they rest on.

  Welford == sample covariance     tests/testthat/test-adaptive.R:6-22
  order invariance                 test-adaptive.R:24-31
  seed Sigma below two draws       test-adaptive.R:33-38
  Robbins-Monro direction, decay   test-adaptive.R:40-60
  Cholesky reconstruction          test-adaptive.R:62-69
  singular Sigma regularised       test-adaptive.R:71-78
  geometric ladder                 tests/testthat/test-pt.R:6-14
  warmup batch plan                R/adaptive.R:112-118 (closed-form sizes)
"""
import numpy as np
import pytest

from gpumetropolis_amd import adaptive as ad


def _welford_one(X):
    """Fold the rows of X (T, d) one draw at a time through a 1-chain state."""
    T, d = X.shape
    st = ad.WelfordState.init(1, d, np.ones(d))
    st.update_batch(X[:, None, :])
    return st


def test_welford_reproduces_var_d1():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(500)
    st = _welford_one(x[:, None])
    assert st.mean[0, 0] == pytest.approx(np.mean(x), rel=1e-12, abs=1e-15)
    assert st.covariance()[0, 0, 0] == pytest.approx(np.var(x, ddof=1), rel=1e-12)


def test_welford_batch_reproduces_cov_d3():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((300, 3))
    st = _welford_one(X)
    np.testing.assert_allclose(st.mean[0], X.mean(axis=0), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(st.covariance()[0], np.cov(X, rowvar=False),
                               rtol=1e-11, atol=1e-14)


def test_welford_order_invariant():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((300, 2))
    s1 = _welford_one(X)
    s2 = _welford_one(X[rng.permutation(300)])
    np.testing.assert_allclose(s1.mean, s2.mean, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(s1.covariance(), s2.covariance(), rtol=1e-12,
                               atol=1e-14)


def test_welford_chains_are_independent():
    """Per-chain state: chain c's statistics depend on chain c's draws only."""
    rng = np.random.default_rng(9)
    X = rng.standard_normal((200, 4, 3)) * np.array([1.0, 2.0, 0.5, 3.0])[None, :, None]
    st = ad.WelfordState.init(4, 3, np.ones(3))
    st.update_batch(X)
    for c in range(4):
        np.testing.assert_allclose(st.covariance()[c], np.cov(X[:, c], rowvar=False),
                                   rtol=1e-11, atol=1e-14)


def test_covariance_is_seed_sigma_below_two_draws():
    st = ad.WelfordState.init(1, 2, np.full(2, 0.5))
    np.testing.assert_array_equal(st.covariance()[0], np.diag([0.5, 0.5]))
    st.update_batch(np.array([[[1.0, 2.0]]]))
    np.testing.assert_array_equal(st.covariance()[0], np.diag([0.5, 0.5]))
    st.update_batch(np.array([[[3.0, 2.0]]]))
    np.testing.assert_allclose(st.covariance()[0], [[2.0, 0.0], [0.0, 0.0]])


def test_pooled_welford_equals_cov_of_all_draws():
    """The pooled accumulator (R/adaptive.R:140-177) folds every chain's batch;
    after any sequence of batches it equals the covariance of all draws."""
    rng = np.random.default_rng(5)
    pw = ad.PooledWelford.init(3, np.ones(3))
    np.testing.assert_array_equal(pw.covariance(), np.eye(3))
    allx = []
    for T in (7, 1, 40, 13):
        X = rng.standard_normal((T, 5, 3)) + rng.standard_normal(3)
        pw.update_batch(X)
        allx.append(X.transpose(1, 0, 2).reshape(-1, 3))
    A = np.concatenate(allx)
    assert pw.n == A.shape[0]
    np.testing.assert_allclose(pw.mean, A.mean(axis=0), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(pw.covariance(), np.cov(A, rowvar=False),
                               rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("d", [1, 2, 5])
def test_robbins_monro_direction(d):
    t = ad.target_accept(d)
    assert ad.robbins_monro_scale(1.0, t + 0.1, 1, d) > 1.0
    assert ad.robbins_monro_scale(1.0, t - 0.1, 1, d) < 1.0
    assert ad.robbins_monro_scale(1.0, t, 1, d) == 1.0


def test_robbins_monro_step_decays():
    early = ad.robbins_monro_scale(1.0, 0.5, 1, 2) - 1.0
    late = ad.robbins_monro_scale(1.0, 0.5, 100, 2) - 1.0
    assert early > late > 0.0
    # gain b^(-2/3): log-scale moves are exactly 100^(-2/3) apart
    r = np.log(ad.robbins_monro_scale(1.0, 0.5, 100, 2)) / \
        np.log(ad.robbins_monro_scale(1.0, 0.5, 1, 2))
    assert r == pytest.approx(100.0 ** (-2.0 / 3.0), rel=1e-12)


def test_target_accept_table():
    """Roberts-Gelman-Gilks: 0.44 in d=1 falling towards 0.234 (R/adaptive.R:69-72)."""
    t = [ad.target_accept(d) for d in range(1, 12)]
    assert t[0] == 0.44 and t[-1] == 0.234
    assert all(a >= b for a, b in zip(t, t[1:]))


def test_cholesky_reconstructs_sigma():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((3, 3))
    S = A.T @ A + 0.1 * np.eye(3)
    L = ad.regularised_cholesky(S, eps=1e-12)
    np.testing.assert_allclose(L @ L.T, S, rtol=1e-6, atol=1e-9)
    assert np.all(L[np.triu_indices(3, 1)] == 0.0)
    assert np.all(np.diag(L) > 0)


def test_cholesky_regularises_singular():
    S = np.ones((2, 2))
    L = ad.regularised_cholesky(S, eps=1e-3)
    assert L.shape == (2, 2) and np.all(np.isfinite(L))
    assert np.all(np.linalg.eigvalsh(L @ L.T) > 0)
    # the ridge is eps * mean(diag): L L' = S + 1e-3 I
    np.testing.assert_allclose(L @ L.T, S + 1e-3 * np.eye(2), rtol=1e-9)


def test_cholesky_indefinite_falls_back_to_diagonal():
    S = np.array([[1.0, 2.0], [2.0, 1.0]])
    L = ad.regularised_cholesky(S)
    assert np.all(np.isfinite(L))
    assert L[0, 1] == 0.0 and L[1, 0] == 0.0


def test_default_ladder_geometric():
    T = ad.default_ladder(8, t_max=10.0)
    assert len(T) == 8 and T[0] == 1.0
    assert T[-1] == pytest.approx(10.0, rel=1e-12)
    assert np.all(np.diff(T) > 0)
    r = T[1:] / T[:-1]
    assert np.max(np.abs(r - r[0])) < 1e-9
    assert list(ad.default_ladder(1)) == [1.0]


@pytest.mark.parametrize("warmup,expect", [
    (1250, [50] * 24 + [50]),          # 25 batches of 50
    (100, [10] * 10),                   # at least 10 batches
    (5000, [125] * 40),                 # at most 40 batches
    (523, [52] * 9 + [55]),             # the last absorbs the remainder
    (7, [1] * 6 + [1]),                 # fewer draws than batches
])
def test_warmup_batches(warmup, expect):
    b = ad.warmup_batches(warmup)
    assert sum(b) == warmup
    assert b == expect


def test_split_batches():
    assert ad.split_batches(4000, 10) == ([10] * 400, 400)
    sizes, nb = ad.split_batches(105, 10)
    assert nb == 10 and sum(sizes) == 105 and sizes[-1] == 15
    # R/adaptive.R:284-289 adds only a positive remainder: fewer iterations
    # than one batch still run one whole batch of `every`.
    assert ad.split_batches(5, 10) == ([10], 1)
