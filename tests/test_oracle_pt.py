"""The parallel-tempering oracle (oracle/pt.py) against what the mathematics
fixes, so that it can serve as the reference for the fused PT kernel.

* Invariant: on the target exp(-mu^2 / 2) every rung k samples the tempered
  density exp(-mu^2 / (2 T_k)), i.e. N(0, T_k).  With swaps on, the product
  of the tempered targets is the stationary law, so each rung's variance is
  T_k.  A flipped sign of the swap statistic, or 1/T taken as T, moves states
  towards the wrong rungs and breaks this.  Independent ladder groups give
  independent estimates, so the standard error needs no autocorrelation
  model.
* The sweep order is the reference's (R/adaptive.R:362-372): odd batches try
  pairs 0, 2, 4, ..., even batches pairs 1, 3, ...
* Adaptation: the oracle's in-loop Welford and Robbins-Monro equal the
  pinned primitives of gpumetropolis_amd/adaptive.py applied to its own
  draws and batch acceptances (R/adaptive.R:342-356).
"""
import numpy as np
import pytest

from gpumetropolis_amd import adaptive as ad
from gpumetropolis_amd.dsl import gpum_model
from oracle import pt

STD_NORMAL = gpum_model("~ -((y - mu)^2) / 2", params=["mu"], data=["y"])
DATA = np.zeros((1, 1))


def _run(temps_group, n_groups, n_iter, swap_every, seed, adapt_batches=0,
         sd_scale=2.4):
    K = len(temps_group)
    temps = np.tile(temps_group, n_groups)
    rng = np.random.default_rng(seed)
    init = rng.standard_normal(K * n_groups) * np.sqrt(temps)
    sd = sd_scale * np.sqrt(temps)
    m = STD_NORMAL
    return pt.run_pt(m.loglik.code, m.loglik.consts, m.prior.code,
                     m.prior.consts, 1, DATA, 1, 1, init, sd, temps, n_iter,
                     seed, swap_every, K, adapt_batches=adapt_batches), temps


def test_each_rung_samples_its_tempered_target():
    ladder = ad.default_ladder(4, t_max=10.0)
    G = 64
    out, temps = _run(ladder, G, 4000, 10, seed=7)
    x = out["draws"][500:, :, 0].astype(np.float64)    # (T, C)
    m2 = (x * x).mean(axis=0).reshape(G, 4)               # per group, per rung
    est = m2.mean(axis=0)
    se = m2.std(axis=0, ddof=1) / np.sqrt(G)
    for k in range(4):
        assert abs(est[k] - ladder[k]) < 3 * se[k], (k, est[k], ladder[k], se[k])
    # swaps happen, and at a rate below one
    sh = out["swap_hist"]
    tried = sh[sh != 255]
    assert 0.05 < tried.mean() < 0.95


def test_swaps_move_the_cold_rung():
    """A cold chain started far in the tail comes back through swaps within a
    few batches; without swaps (one-chain groups) its random walk is slowed by
    a tiny step."""
    ladder = ad.default_ladder(4, t_max=10.0)
    temps = np.tile(ladder, 8)
    init = np.tile([6.0, 0.0, 0.0, 0.0], 8)
    sd = 0.05 * np.sqrt(temps)
    m = STD_NORMAL
    args = (m.loglik.code, m.loglik.consts, m.prior.code, m.prior.consts, 1,
            DATA, 1, 1, init, sd, temps, 400, 3, 10)
    on = pt.run_pt(*args, 4)
    off = pt.run_pt(*args, 1)
    cold_on = np.abs(on["draws"][-1, 0::4, 0]).mean()
    cold_off = np.abs(off["draws"][-1, 0::4, 0]).mean()
    assert cold_on < 3.0 < cold_off


def test_sweep_alternates_even_and_odd_pairs():
    ladder = ad.default_ladder(5, t_max=10.0)
    out, _ = _run(ladder, 3, 60, 10, seed=11)
    sh = out["swap_hist"]                     # (batch, chain): 255 = untried
    for b in range(sh.shape[0]):
        tried = [i for i in range(5) if sh[b, i] != 255]
        assert tried == ([0, 2] if (b + 1) % 2 == 1 else [1, 3]), (b, tried)
        assert np.all(sh[b, 4::5] == 255)     # the top rung opens no pair


def test_adaptation_equals_pinned_primitives():
    ladder = ad.default_ladder(4, t_max=10.0)
    G, A, S = 4, 6, 10
    out, temps = _run(ladder, G, 100, S, seed=5, adapt_batches=A, sd_scale=0.3)
    C = 4 * G
    sd0 = 0.3 * np.sqrt(temps)
    st = ad.WelfordState.init(C, 1, (sd0 * sd0)[:, None])
    scale = np.ones(C)
    X = out["draws"].astype(np.float64)
    for b in range(1, A + 1):
        st.update_batch(X[(b - 1) * S:b * S])
        acc = out["acc_hist"][b - 1].astype(np.float64)
        scale = ad.robbins_monro_scale(scale, acc, b, 1)
        var = st.covariance()[:, 0, 0]
        sd = scale * np.sqrt(np.maximum(var, 1e-12))
    np.testing.assert_allclose(out["final_scale"], scale, rtol=1e-13)
    np.testing.assert_allclose(out["final_sd"][:, 0], sd, rtol=1e-12)
