"""CPU checks of the surrogate-network and delayed-acceptance oracles.

oracle/mlp.py is pinned against torch (its bf16 cast and float64 autograd,
library routines independent of the restatement) and against finite
differences; oracle/da.py's proposal / uniform / decision logic is pinned
against the C oracle's own sampler run on the same Philox stream.
"""
import math

import numpy as np
import pytest

from oracle import mlp


def test_bf16_matches_torch_round_to_nearest_even():
    import torch  # imported here: GPU test runs never load its runtime
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000),
        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 65504.0,
         3.4e38, -3.4e38, 1e-40, np.inf, -np.inf]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    got = mlp.bf16(x)
    assert np.array_equal(got, ref)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6
    assert mlp.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert mlp.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    assert np.isnan(mlp.bf16(np.float32(np.nan)))


def _net(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    w = np.zeros(mlp.N_WEIGHTS)
    w[mlp.OFF_W1:mlp.OFF_B1] = rng.uniform(-1, 1, 16 * 256) * 0.4 * scale
    w[mlp.OFF_B1:mlp.OFF_W2] = rng.uniform(-0.1, 0.1, 256)
    w[mlp.OFF_W2:mlp.OFF_B2] = rng.uniform(-1, 1, 256 * 256) * 0.1 * scale
    w[mlp.OFF_B2:mlp.OFF_W3] = rng.uniform(-0.1, 0.1, 256)
    w[mlp.OFF_W3:mlp.OFF_B3] = rng.uniform(-1, 1, 256) * 0.1
    w[mlp.OFF_B3] = 0.3
    return w


def test_exact_network_matches_torch_autograd():
    import torch
    w = _net(1)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((37, 16))
    t = rng.standard_normal(37)
    loss, g = mlp.loss_and_grad(w, X, t, rounding=False)
    p = mlp.unpack(w)
    tw = {k: torch.tensor(np.atleast_1d(v), dtype=torch.float64, requires_grad=True)
          for k, v in p.items()}
    Xt = torch.tensor(X)
    h1 = torch.nn.functional.silu(Xt @ tw["W1"] + tw["b1"])
    h2 = torch.nn.functional.silu(h1 @ tw["W2"] + tw["b2"])
    s = h2 @ tw["w3"] + tw["b3"]
    lt = torch.mean((s - torch.tensor(t)) ** 2)
    lt.backward()
    assert abs(loss - lt.item()) <= 1e-12 * max(1.0, abs(loss))
    ref = mlp.pack({k: (tw[k].grad.numpy() if k != "b3" else float(tw[k].grad[0]))
                    for k in tw})
    assert np.allclose(g, ref, rtol=1e-10, atol=1e-13)


def test_exact_backward_matches_finite_differences():
    w = _net(3)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((11, 16))
    t = rng.standard_normal(11)
    _, g = mlp.loss_and_grad(w, X, t, rounding=False)
    idx = np.concatenate([rng.integers(lo, hi, 6) for lo, hi in [
        (mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2),
        (mlp.OFF_W2, mlp.OFF_B2), (mlp.OFF_B2, mlp.OFF_W3),
        (mlp.OFF_W3, mlp.OFF_B3)]] + [[mlp.OFF_B3]])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        lp = mlp.loss_and_grad(wp, X, t, rounding=False)[0]
        lm = mlp.loss_and_grad(wm, X, t, rounding=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])


def test_rounded_network_is_close_to_exact():
    """The bf16 rounding points move the output by bf16 precision only."""
    w = _net(5)
    X = np.random.default_rng(6).standard_normal((64, 16))
    s_r, _ = mlp.forward(w, X, rounding=True)
    s_e, _ = mlp.forward(w, X, rounding=False)
    assert np.max(np.abs(s_r - s_e)) < 0.05 * np.max(np.abs(s_e)) + 1e-3


def test_features_whiten():
    rng = np.random.default_rng(7)
    A = rng.standard_normal((3, 3))
    cov = A @ A.T + np.eye(3)
    th = rng.multivariate_normal([1.0, -2.0, 0.5], cov, size=20000)
    from gpumetropolis_amd.surrogate import whitening
    mean, W = whitening(th)
    x = mlp.features(th, mean, W, rounding=False)
    assert x.shape == (20000, 16)
    assert np.allclose(np.cov(x[:, :3], rowvar=False), np.eye(3), atol=1e-8)
    assert np.all(x[:, 3:] == 0)


def test_da_replay_reproduces_the_oracle_sampler(orc):
    """replay_da with the surrogate off walks the C oracle's own Philox run
    (modes 0 and 2) decision for decision."""
    from oracle import da
    from gpumetropolis_amd import workloads
    wl = workloads.gaussian_mean(n=500, seed=3)
    m = wl.model
    rng = np.random.default_rng(8)
    for mode in (0, 2):
        C, T = 4, 60
        init = rng.normal(3.0, 0.5, (C, 1))
        sd = np.full((C, 1), 0.15)
        L = np.full((C, 1, 1), 0.15) if mode == 2 else None
        r = orc.run(m.loglik.code, m.loglik.consts, m.prior.code,
                    m.prior.consts, 1, wl.data, 1, wl.n_obs, init, sd,
                    np.ones(C), T, 17, proposal_mode=mode,
                    proposal_l=None if L is None else L.ravel(),
                    stream=orc.STREAM_PHILOX, trace=True)

        class Fake:
            trace_stage = np.where(r.trace_accept, 2, 1).astype(np.uint8)
            trace_plp = r.trace_plp
            trace_s = None

        rep = da.replay_da(Fake, m, wl.data, 1, wl.n_obs, init, np.ones(C), 17,
                           proposal_sd=sd if mode == 0 else None,
                           proposal_l=L, surrogate=False)
        assert rep.exact_checked == C * T
        assert rep.certified_flips == 0 and rep.uncertified_flips == 0
        assert rep.max_plp_rel < 1e-12
        assert 0.2 < r.trace_accept.mean() < 0.95
