"""Surrogate network kernels and the stepwise / delayed-acceptance sampler.

The network's forward pass and gradients are compared with the numpy
restatement in oracle/mlp.py (same bf16 rounding points, FP64 elsewhere);
the sampler is replayed decision by decision by oracle/da.py on the same
Philox uniforms, and its posterior is checked against closed forms.
"""
import math

import numpy as np
import pytest

from gpumetropolis_amd import diagnostics as dg
from gpumetropolis_amd import dsl, workloads
from gpumetropolis_amd.surrogate import fit_surrogate, whitening

pytestmark = pytest.mark.gpu


def _weights(seed):
    from oracle import mlp
    rng = np.random.default_rng(seed)
    w = np.zeros(mlp.N_WEIGHTS, np.float32)
    w[mlp.OFF_W1:mlp.OFF_B1] = rng.uniform(-0.5, 0.5, 16 * 256)
    w[mlp.OFF_B1:mlp.OFF_W2] = rng.uniform(-0.2, 0.2, 256)
    w[mlp.OFF_W2:mlp.OFF_B2] = rng.uniform(-0.1, 0.1, 256 * 256)
    w[mlp.OFF_B2:mlp.OFF_W3] = rng.uniform(-0.2, 0.2, 256)
    w[mlp.OFF_W3:mlp.OFF_B3] = rng.uniform(-0.2, 0.2, 256)
    w[mlp.OFF_B3] = 0.25
    return w


def _net(hip, d, seed, n=700):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0.5, 1.3, (n, d))
    mean, W = whitening(theta)
    net = hip.Surrogate(d, seed)
    net.set_scaling(mean, W, -3.0, 2.0)
    w = _weights(seed)
    if d < 16:
        w[d * 256:16 * 256] = 0.0  # rows of absent inputs
    net.set_weights(w)
    return net, theta, mean, W, w


@pytest.mark.parametrize("d", [16, 5])
def test_mlp_forward_matches_oracle(hip, d):
    from oracle import mlp
    print(f"[mlp] forward d={d}", flush=True)
    net, theta, mean, W, w = _net(hip, d, 1)
    got = net.predict(theta)
    X = mlp.features(theta, mean, W)
    s, _ = mlp.forward(w.astype(np.float64), X)
    want = -3.0 + 2.0 * s
    err = np.abs(got - want)
    # FP32 accumulation and rare one-ulp bf16 flips of a hidden unit
    assert err.max() <= 2e-3 * max(1.0, np.abs(want).max()), err.max()
    assert np.array_equal(net.get_weights(), w)


def test_mlp_gradient_matches_oracle(hip):
    from oracle import mlp
    print("[mlp] gradient", flush=True)
    net, theta, mean, W, w = _net(hip, 16, 2, n=300)
    rng = np.random.default_rng(3)
    ll = -3.0 + 2.0 * rng.standard_normal(300)
    g, loss = net.gradient(theta, ll)
    X = mlp.features(theta, mean, W)
    t = (ll - (-3.0)) / 2.0
    loss_o, g_o = mlp.loss_and_grad(w.astype(np.float64), X, t)
    assert abs(loss - loss_o) <= 1e-4 * loss_o
    blocks = {"W1": (mlp.OFF_W1, mlp.OFF_B1), "b1": (mlp.OFF_B1, mlp.OFF_W2),
              "W2": (mlp.OFF_W2, mlp.OFF_B2), "b2": (mlp.OFF_B2, mlp.OFF_W3),
              "w3": (mlp.OFF_W3, mlp.OFF_B3), "b3": (mlp.OFF_B3, mlp.N_WEIGHTS)}
    for name, (lo, hi) in blocks.items():
        a, b = g[lo:hi].astype(np.float64), g_o[lo:hi]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(name, rel)
        assert rel <= 5e-3, (name, rel)


def test_mlp_training_fits_a_quadratic(hip):
    """Training on exact pairs drives the held-out error of a 6-dim
    Gaussian log-likelihood surface well below its spread."""
    print("[mlp] train", flush=True)
    rng = np.random.default_rng(4)
    d = 6
    A = rng.standard_normal((d, d))
    P = A @ A.T / d + np.eye(d)
    theta = rng.multivariate_normal(np.arange(d) * 0.1, np.linalg.inv(P), 60000)
    ll = -5000.0 - 0.5 * np.einsum("ni,ij,nj->n", theta, P, theta)
    fit = fit_surrogate(theta, ll, steps=3000, seed=5)
    print(fit.summary(), fit.loss_trace[:3], fit.loss_trace[-3:])
    assert fit.usable
    assert fit.holdout_err < 0.05 * np.std(ll)
    assert fit.loss_trace[-200:].mean() < 0.05 * fit.loss_trace[:20].mean()
    fit.net.close()


def _replay_plain(hip, wl, mode, C=8, T=80, seed=21, work_items=0):
    from oracle import da
    m = wl.model
    np_ = m.n_params
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                         m.prior.consts, np_, wl.data, wl.n_cols, wl.n_obs)
    rng = np.random.default_rng(seed)
    truth = np.asarray(wl.meta.get("true", np.zeros(np_)), float).ravel()
    init = truth + rng.normal(0, 0.02, (C, np_))
    sd = np.full((C, np_), 0.01)
    L = None
    if mode == 2:
        L = np.zeros((C, np_, np_))
        L[:] = np.eye(np_) * 0.01
        L[:, -1, 0] = 0.004
    r = dm.run_da(None, init, np.ones(C), T, seed,
                  proposal_sd=sd if mode == 0 else None,
                  proposal_l=None if L is None else L.ravel(), trace=True,
                  work_items=work_items)
    rep = da.replay_da(r, m, wl.data, wl.n_cols, wl.n_obs, init, np.ones(C),
                       seed, proposal_sd=sd if mode == 0 else None,
                       proposal_l=L, surrogate=False)
    dm.close()
    return r, rep


@pytest.mark.parametrize("name,mode,work_items", [("gamma", 0, 0),
                                                  ("gamma", 2, 3),
                                                  ("logistic", 2, 0)])
def test_stepwise_plain_replays_oracle(hip, name, mode, work_items):
    print(f"[stepwise] {name} mode {mode} items {work_items}", flush=True)
    if name == "gamma":
        wl = workloads.gamma(n=30011)
    else:
        wl = workloads.logistic(n=20000, d=4, seed=3)
    r, rep = _replay_plain(hip, wl, mode, work_items=work_items)
    print(rep)
    assert rep.max_plp_rel <= 1e-6
    assert rep.uncertified_flips == 0
    assert rep.exact_checked >= 0.9 * r.trace_stage.size
    # the traced proposals and data terms are what the replay evaluated
    from oracle import oracle as orc
    m = wl.model
    ll = orc.loglik_batch(m.loglik.code, m.loglik.consts, m.n_params, wl.data,
                          wl.n_cols, wl.n_obs, r.trace_prop[:3].reshape(-1, m.n_params))
    assert np.allclose(r.trace_ll[:3].ravel(), ll, rtol=1e-6, atol=0)
    assert np.all(r.stage1_rate == 1.0)
    assert 0.05 < r.accept_rate.mean() < 0.95


def test_stepwise_plain_matches_fused_kernel(hip):
    """Without a surrogate the stepwise sampler draws the fused kernel's
    uniforms: identical draws unless an FP64 near-tie separates them."""
    print("[stepwise] vs fused", flush=True)
    wl = workloads.gaussian_mean(n=5000)
    m = wl.model
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                         m.prior.consts, 1, wl.data, 1, wl.n_obs)
    C, T = 16, 200
    init = np.linspace(3.0, 3.8, C).reshape(C, 1)
    sd = np.full((C, 1), 0.05)
    a = dm.run(init, sd, np.ones(C), T, 9)
    b = dm.run_da(None, init, np.ones(C), T, 9, proposal_sd=sd)
    dm.close()
    same = np.all(a.draws == b.draws, axis=(0, 2))
    assert same.mean() >= 0.9
    assert np.allclose(a.last_log_post[same], b.last_log_post[same], rtol=1e-9)


def _logistic_fit(hip, d=6, n=40000, C=256, warm=400, seed=7):
    """Warmup pairs of a small logistic regression and a surrogate on them."""
    from gpumetropolis_amd import adaptive as ad
    from gpumetropolis_amd.api import _Caller, adaptive_warmup
    wl = workloads.logistic(n=n, d=d, seed=seed)
    m = wl.model
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                         m.prior.consts, d, wl.data, wl.n_cols, wl.n_obs)
    call = _Caller(dm, C, None)
    rng = np.random.default_rng(seed)
    init = wl.meta["true"] + rng.normal(0, 0.02, (C, d))
    w = adaptive_warmup(call, d, C, init, np.full((C, d), 0.01),
                        ad.warmup_batches(warm), seed, collect=True)
    fit = fit_surrogate(w.pairs_theta, w.pairs_loglik, seed=seed, steps=3000)
    return wl, dm, w, fit


@pytest.mark.parametrize("d,n,C", [(6, 40000, 64), (16, 20000, 128)])
def test_da_replays_oracle(hip, d, n, C):
    """Every stage-1 decision follows from the trace's surrogate values and
    the Philox uniforms; every exact value matches the oracle to 1e-6; every
    stage-2 decision matches or is a certified near-tie.  d = 16 is the
    dimension of the 16-parameter logistic regression workload."""
    from oracle import da
    print(f"[da] replay d={d} n={n} C={C}", flush=True)
    wl, dm, w, fit = _logistic_fit(hip, d=d, n=n, C=C)
    print(fit.summary())
    assert fit.usable
    assert wl.model.n_params == d
    T = 120
    L = w.L.reshape(C, d, d)
    r = dm.run_da(fit.net, w.init, np.ones(C), T, 99, proposal_l=w.L,
                  trace=True)
    init_s = fit.net.predict(w.init)
    from oracle import oracle as orc
    pr0 = np.array([orc.vm_eval(wl.model.prior.code, wl.model.prior.consts,
                                w.init[c], [0.0]) for c in range(C)])
    rep = da.replay_da(r, wl.model, wl.data, wl.n_cols, wl.n_obs, w.init,
                       np.ones(C), 99, init_s=init_s + pr0, proposal_l=L)
    print(rep, r.stage1_rate.mean(), r.accept_rate.mean())
    dm.close()
    fit.net.close()
    assert rep.stage1_mismatch == 0
    assert rep.uncertified_flips == 0
    assert rep.max_plp_rel <= 1e-6
    assert rep.exact_checked >= 0.1 * C * T
    # the exact stage ran only for stage-1 survivors
    assert 0.05 < r.stage1_rate.mean() < 0.8


def test_da_gains_over_plain(hip):
    """What the trained network buys: the same posterior from a fraction of
    the exact evaluations (ESS per exact evaluation at least twice plain)."""
    print("[da] gain", flush=True)
    wl, dm, w, fit = _logistic_fit(hip, C=256)
    assert fit.usable
    C, T = 256, 600
    plain = dm.run_da(None, w.init, np.ones(C), T, 5, proposal_l=w.L)
    dar = dm.run_da(fit.net, w.init, np.ones(C), T, 5, proposal_l=w.L)
    dm.close()
    fit.net.close()
    dp = plain.draws.astype(np.float64)[100:]
    dd = dar.draws.astype(np.float64)[100:]
    sp_, sd_ = dg.summarize(dp)["params"], dg.summarize(dd)["params"]
    ess_p = min(r["ess"] for r in sp_)
    ess_d = min(r["ess"] for r in sd_)
    per_p = ess_p / plain.exact_evals
    per_d = ess_d / dar.exact_evals
    print("plain", ess_p, plain.exact_evals, "da", ess_d, dar.exact_evals,
          "ratio", per_d / per_p)
    assert per_d / per_p >= 2.0
    # same posterior: means within 4 combined MCSE
    for a, b in zip(sp_, sd_):
        mcse = math.hypot(a["mcse"], b["mcse"])
        assert abs(a["mean"] - b["mean"]) <= 4 * mcse, (a, b)


def test_da_keeps_the_target_with_a_poor_surrogate(hip):
    """A surrogate of the wrong curvature changes efficiency, not the
    target: the Gaussian-mean posterior N(mean(y), 4/n) within 3 MCSE."""
    print("[da] poor surrogate", flush=True)
    wl = workloads.gaussian_mean(n=20000)
    m = wl.model
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                         m.prior.consts, 1, wl.data, 1, wl.n_obs)
    post_m, post_s = wl.meta["post_mean"], wl.meta["post_sd"]
    rng = np.random.default_rng(3)
    th = rng.normal(post_m, 3 * post_s, (20000, 1))
    # twice the true curvature, shifted by half a posterior sd
    ll_wrong = -((th[:, 0] - post_m - 0.5 * post_s) ** 2) / post_s ** 2
    net = hip.Surrogate(1, 3)
    mean, W = whitening(th)
    net.set_scaling(mean, W, float(ll_wrong.mean()), float(ll_wrong.std()))
    net.train(th, ll_wrong, steps=1500)
    C, T = 64, 1500
    init = rng.normal(post_m, post_s, (C, 1))
    r = dm.run_da(net, init, np.ones(C), T, 11,
                  proposal_sd=np.full((C, 1), 1.7 * post_s))
    net.close()
    dm.close()
    x = r.draws.astype(np.float64)[200:]
    s = dg.summarize(x)["params"][0]
    print(s, r.stage1_rate.mean(), r.accept_rate.mean())
    assert abs(s["mean"] - post_m) <= 3 * s["mcse"]
    assert abs(s["sd"] - post_s) / post_s < 0.1
    assert r.accept_rate.mean() < 0.9 * r.stage1_rate.mean()


def test_api_surrogate_used_and_fallback(hip):
    print("[da] api", flush=True)
    wl = workloads.logistic(n=20000, d=3, seed=9)
    kw = dict(n_iter=1200, n_chains=64, seed=4, proposal_sd=0.02,
              init=np.tile(wl.meta["true"], (64, 1)))
    from gpumetropolis_amd.api import gpu_metropolis
    fit = gpu_metropolis(wl.model, wl.data, surrogate=True, **kw)
    info = fit.adaptation["surrogate"]
    print(info)
    assert info["used"] and info["holdout_err"] <= info["bound"]
    assert 0.0 < info["stage1_rate"] < 1.0
    assert np.all(np.isfinite(fit.draws))
    off = gpu_metropolis(wl.model, wl.data, surrogate={"bound": 1e-9}, **kw)
    assert off.adaptation["surrogate"]["used"] is False
    assert "stage1_rate" not in off.adaptation["surrogate"]
    with pytest.raises(ValueError, match="surrogate"):
        gpu_metropolis(wl.model, wl.data, surrogate=True, adapt=False, **kw)
