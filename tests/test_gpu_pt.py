"""Parallel tempering and per-chain adaptation fused into one launch.

The kernel's batches, Robbins-Monro adaptation and replica-exchange sweeps
are checked against the oracle's restatement of the reference's host loop
(oracle/pt.py over the C oracle), against the same kernel driven one batch
per launch, and statistically against test-pt.R's bimodal target.
"""
import math

import numpy as np
import pytest

from gpumetropolis_amd import adaptive as ad, dsl
from gpumetropolis_amd.api import gpu_metropolis

pytestmark = pytest.mark.gpu

BIMODAL = "~ log(exp(-((y - mu)^2) / 2) + exp(-((y + mu)^2) / 2))"


def _bimodal(n=400, seed=1):
    y = np.random.default_rng(seed).normal(3.0, 1.0, n)
    return dsl.gpum_model(BIMODAL, ["mu"], ["y"]), y


def _first_divergence(g, o):
    """First (t, c) where the decisions differ, or None."""
    diff = np.argwhere(g.trace_accept != o["trace_accept"])
    return None if diff.size == 0 else tuple(diff[diff[:, 0].argmin()])


def test_pt_fused_matches_oracle(hip, orc):
    from oracle import pt as opt
    m, y = _bimodal()
    C, T, se, G = 8, 300, 10, 8
    init = np.linspace(-6, 6, C).reshape(C, 1)
    sd = np.full((C, 1), 0.15)
    temps = ad.default_ladder(C)
    args = (m.loglik.code, m.loglik.consts, [], [], 1, y, 1, y.size, init, sd,
            temps, T, 5)
    print("[pt] fused vs oracle", flush=True)
    g = hip.run_chains(*args, trace=True, final_state=True,
                       batches={"swap_every": se, "swap_group": G,
                                "adapt_batches": 10})
    o = opt.run_pt(*args, swap_every=se, group=G, adapt_batches=10, trace=True)
    first = _first_divergence(g, o)
    T_ok = T if first is None else first[0]
    print("first divergence:", first, flush=True)
    # Everything before the first differing decision is identical.
    assert np.allclose(g.trace_prop[:T_ok], o["trace_prop"][:T_ok], rtol=1e-12,
                       atol=1e-12)
    rel = np.abs(g.trace_plp[:T_ok] - o["trace_plp"][:T_ok]) / np.abs(o["trace_plp"][:T_ok])
    assert rel.max() <= 1e-6
    nb_ok = T_ok // se
    assert np.array_equal(g.swap_hist[:nb_ok], o["swap_hist"][:nb_ok])
    assert np.array_equal(g.acc_hist[:nb_ok], o["acc_hist"][:nb_ok])
    # FP32 terms move the decision margin by ~1e-9 here: no decision of
    # this run sits that close, so the whole run must agree.
    assert first is None
    assert np.array_equal(g.draws, o["draws"])
    assert np.allclose(g.final_sd, o["final_sd"], rtol=1e-12)
    assert np.allclose(g.final_scale, o["final_scale"], rtol=1e-12)
    assert (g.swap_hist[g.swap_hist != 255] == 1).mean() > 0.05


def _current_lp(init_lp, plp, acc, swap_hist, sizes):
    """Per-iteration current log-posterior of every chain, rebuilt from a
    traced PT run: accepted proposals replace it, and at each batch end the
    accepted swaps exchange it between the pair.  Also the end-of-batch
    values the swap sweep saw."""
    T, C = plp.shape
    cur = np.asarray(init_lp, float).copy()
    out = np.empty((T, C))
    ends = np.empty((len(sizes), C))
    t = 0
    for b, n in enumerate(sizes):
        for _ in range(n):
            out[t] = cur
            cur = np.where(acc[t], plp[t], cur)
            t += 1
        ends[b] = cur
        for lo in np.nonzero(swap_hist[b] == 1)[0]:
            cur[lo], cur[lo + 1] = cur[lo + 1], cur[lo]
    return out, ends


def test_pt_mixture_ladders_match_oracle(hip, orc):
    """The configuration's ladder shape: two independent 16-rung ladders,
    beta 1 .. 0.01 (T = 1 .. 100), on the 2-parameter mixture (20,000 rows).
    Per ladder, everything up to its first differing decision is identical
    (proposals, log-posteriors to 1e-6, swap decisions, batch acceptances),
    and that first decision is certified: its FP64 margin lies within the
    GPU's own evaluation error.  |log-post| ~ 4e4 here, so FP32 sums leave
    ~1e-3 nats and a few near-ties are expected."""
    from oracle import pt as opt
    from gpumetropolis_amd import workloads
    w = workloads.mixture(n=20000)
    m = w.model
    G, NG, T, se, seed = 16, 2, 400, 10, 8
    C, P = G * NG, 2
    temps = np.tile(1.0 / np.geomspace(1.0, 0.01, G), NG)
    rng = np.random.default_rng(6)
    init = rng.normal(0.0, 3.0, (C, P))
    sd = 0.02 * np.sqrt(temps)[:, None] * np.ones((C, P))
    args = (m.loglik.code, m.loglik.consts, [], [], P, w.data, 1, w.n_obs,
            init, sd, temps, T, seed)
    print("[pt] mixture 2 x 16 ladders vs oracle", flush=True)
    g = hip.run_chains(*args, trace=True, final_state=True,
                       batches={"swap_every": se, "swap_group": G,
                                "adapt_batches": 10})
    o = opt.run_pt(*args, swap_every=se, group=G, adapt_batches=10, trace=True)
    sizes = o["batch_sizes"]
    nb = len(sizes)
    ll0 = orc.loglik_batch(m.loglik.code, m.loglik.consts, P, w.data, 1,
                           w.n_obs, init)
    cur_o, end_o = _current_lp(ll0, o["trace_plp"], o["trace_accept"],
                               o["swap_hist"], sizes)
    cur_g, end_g = _current_lp(g.init_log_post, g.trace_plp, g.trace_accept,
                               g.swap_hist, sizes)
    compared = 0
    for g0 in range(0, C, G):
        cs = slice(g0, g0 + G)
        dif = np.argwhere(g.trace_accept[:, cs] != o["trace_accept"][:, cs])
        t_acc = int(dif[:, 0].min()) if dif.size else T
        sdif = np.argwhere(g.swap_hist[:, cs] != o["swap_hist"][:, cs])
        b_sw = int(sdif[:, 0].min()) if sdif.size else nb
        t_ok = min(t_acc, (b_sw + 1) * se)
        print(f"group {g0 // G}: first decision flip t={t_acc}, first swap "
              f"flip batch={b_sw}", flush=True)
        assert np.allclose(g.trace_prop[:t_ok, cs], o["trace_prop"][:t_ok, cs],
                           rtol=1e-12, atol=1e-12)
        rel = np.abs(g.trace_plp[:t_ok, cs] - o["trace_plp"][:t_ok, cs]) / \
            np.abs(o["trace_plp"][:t_ok, cs])
        assert rel.max() <= 1e-6
        b_ok = t_ok // se
        b_sw_ok = min(b_ok, b_sw)
        assert np.array_equal(g.swap_hist[:b_sw_ok, cs], o["swap_hist"][:b_sw_ok, cs])
        assert np.array_equal(g.acc_hist[:b_ok, cs], o["acc_hist"][:b_ok, cs])
        if dif.size and t_acc < (b_sw + 1) * se:   # a decision flipped
            c = g0 + int(dif[dif[:, 0] == t_acc][0, 1])
            u = orc.uniform_philox(seed, 0, 0, t_acc, c, 2 * P)
            d64 = (o["trace_plp"][t_acc, c] - cur_o[t_acc, c]) / temps[c]
            dg = (g.trace_plp[t_acc, c] - cur_g[t_acc, c]) / temps[c]
            assert abs(math.log(u) - d64) <= abs(dg - d64) * (1 + 1e-9) + 1e-12
        elif sdif.size:                       # a swap decision flipped
            lo = g0 + int(sdif[sdif[:, 0] == b_sw][0, 1])
            b1 = b_sw + 1
            start = 0 if b1 % 2 == 1 else 1
            k = (lo - g0 - start) // 2
            wv = orc.philox([b1, k, g0 // G, opt.TAG_SWAP], [seed, 0])
            u = (float(wv[0]) + 0.5) / 4294967296.0
            dt = 1.0 / temps[lo] - 1.0 / temps[lo + 1]
            d64 = (end_o[b_sw, lo + 1] - end_o[b_sw, lo]) * dt
            dg = (end_g[b_sw, lo + 1] - end_g[b_sw, lo]) * dt
            assert abs(math.log(u) - d64) <= abs(dg - d64) * (1 + 1e-9) + 1e-12
        compared += t_ok
    assert compared >= T * NG // 3
    assert (g.swap_hist[g.swap_hist != 255] == 1).mean() > 0.05


def test_pt_fused_equals_batched_launches(hip):
    """The in-kernel batches reproduce one launch per batch exactly (same
    tiling, same stream)."""
    m, y = _bimodal()
    C = 8
    init = np.linspace(-6, 6, C).reshape(C, 1)
    kw = dict(n_iter=400, method="pt", init=init, proposal_sd=0.15, seed=3,
              swap_every=10, tiling=(8, 2))
    fused = gpu_metropolis(m, {"y": y}, **kw)
    # the batched orchestrator (used for ladders longer than 16), driven
    # with the same arguments and tiling
    from gpumetropolis_amd import api, _native
    dm = _native.DeviceModel(m.loglik.code, m.loglik.consts, [], [], 1, y, 1,
                             y.size)
    call = api._Caller(dm, C, (8, 2))
    sizes, nb = ad.split_batches(400, 10)
    wb = 200 // 10
    d, acc_hist, swap_hist, sd, sc = api._pt_batched(
        call, 1, C, init, np.full((C, 1), 0.15), ad.default_ladder(C), sizes,
        wb, True, 3, C)
    dm.close()
    assert np.array_equal(fused.draws, d[200:])
    assert np.array_equal(fused.adaptation["swap_history"],
                          np.where(swap_hist[:, :C - 1] == 255, np.nan,
                                   swap_hist[:, :C - 1]).T, equal_nan=True)
    assert np.allclose(fused.adaptation["final_proposal_sd"], sd, rtol=1e-12)


def test_pt_cold_chain_visits_both_modes(hip):
    """test-pt.R:29-48."""
    m, y = _bimodal()
    init = np.linspace(-6, 6, 8).reshape(8, 1)
    fit = gpu_metropolis(m, {"y": y}, init=init, proposal_sd=0.15,
                         n_iter=4000, method="pt", seed=1)
    assert fit.method == "pt" and len(fit.temperatures) == 8
    assert fit.temperatures[0] == 1.0
    cold = fit.draws[:, 0, 0]
    assert np.any(cold > 1.5) and np.any(cold < -1.5)
    assert abs(cold[cold > 0].mean() - 3) < 0.4
    assert abs(cold[cold < 0].mean() + 3) < 0.4


def test_pt_grouped_ladders(hip):
    """Independent ladders of 4 chains: every group swaps within itself."""
    m, y = _bimodal(n=200, seed=2)
    C, G = 16, 4
    init = np.tile(np.linspace(-4, 4, G), C // G).reshape(C, 1)
    fit = gpu_metropolis(m, {"y": y}, init=init, proposal_sd=0.2, n_iter=800,
                         method="pt", ladder_size=G, seed=7, swap_every=10)
    sh = fit.adaptation["swap_history"]
    assert sh.shape[0] == C - C // G  # pairs inside the groups only
    assert np.nanmean(sh) > 0.05
    assert np.all(np.isfinite(fit.draws))


def test_pt_rejects_bad_ladders(hip):
    m, y = _bimodal(n=50)
    with pytest.raises(ValueError, match="one value per chain"):
        gpu_metropolis(m, {"y": y}, n_iter=100, n_chains=4, method="pt",
                       temperatures=[1, 2])
    with pytest.raises(ValueError, match="positive and finite"):
        gpu_metropolis(m, {"y": y}, n_iter=100, n_chains=4, method="pt",
                       temperatures=[1, 2, -1, 3])
