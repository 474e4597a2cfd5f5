"""Many models in one persistent batched launch.

The batched run must reproduce, draw for draw, the reference's per-model
DE orchestration (R/adaptive.R:427-524, restated by the host mirror) run
one model at a time on the same tiling, and its posteriors must match the
closed form of flat-prior linear regression.
"""
import math

import numpy as np
import pytest

from gpumetropolis_amd import api, workloads
from gpumetropolis_amd.batch import gpu_metropolis_batch
from gpumetropolis_amd.diagnostics import summarize

pytestmark = pytest.mark.gpu


def _models():
    g = workloads.gaussian_mean(n=30000, seed=4)
    ga = workloads.gamma(n=30011, seed=5)
    lo = workloads.logistic(n=26000, d=3, seed=6)
    rng = np.random.default_rng(1)
    inits = [rng.normal(3.4, 0.05, (16, 1)),
             np.log([2.5, 0.7]) + rng.normal(0, 0.02, (24, 2)),
             lo.meta["true"] + rng.normal(0, 0.02, (32, 3))]
    sds = [np.full((16, 1), 0.02), np.full((24, 2), 0.01), np.full((32, 3), 0.01)]
    return [g, ga, lo], inits, sds


def _per_model(hip, wl, k, init, sd, n_iter, warmup, de_every, seed):
    m = wl.model
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                         m.prior.consts, m.n_params, wl.data, wl.n_cols,
                         wl.n_obs)
    C = init.shape[0]

    def call(init_m, sd_m, iters, s, temps=None, mode=0, g=0.0, noise=0.0,
             L=None, iter_offset=0, trace=False):
        return dm.run(init_m, sd_m, np.ones(C), iters, int(s) & 0xFFFFFFFF,
                      proposal_mode=mode, gamma=g, de_noise=noise,
                      final_state=True, tiling=(16, 1), stream_id=k)

    fit = api._de(call, m, init, sd, None, 1e-3, n_iter, warmup, de_every, True,
                  seed, C)
    dm.close()
    return fit


def test_batch_reproduces_per_model_orchestration(hip):
    print("[batch] vs per-model", flush=True)
    wls, inits, sds = _models()
    n_iter, warmup, de_every, seed = 200, 80, 40, 9
    bf = gpu_metropolis_batch([w.model for w in wls], [w.data for w in wls],
                              n_chains=[16, 24, 32], n_iter=n_iter,
                              warmup=warmup, de_every=de_every, init=inits,
                              proposal_sd=sds, seed=seed, tiling=(16, 1))
    nog = gpu_metropolis_batch([w.model for w in wls], [w.data for w in wls],
                               n_chains=[16, 24, 32], n_iter=n_iter,
                               warmup=warmup, de_every=de_every, init=inits,
                               proposal_sd=sds, seed=seed, tiling=(16, 1),
                               use_graph=False)
    for k, w in enumerate(wls):
        ref = _per_model(hip, w, k, inits[k], sds[k], n_iter, warmup, de_every,
                         seed)
        got = bf.fits[k]
        print(k, got.draws.shape, ref.draws.shape,
              float(np.mean(got.accept_rate)))
        assert got.draws.shape == ref.draws.shape
        assert np.array_equal(got.draws, ref.draws), k
        assert np.array_equal(nog.fits[k].draws, got.draws), k
        assert np.allclose(got.adaptation["final_proposal_sd"],
                           ref.adaptation["final_proposal_sd"], rtol=0, atol=0)
        assert np.allclose(got.accept_rate, ref.accept_rate, rtol=1e-6)
        assert 0.02 < float(np.mean(got.accept_rate)) < 0.98
    # ceil(16/16), ceil(24/16), ceil(32/16) chain groups; each model goes
    # to its own XCD queue, which gets one workgroup per item of a batch
    # (at most 2), and the grid is 8 XCDs x 2 slots
    assert bf.n_workgroups == 8 * 2


def test_batch_queues_with_many_models_match_per_model(hip):
    """More models than XCDs: several models share each XCD's queue and
    their items interleave batch by batch, each waiting only for its own
    model's previous batch; every model still reproduces its per-model
    orchestration bit for bit."""
    print("[batch] 12 models, shared queues", flush=True)
    wls = workloads.regressions(n_models=12, seed=8, n_min=1000, n_max=40000)
    post = [workloads.regression_posterior(w.data) for w in wls]
    rng = np.random.default_rng(6)
    inits = [p["mean"] + 2 * p["sd"] * rng.standard_normal((32, 10))
             for p in post]
    sds = [np.tile(p["sd"], (32, 1)) for p in post]
    n_iter, warmup, de_every, seed = 160, 64, 32, 13
    bf = gpu_metropolis_batch([w.model for w in wls], [w.data for w in wls],
                              n_chains=32, n_iter=n_iter, warmup=warmup,
                              de_every=de_every, init=inits, proposal_sd=sds,
                              seed=seed, tiling=(16, 1))
    for k, w in enumerate(wls):
        ref = _per_model(hip, w, k, inits[k], sds[k], n_iter, warmup,
                         de_every, seed)
        got = bf.fits[k]
        assert np.array_equal(got.draws, ref.draws), k
        assert np.array_equal(got.adaptation["final_proposal_sd"],
                              ref.adaptation["final_proposal_sd"]), k
    # two chain groups of 16 per model, 12 models over 8 queues
    assert bf.n_workgroups == 8 * 4


def test_batch_regressions_match_closed_form(hip):
    """Six uneven regressions, DE-MC with 64 chains each: posterior means
    within 4 MCSE and sds within 15% of the closed form."""
    print("[batch] regressions", flush=True)
    wls = workloads.regressions(n_models=6, seed=21, n_min=1000, n_max=30000)
    post = [workloads.regression_posterior(w.data) for w in wls]
    rng = np.random.default_rng(2)
    inits = [p["mean"] + 3 * p["sd"] * rng.standard_normal((64, 10)) for p in post]
    sds = [np.tile(p["sd"], (64, 1)) for p in post]
    bf = gpu_metropolis_batch([w.model for w in wls], [w.data for w in wls],
                              n_chains=64, n_iter=3000, warmup=1000,
                              init=inits, proposal_sd=sds, seed=3,
                              tiling="auto")
    print("tiling", bf.tiling, bf.tile_timings, "ms", bf.kernel_ms)
    for k, (fit, p) in enumerate(zip(bf.fits, post)):
        s = summarize(fit.draws)
        for j, row in enumerate(s["params"]):
            z = abs(row["mean"] - p["mean"][j]) / row["mcse"]
            r = row["sd"] / p["sd"][j]
            assert z < 4.0, (k, j, row, p["mean"][j])
            assert 0.85 < r < 1.15, (k, j, row["sd"], p["sd"][j])
        assert s["max_rhat"] < 1.05


def test_batch_rejects_bad_input(hip):
    wls, inits, sds = _models()
    with pytest.raises(ValueError, match="at least 4 chains"):
        gpu_metropolis_batch([wls[0].model], [wls[0].data], n_chains=2)
    with pytest.raises(ValueError, match="method"):
        gpu_metropolis_batch([wls[0].model], [wls[0].data], method="pt")
