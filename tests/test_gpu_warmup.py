"""The adaptive warmup run on the device (gpum_hip_model_warmup) against
the host orchestration of the same warmup (api.adaptive_warmup, the
restatement of R/adaptive.R:101-254 driving one launch per batch).

Both draw the same Philox uniforms through the same kernel, so every
accept decision matches unless a chain meets a near-tie; the moments,
scalars and Cholesky factors then agree to rounding (the device merges the
pooled statistics over groups of chains, the host with numpy, and each
side rounds exp/log in its own library)."""
import numpy as np
import pytest

from gpumetropolis_amd import adaptive as ad
from gpumetropolis_amd import api, workloads

pytestmark = pytest.mark.gpu


def _model(hip, wl):
    m = wl.model
    return hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                           m.prior.consts, m.n_params, wl.data, wl.n_cols,
                           wl.n_obs)


CASES = [
    # name, workload, chains, warmup, auto_stop, init around, sd
    ("gaussian", lambda: workloads.gaussian_mean(n=20000), 8, 1500, False),
    ("gamma", lambda: workloads.gamma(n=40000, seed=3), 64, 1000, False),
    ("logistic3", lambda: workloads.logistic(n=20000, d=3, seed=9), 32, 1200,
     False),
    ("logistic3_freeze", lambda: workloads.logistic(n=20000, d=3, seed=9), 64,
     4000, True),
    # d = 16: a whole workgroup per chain's 16 x 16 moments, four merge
    # groups in the pooled accumulator
    ("logistic16", lambda: workloads.logistic(n=8000, d=16, seed=4), 24, 800,
     False),
]


def _start(wl, C, rng):
    np_ = wl.model.n_params
    if "true" in wl.meta:
        centre = np.asarray(wl.meta["true"], float)
    elif "post_mean" in wl.meta:
        centre = np.full(np_, wl.meta["post_mean"])
    else:
        centre = np.log([2.5, 0.7])
    init = centre + rng.normal(0, 0.02, (C, np_))
    return init, np.full((C, np_), 0.01)


@pytest.mark.parametrize("name,make,C,warmup,auto", CASES,
                         ids=[c[0] for c in CASES])
def test_device_warmup_matches_host(hip, name, make, C, warmup, auto):
    print(f"[warmup] {name} C={C} warmup={warmup} auto_stop={auto}",
          flush=True)
    wl = make()
    np_ = wl.model.n_params
    init, sd = _start(wl, C, np.random.default_rng(5))
    dm = _model(hip, wl)
    seed = 77
    host = api.adaptive_warmup(api._Caller(dm, C, None), np_, C, init, sd,
                               ad.warmup_batches(warmup), seed,
                               auto_stop=auto)
    dev = api.device_warmup(dm, np_, C, init, sd, warmup, seed,
                            auto_stop=auto)
    dm.close()
    print("batches", host.batches_run, dev.batches_run, "used", host.used,
          dev.used)
    assert dev.batches_run == host.batches_run
    assert dev.used == host.used
    assert dev.batch_sizes == host.batch_sizes
    if auto:
        assert dev.batches_run < len(ad.warmup_batches(warmup))
    # chains whose every batch acceptance matches took the same decisions
    same = np.all(dev.accept_history == host.accept_history, axis=1)
    print("chains with identical acceptance history", same.mean())
    assert same.mean() >= 0.95
    assert np.array_equal(dev.init[same], host.init[same])
    np.testing.assert_allclose(dev.scales[same], host.scales[same],
                               rtol=1e-12)
    np.testing.assert_allclose(dev.sd[same], host.sd[same], rtol=1e-12)
    if np_ > 1:
        assert dev.L is not None and host.L is not None
        Ld = dev.L.reshape(C, np_, np_)[same]
        Lh = host.L.reshape(C, np_, np_)[same]
        np.testing.assert_allclose(Ld, Lh, rtol=1e-10, atol=1e-14 * np.abs(Lh).max())
        assert np.all(np.triu(Ld, 1) == 0.0)
    else:
        assert dev.L is None and host.L is None
    # the warmup adapted: the last batch accepts near the target
    last = dev.accept_history[:, -1].mean()
    print("last batch acceptance", last, "target", ad.target_accept(np_))
    assert abs(last - ad.target_accept(np_)) < 0.15


def test_device_warmup_in_gpu_metropolis(hip):
    """The API's adaptive random-walk run takes the device warmup and still
    meets the Gaussian-mean closed form N(mean(y), 4/n) within 4 MCSE."""
    from gpumetropolis_amd.diagnostics import summarize
    print("[warmup] api", flush=True)
    wl = workloads.gaussian_mean(n=20000)
    fit = api.gpu_metropolis(wl.model, wl.data, n_iter=3000, n_chains=8,
                             warmup=1500, seed=4, proposal_sd=0.05,
                             init=np.full((8, 1), 3.4))
    s = summarize(fit.draws)["params"][0]
    print(s, fit.adaptation["n_batches"])
    assert abs(s["mean"] - wl.meta["post_mean"]) <= 4 * s["mcse"]
    assert abs(s["sd"] / wl.meta["post_sd"] - 1) < 0.1
    assert fit.adaptation["n_batches"] == len(ad.warmup_batches(1500))


def test_device_warmup_rejects_bad_input(hip):
    wl = workloads.gaussian_mean(n=1000)
    dm = _model(hip, wl)
    with pytest.raises(hip.BackendError, match="warmup > 0"):
        dm.warmup(np.zeros((4, 1)), np.full((4, 1), 0.1), 0, 1)
    with pytest.raises(hip.BackendError, match="positive and finite"):
        dm.warmup(np.zeros((4, 1)), np.zeros((4, 1)), 100, 1)
    dm.close()
