"""Synchronous Differential Evolution (path B, cpu_native.rs:400-506) on the
GPU's stepwise sampler, through the C ABI, against the oracle driven by the
same Philox uniforms: the same pairs, scales and deviates give the same
proposals (1e-12 relative), the proposal log-posterior agrees to 1e-6
relative and every accept decision is identical up to the first certified
near-tie (after which the population legitimately parts, for every chain,
since each proposal reads the whole previous generation)."""
import math

import numpy as np
import pytest

from gpumetropolis_amd import workloads
from gpumetropolis_amd.diagnostics import summarize

pytestmark = pytest.mark.gpu


def _init_lp(orc, m, data, init):
    """FP64 log-posterior of the initial states (oracle)."""
    ll = orc.loglik_batch(m.loglik.code, m.loglik.consts, m.n_params, data,
                          data.shape[1], data.shape[0], init)
    if len(m.prior.code):
        ll = ll + orc.loglik_batch(m.prior.code, m.prior.consts, m.n_params,
                                   np.zeros((1, 0)), 0, 1, init)
    return ll


@pytest.mark.parametrize("name,centre,spread", [
    ("gaussian_mean", [3.4], 0.3),
    ("gamma", [np.log(2.5), np.log(0.7)], 0.05),
])
def test_de_sync_matches_oracle_draw_for_draw(hip, orc, name, centre, spread):
    w = workloads.by_name(name, n=4000)
    m = w.model
    C, T, seed = 16, 60, 31
    g, noise = 2.38 / math.sqrt(2 * m.n_params), 1e-3
    rng = np.random.default_rng(2)
    init = np.asarray(centre) + spread * rng.standard_normal((C, m.n_params))
    sd = np.full((C, m.n_params), 0.05)
    print(f"[de_sync] {name} C={C} T={T}", flush=True)
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                         m.prior.consts, m.n_params, w.data, w.n_cols, w.n_obs)
    gpu = dm.run_da(None, init, np.ones(C), T, seed, proposal_sd=sd,
                    trace=True, de_sync=(g, noise))
    dm.close()
    o = orc.run(m.loglik.code, m.loglik.consts, m.prior.code, m.prior.consts,
                m.n_params, w.data, w.n_cols, w.n_obs, init, sd, np.ones(C), T,
                seed, proposal_mode=1, gamma=g, de_noise=noise, de_sync=True,
                trace=True)
    acc_g = gpu.trace_stage == 2
    cur_o = _init_lp(orc, m, w.data, init)
    cur_g = cur_o.copy()
    compared = 0
    for t in range(T):
        scale = np.maximum(np.abs(o.trace_prop[t]), 1.0)
        assert np.max(np.abs(gpu.trace_prop[t] - o.trace_prop[t]) / scale) <= 1e-12, t
        rel = np.abs(gpu.trace_plp[t] - o.trace_plp[t]) / np.abs(o.trace_plp[t])
        assert rel.max() <= 1e-6, (t, rel.max())
        flips = np.flatnonzero(acc_g[t] != o.trace_accept[t])
        if flips.size:
            # Every mismatch must be a near-tie the evaluation error cannot
            # resolve; the populations part from here.
            for c in flips:
                u = orc.uniform_philox(seed, 0, 0, t, c, 2 * m.n_params + 3)
                d64 = o.trace_plp[t, c] - cur_o[c]
                dg = gpu.trace_plp[t, c] - cur_g[c]
                assert abs(math.log(u) - d64) <= abs(dg - d64) * (1 + 1e-9) + 1e-12
            break
        cur_o = np.where(o.trace_accept[t], o.trace_plp[t], cur_o)
        cur_g = np.where(acc_g[t], gpu.trace_plp[t], cur_g)
        compared += 1
    assert compared >= T // 2, compared
    if compared == T:
        # Same decisions: the f32 draws agree (FP64 proposals equal to 1e-12).
        assert np.allclose(gpu.draws, o.draws, rtol=1e-6, atol=0.0)
        assert np.array_equal(gpu.accept_rate, o.accept_rate)


def test_de_sync_entry_point_equals_resident_model(hip):
    """gpum_hip_run_de_sync (the stateless mirror of
    rust_gpu_metropolis_de_sync) and the resident model's mode 1 are the
    same computation: identical draws."""
    w = workloads.gaussian_mean(n=3000)
    m = w.model
    C, T = 8, 40
    init = np.linspace(2.5, 4.5, C).reshape(C, 1)
    sd = np.full((C, 1), 0.1)
    a = hip.run_de_sync(m.loglik.code, m.loglik.consts, m.prior.code,
                        m.prior.consts, 1, w.data, 1, w.n_obs, init, sd, T, 5,
                        1.68, 1e-3)
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                         m.prior.consts, 1, w.data, 1, w.n_obs)
    b = dm.run_da(None, init, np.ones(C), T, 5, proposal_sd=sd,
                  de_sync=(1.68, 1e-3))
    dm.close()
    assert np.array_equal(a.draws, b.draws)
    assert np.array_equal(a.last_log_post, b.last_log_post)


def test_de_sync_pair_excludes_self(hip):
    """Three chains, no jitter: each step is +-g times the difference of the
    two other chains of the previous generation."""
    w = workloads.gaussian_mean(n=500)
    m = w.model
    C, T, gam = 3, 40, 0.7
    init = np.array([[2.0], [3.1], [4.5]])
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                         m.prior.consts, 1, w.data, 1, w.n_obs)
    r = dm.run_da(None, init, np.ones(C), T, 9, proposal_sd=np.full((C, 1), 0.1),
                  trace=True, de_sync=(gam, 0.0))
    dm.close()
    prev = init[:, 0].copy()
    for t in range(T):
        for c in range(C):
            o1, o2 = [k for k in range(C) if k != c]
            step = abs(r.trace_prop[t, c, 0] - prev[c])
            d = abs(prev[o1] - prev[o2])
            assert min(abs(step - d), abs(step - gam * d)) <= 1e-12 * (1 + d)
        prev = np.where(r.trace_stage[t] == 2, r.trace_prop[t, :, 0], prev)


def test_de_sync_posterior_closed_form(hip):
    """Gaussian mean, known variance: the posterior mean is the sample mean
    (flat prior); the DE-B draws after warmup lie within 3 MCSE of it."""
    from gpumetropolis_amd.api import gpu_metropolis
    w = workloads.gaussian_mean(n=4000)
    y = w.data[:, 0]
    fit = gpu_metropolis(w.model, w.data, init=np.linspace(3.0, 3.8, 8),
                         n_iter=6000, warmup=1000, method="de", de_sync=True,
                         proposal_sd=0.05, seed=3, jit=False)
    s = summarize(fit.draws)["params"][0]
    post_mean = y.mean()
    print(f"[de_sync] mean {s['mean']:.5f} vs {post_mean:.5f} mcse {s['mcse']:.2e}")
    assert abs(s["mean"] - post_mean) <= 3 * s["mcse"]
    assert fit.adaptation["de_sync"]


def test_de_sync_rejects_bad_arguments(hip):
    w = workloads.gaussian_mean(n=100)
    m = w.model
    sd = np.full((2, 1), 0.1)
    with pytest.raises(hip.BackendError, match="at least 3 chains"):
        hip.run_de_sync(m.loglik.code, m.loglik.consts, m.prior.code,
                        m.prior.consts, 1, w.data, 1, w.n_obs,
                        np.array([[3.0], [3.5]]), sd, 10, 1, 1.0, 0.0)
    with pytest.raises(hip.BackendError, match="gamma must be positive"):
        hip.run_de_sync(m.loglik.code, m.loglik.consts, m.prior.code,
                        m.prior.consts, 1, w.data, 1, w.n_obs,
                        np.full((4, 1), 3.0), np.full((4, 1), 0.1), 10, 1,
                        0.0, 0.0)
