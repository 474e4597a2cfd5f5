"""Per-model compiled kernels against the interpreter.

The compiled evaluator calls the interpreter's own per-instruction code
with constant operands and both are built with contraction limited to one
expression, so every launch kind must reproduce the interpreter bit for
bit: data terms, decisions and draws.
"""
import numpy as np
import pytest

from gpumetropolis_amd import dsl, workloads
from gpumetropolis_amd.batch import gpu_metropolis_batch

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(0)
    out = []
    g = workloads.gaussian_mean(n=3000)                     # resident
    out.append(("gaussian", g, rng.normal(3.4, 0.05, (16, 1)), 0.03))
    ga = workloads.gamma(n=40000)                           # streamed
    out.append(("gamma", ga, np.log([2.5, 0.7]) + rng.normal(0, 0.01, (32, 2)), 0.005))
    mx = workloads.mixture(n=30000)
    out.append(("mixture", mx, rng.normal(0, 2, (16, 2)), 0.02))
    lo = workloads.logistic(n=30000, d=5, seed=2)
    out.append(("logistic", lo, lo.meta["true"] + rng.normal(0, 0.01, (32, 5)), 0.005))
    # a deeper program: several stack slots, pow, sqrt, lgamma, divisions
    m = dsl.gpum_model("~ lgamma(y + a) - sqrt(y * y + b ^ 2) / (1 + (y - a) ^ 2) "
                       "- log(2 + exp(-y / b)) * (a - y) ^ 3 / 100",
                       params=["a", "b"], data=["y"])
    y = rng.gamma(2.0, 1.0, 20000)
    wl = workloads.Workload("deep", m, y.reshape(-1, 1), 16, 100)
    out.append(("deep", wl, np.column_stack([rng.uniform(1, 2, 16), rng.uniform(0.5, 1, 16)]), 0.01))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_compiled_chain_kernel_equals_interpreter(hip, case):
    name, wl, init, step = case
    print(f"[jit] fused {name}", flush=True)
    m = wl.model
    args = (m.loglik.code, m.loglik.consts, m.prior.code, m.prior.consts,
            m.n_params, wl.data, wl.n_cols, wl.n_obs)
    vm = hip.DeviceModel(*args)
    jt = hip.DeviceModel(*args, jit=True)
    C = init.shape[0]
    sd = np.full_like(init, step)
    a = vm.run(init, sd, np.ones(C), 60, 5, trace=True)
    b = jt.run(init, sd, np.ones(C), 60, 5, trace=True)
    assert np.array_equal(a.trace_plp, b.trace_plp, equal_nan=True)
    assert np.array_equal(a.draws, b.draws)
    assert np.array_equal(a.last_log_post, b.last_log_post)
    pts = init + 0.01
    assert np.array_equal(vm.loglik(pts), jt.loglik(pts))
    # the stepwise sampler (points kernel)
    c = vm.run_da(None, init, np.ones(C), 30, 7, proposal_sd=sd, trace=True)
    d = jt.run_da(None, init, np.ones(C), 30, 7, proposal_sd=sd, trace=True)
    assert np.array_equal(c.trace_plp, d.trace_plp, equal_nan=True)
    assert np.array_equal(c.draws, d.draws)
    vm.close()
    jt.close()


def test_compiled_loglik_matches_fp64(hip, orc):
    wl = workloads.logistic(n=20000, d=16, seed=4)
    m = wl.model
    jt = hip.DeviceModel(m.loglik.code, m.loglik.consts, [], [], 16, wl.data,
                         wl.n_cols, wl.n_obs, jit=True)
    pts = wl.meta["true"] + np.random.default_rng(1).normal(0, 0.05, (8, 16))
    got = jt.loglik(pts)
    ref = orc.loglik_batch(m.loglik.code, m.loglik.consts, 16, wl.data,
                           wl.n_cols, wl.n_obs, pts)
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-6
    jt.close()


def test_compiled_batch_equals_interpreter(hip):
    print("[jit] batch", flush=True)
    wls = workloads.regressions(n_models=5, seed=31, n_min=2000, n_max=40000)
    rng = np.random.default_rng(3)
    inits = [w.meta["true"] + 0.01 * rng.standard_normal((32, 10)) for w in wls]
    kw = dict(n_chains=32, n_iter=120, warmup=40, de_every=40, init=inits,
              proposal_sd=0.003, seed=2, tiling=(16, 1))
    a = gpu_metropolis_batch([w.model for w in wls], [w.data for w in wls], **kw)
    b = gpu_metropolis_batch([w.model for w in wls], [w.data for w in wls],
                             jit=True, **kw)
    for fa, fb in zip(a.fits, b.fits):
        assert np.array_equal(fa.draws, fb.draws)


def _pair_cases():
    rng = np.random.default_rng(11)
    out = []
    g = workloads.gaussian_mean(n=3000)                     # resident, square
    out.append(("gaussian", g, rng.normal(3.4, 0.05, (7, 1)), 0.03))
    ga = workloads.gamma(n=40000)                           # streamed, affine
    out.append(("gamma", ga, np.log([2.5, 0.7]) + rng.normal(0, 0.01, (7, 2)), 0.005))
    rg = workloads.regressions(n_models=1, seed=5, n_min=30011, n_max=30011)[0]
    out.append(("regression", rg, rg.meta["true"] + 0.01 * rng.standard_normal((7, 10)), 0.003))
    return out


@pytest.mark.parametrize("jit", [False, True], ids=["vm", "jit"])
@pytest.mark.parametrize("case", _pair_cases(), ids=lambda c: c[0])
def test_two_chains_per_wave_equal_one(hip, case, jit):
    """A wave evaluating two (or four) chains on each observation it reads
    does each chain's arithmetic exactly as a wave of one chain: the same
    waves per chain give identical data terms, decisions and draws (7
    chains: the last group has idle members)."""
    name, wl, init, step = case
    print(f"[pairs] {name} jit={jit}", flush=True)
    m = wl.model
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                         m.prior.consts, m.n_params, wl.data, wl.n_cols,
                         wl.n_obs, jit=jit)
    C = init.shape[0]
    sd = np.full_like(init, step)
    # (waves per chain group, chains per wave); four on compiled kernels
    # and, for affine programs, on the affine interpreter
    quads = jit or name == "gamma"
    pairs = [(2, 2), (4, 2)] + ([(4, 4)] if quads else [])
    for wpc, cw in pairs:
        a = dm.run(init, sd, np.ones(C), 40, 3, trace=True, tiling=(16 // wpc, wpc, 1))
        b = dm.run(init, sd, np.ones(C), 40, 3, trace=True,
                   tiling=(16 * cw // wpc, wpc, cw))
        assert np.array_equal(a.trace_plp, b.trace_plp, equal_nan=True), (wpc, cw)
        assert np.array_equal(a.draws, b.draws), (wpc, cw)
    # the points kernel (one slice, so the sums are ordered alike)
    c = dm.run_da(None, init, np.ones(C), 20, 7, proposal_sd=sd, trace=True,
                  work_items=1, chains_per_wave=1)
    for cw in (2, 4) if quads else (2,):
        d = dm.run_da(None, init, np.ones(C), 20, 7, proposal_sd=sd, trace=True,
                      work_items=1, chains_per_wave=cw)
        assert np.array_equal(c.trace_plp, d.trace_plp, equal_nan=True), cw
        assert np.array_equal(c.draws, d.draws), cw
    dm.close()


def test_two_chains_per_wave_batch(hip):
    print("[pairs] batch", flush=True)
    wls = workloads.regressions(n_models=3, seed=32, n_min=2000, n_max=40000)
    rng = np.random.default_rng(4)
    inits = [w.meta["true"] + 0.01 * rng.standard_normal((32, 10)) for w in wls]
    kw = dict(n_chains=32, n_iter=80, warmup=40, de_every=40, init=inits,
              proposal_sd=0.003, seed=2)
    ms, ds = [w.model for w in wls], [w.data for w in wls]
    a = gpu_metropolis_batch(ms, ds, tiling=(8, 2), **kw)
    for jit in (False, True):
        b = gpu_metropolis_batch(ms, ds, tiling=(16, 2, 2), jit=jit, **kw)
        for fa, fb in zip(a.fits, b.fits):
            assert np.array_equal(fa.draws, fb.draws)


def test_four_chains_per_wave_batch(hip):
    """Affine programs (gamma) batched with four chains per wave, on the
    compiled kernel and on the affine interpreter: the draws of one chain
    per wave, bit for bit.  The general interpreter's batched kernels
    refuse four (a square-sum program)."""
    print("[quads] batch", flush=True)
    wls = [workloads.gamma(n=n, seed=40 + k) for k, n in enumerate((3000, 7000))]
    rng = np.random.default_rng(5)
    inits = [np.array([np.log(2.5), np.log(0.7)]) + 0.01 * rng.standard_normal((16, 2))
             for _ in wls]
    kw = dict(n_chains=16, n_iter=60, warmup=20, de_every=20, init=inits,
              proposal_sd=0.01, seed=4)
    ms, ds = [w.model for w in wls], [w.data for w in wls]
    a = gpu_metropolis_batch(ms, ds, tiling=(16, 1), **kw)
    for jit in (True, False):
        b = gpu_metropolis_batch(ms, ds, tiling=(16, 4, 4), jit=jit, **kw)
        for fa, fb in zip(a.fits, b.fits):
            assert np.array_equal(fa.draws, fb.draws), jit
    rg = workloads.regressions(n_models=1, seed=5, n_min=3000, n_max=3000)[0]
    kw.update(init=[rg.meta["true"] + 0.01 * rng.standard_normal((16, 10))])
    with pytest.raises(hip.BackendError, match="affine programs only"):
        gpu_metropolis_batch([rg.model], [rg.data], tiling=(16, 4, 4),
                             jit=False, **kw)


def test_four_chains_per_wave_square_batch(hip):
    """Square-sum programs (regressions, 4 observations per lane) batched
    with four chains per wave on the compiled kernel: the draws of the
    interpreter at one chain per wave, bit for bit, and the closed-form
    posterior mean of b0 within 4 MCSE."""
    print("[quads] square batch", flush=True)
    wls = workloads.regressions(n_models=3, n_min=1500, n_max=6000, seed=7)
    rng = np.random.default_rng(6)
    inits = [w.meta["true"] + 0.01 * rng.standard_normal((16, 10)) for w in wls]
    kw = dict(n_chains=16, n_iter=1600, warmup=600, de_every=20, init=inits,
              proposal_sd=0.003, seed=9)
    ms, ds = [w.model for w in wls], [w.data for w in wls]
    a = gpu_metropolis_batch(ms, ds, tiling=(16, 1), jit=False, **kw)
    b = gpu_metropolis_batch(ms, ds, tiling=(16, 4, 4), jit=True, **kw)
    for fa, fb in zip(a.fits, b.fits):
        assert np.array_equal(fa.draws, fb.draws)
    from gpumetropolis_amd.diagnostics import summarize
    post = workloads.regression_posterior(wls[0].data)
    s = summarize(b.fits[0].draws)["params"][0]
    print(f"[quads] b0 {s['mean']:.5f} vs {post['mean'][0]:.5f} "
          f"mcse {s['mcse']:.2e}", flush=True)
    assert abs(s["mean"] - post["mean"][0]) <= 4 * s["mcse"]
