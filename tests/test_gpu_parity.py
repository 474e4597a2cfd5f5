"""HIP backend vs the CPU oracle, through the C ABI.

Per-proposal log-likelihood within 1e-6 relative of the FP64 evaluation;
identical accept/reject sequences under the shared Philox stream (up to
certified ambiguous decisions); posterior moments within 3 MCSE of closed
forms.
"""
import math

import numpy as np
import pytest

from gpumetropolis_amd import dsl, workloads
from gpumetropolis_amd.diagnostics import ess, rhat
from tests.helpers import replay_compare

pytestmark = pytest.mark.gpu

REL_TOL = 1e-6


def _points(rng, center, scale, n):
    center = np.asarray(center, float)
    return center + scale * rng.standard_normal((n, center.size))


CASES = {
    # name: (workload factory kwargs, point centre, point scale)
    "gaussian_mean": (dict(n=20000), [3.4], 0.05),
    "gamma": (dict(n=50000), [np.log(2.5), np.log(0.7)], 0.05),
    "mixture": (dict(n=30000), [-3.0, 3.0], 0.2),
    "logistic16": (dict(n=20000), [0.0] * 16, 0.3),
}


@pytest.mark.parametrize("name", list(CASES))
def test_loglik_batch_matches_fp64(hip, orc, name):
    kw, centre, scale = CASES[name]
    w = workloads.by_name(name, **kw)
    m = w.model
    rng = np.random.default_rng(5)
    pts = _points(rng, centre, scale, 64)
    print(f"[loglik] {name} n={w.n_obs} cols={w.n_cols}", flush=True)
    got = hip.loglik_batch(m.loglik.code, m.loglik.consts, m.n_params, w.data,
                           w.n_cols, w.n_obs, pts)
    ref = orc.loglik_batch(m.loglik.code, m.loglik.consts, m.n_params, w.data,
                           w.n_cols, w.n_obs, pts)
    rel = np.abs(got - ref) / np.abs(ref)
    assert np.all(np.isfinite(got))
    assert rel.max() <= REL_TOL, (name, rel.max())


@pytest.mark.parametrize("precompute,fuse", [(0, 0), (1, 1)])
def test_lowering_switches_agree(hip, orc, precompute, fuse):
    w = workloads.gamma(n=20000)
    m = w.model
    pts = _points(np.random.default_rng(1), [np.log(2.5), np.log(0.7)], 0.1, 16)
    ref = orc.loglik_batch(m.loglik.code, m.loglik.consts, 2, w.data, 1,
                           w.n_obs, pts)
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, [], [], 2, w.data, 1,
                         w.n_obs, precompute_data=precompute, fuse=fuse)
    got = dm.loglik(pts)
    assert (np.abs(got - ref) / np.abs(ref)).max() <= REL_TOL


# Every summation form of the lowering (affine programs accumulated
# directly, trailing squares, sum tails of uniform operations) against the
# FP64 evaluation, with and without the fold, on sizes that end in a
# partial chunk.
FOLD_FORMS = [
    ("~ y * mu - 3 * x + 2 - s", [0.7, 1.3]),         # affine, constants in the tail
    ("~ mu * x - y + s", [1.1, -0.4]),                 # affine, SUB_D term
    ("~ 2 * (x - mu) - s", [0.2, 0.5]),                # affine, tail MUL then SUB
    ("~ -log(s) * 2 + 0 * y", [0.0, 1.7]),             # constant term only
    ("~ (y - mu)^2 * s + 4 - s", [0.3, 0.8]),          # square, tail MUL ADD SUB
    ("~ (exp(y - mu) + x) * s", [0.1, 0.9]),           # value, tail MUL
    ("~ log(1 + exp(y * mu)) - x * s", [0.4, 0.2]),    # no fold
]


@pytest.mark.parametrize("n_obs", [1025, 70001])
@pytest.mark.parametrize("form", FOLD_FORMS, ids=lambda f: f[0])
def test_summation_forms_match_fp64(hip, orc, form, n_obs):
    formula, centre = form
    m = dsl.gpum_model(formula, ["mu", "s"], ["y", "x"])
    rng = np.random.default_rng(n_obs)
    data = np.column_stack([rng.normal(0.5, 1.0, n_obs), rng.normal(-0.2, 1.5, n_obs)])
    data = data.astype(np.float32).astype(np.float64)
    pts = _points(rng, centre, 0.1, 8)
    pts[:, 1] = np.abs(pts[:, 1])
    ref = orc.loglik_batch(m.loglik.code, m.loglik.consts, 2, data, 2, n_obs, pts)
    print(f"[fold] {formula} n={n_obs}", flush=True)
    scale = np.abs(ref) + 1e-7 * n_obs
    for fuse in (1, 0):
        dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, [], [], 2, data, 2,
                             n_obs, fuse=fuse)
        got = dm.loglik(pts)
        assert np.max(np.abs(got - ref) / scale) <= REL_TOL, (fuse, got, ref)
        if fuse and dm.info()["vm_depth"] <= 1:
            # two chains per wave: the proposals' data terms
            r = dm.run(pts, np.full_like(pts, 0.01), np.ones(8), 3, 1,
                       trace=True, tiling=(16, 1, 2))
            props = r.trace_prop.reshape(-1, 2)
            ref2 = orc.loglik_batch(m.loglik.code, m.loglik.consts, 2, data, 2,
                                    n_obs, props)
            sc2 = np.abs(ref2) + 1e-7 * n_obs
            assert np.max(np.abs(r.trace_ll.ravel() - ref2) / sc2) <= REL_TOL
        dm.close()


# The FP32 per-observation log-gamma (csrc/device_vm.h): its argument mixes
# data and parameters, so it is neither hoisted nor precomputed and runs on
# the VM.  Data terms against the oracle's FP64 Lanczos log-gamma
# (cpu_native.rs:34-60), interpreted and compiled.  Arguments: y > 0 from a
# gamma law, x in two bands that keep negative arguments at least 0.1 away
# from the poles (the argument's own FP32 rounding is amplified by
# 1/distance there).
LGAMMA_FORMS = [
    ("~ lgamma(y + a) - b * y", [2.5, 0.7]),        # arguments in (2.5, 15)
    ("~ lgamma(y * y * b + a) - y", [0.1, 0.5]),    # from 0.1 across 0.5
    ("~ lgamma(x + a) * b + y", [0.0, 0.6]),        # negative: reflection
    ("~ lgamma(y * 4 + a) * b", [1.0, 0.5]),        # beyond 8: Stirling
]


@pytest.mark.parametrize("n_obs", [1025, 70001])
@pytest.mark.parametrize("form", LGAMMA_FORMS, ids=lambda f: f[0])
def test_lgamma_data_terms_match_fp64(hip, orc, form, n_obs):
    formula, centre = form
    m = dsl.gpum_model(formula, ["a", "b"], ["y", "x"])
    rng = np.random.default_rng(n_obs + 3)
    y = rng.gamma(2.0, 1.5, n_obs)
    x = rng.uniform(0.15, 0.85, n_obs) * np.where(rng.random(n_obs) < 0.5, -1, 1)
    x = x - np.where(rng.random(n_obs) < 0.3, 1.0, 0.0)     # (-1.85, -1.15) too
    data = np.column_stack([y, x]).astype(np.float32).astype(np.float64)
    pts = _points(rng, centre, 0.03, 8)
    ref = orc.loglik_batch(m.loglik.code, m.loglik.consts, 2, data, 2, n_obs, pts)
    assert np.all(np.isfinite(ref))
    scale = np.abs(ref) + 1e-7 * n_obs
    for jit in (False, True):
        print(f"[lgamma] {formula} n={n_obs} jit={jit}", flush=True)
        dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, [], [], 2, data, 2,
                             n_obs, jit=jit)
        assert dm.info()["n_vm_instr"] > 0
        got = dm.loglik(pts)
        dm.close()
        assert np.max(np.abs(got - ref) / scale) <= REL_TOL, (jit, got, ref)


# log(c1 exp(A1) + c2 exp(A2)) is evaluated as a log-sum-exp (csrc/lower.cpp,
# VM_LSE_S): it must match FP64 near the data as the plain FP32 expression
# did, and also far from it, where both exponentials underflow in FP32 (the
# plain expression gives -inf there while the FP64 oracle stays finite).
LSE_FORMS = [
    ("~ log(0.5 * exp(-((y - a)^2) / 2) + 0.5 * exp(-((y - b)^2) / 2))",
     [[-0.3, 1.2], [0.5, 0.6], [20.0, 25.0], [-30.0, 14.0]]),
    ("~ log(0.3 * exp(x - a) + exp(y * b) * 0.7)",
     [[0.2, 0.4], [-1.0, 2.0], [120.0, -60.0], [-95.0, 40.0]]),
    ("~ log(exp(y - a) + exp(x * b))",
     [[0.1, 0.3], [130.0, -70.0], [-99.0, 45.0]]),
]


@pytest.mark.parametrize("form", LSE_FORMS, ids=lambda f: f[0])
def test_log_sum_exp_matches_fp64_near_and_far(hip, orc, form):
    formula, pts = form
    m = dsl.gpum_model(formula, ["a", "b"], ["y", "x"])
    rng = np.random.default_rng(17)
    n = 30001
    data = np.column_stack([rng.normal(0.5, 1.0, n), rng.normal(-0.2, 1.5, n)])
    data = data.astype(np.float32).astype(np.float64)
    pts = np.asarray(pts, float)
    ref = orc.loglik_batch(m.loglik.code, m.loglik.consts, 2, data, 2, n, pts)
    assert np.all(np.isfinite(ref))
    print(f"[lse] {formula}", flush=True)
    for jit in (False, True):
        dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, [], [], 2, data, 2,
                             n, jit=jit)
        got = dm.loglik(pts)
        dm.close()
        rel = np.abs(got - ref) / (np.abs(ref) + 1e-7 * n)
        assert np.all(np.isfinite(got)), (jit, got)
        assert rel.max() <= REL_TOL, (jit, got, ref)


NEG_BINOMIAL = ("~ lgamma(y + exp(lr)) - lgamma(exp(lr)) - lgamma(y + 1)"
                " - exp(lr) * log(1 + exp(-lp)) - y * log(1 + exp(lp))")


@pytest.mark.parametrize("jit", [False, True])
def test_replay_negative_binomial_lgamma(hip, orc, jit):
    """Negative-binomial counts (size r = exp(lr), p = logistic(lp)): the
    data-dependent lgamma(y + r) runs per observation on the device; the
    replay must keep every decision."""
    rng = np.random.default_rng(21)
    n = 20000
    y = rng.negative_binomial(3.0, 0.4, n).astype(np.float64)
    m = dsl.gpum_model(NEG_BINOMIAL, ["lr", "lp"], ["y"],
                       prior="~ -0.5 * (lr / 10)^2 - 0.5 * (lp / 10)^2")
    w = workloads.Workload("negbin", m, y.reshape(-1, 1), 16, 120)
    C, T, P = 16, 120, 2
    init = np.array([np.log(3.0), np.log(0.4 / 0.6)]) + 0.02 * rng.standard_normal((C, P))
    sd = np.full((C, P), 0.02)
    temps = np.ones(C)
    print(f"[replay] negative binomial jit={jit}", flush=True)
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                         m.prior.consts, P, w.data, 1, n, jit=jit)
    info = dm.info()
    g = dm.run(init, sd, temps, T, 13, trace=True)
    dm.close()
    o = orc.run(m.loglik.code, m.loglik.consts, m.prior.code, m.prior.consts,
                P, w.data, 1, n, init, sd, temps, T, 13,
                stream=orc.STREAM_PHILOX, trace=True)
    rep = replay_compare(g, o, orc, 13, temps, 0, P,
                         init_lp_gpu=g.init_log_post,
                         init_lp_orc=_init_lp_orc(orc, w, init))
    print(rep, info, flush=True)
    assert info["n_vm_instr"] > 0
    assert rep.uncertified_flips == 0
    assert rep.max_plp_rel <= REL_TOL
    assert rep.steps_compared >= C * T // 2


@pytest.mark.parametrize("n_obs", [1, 7, 255, 1023, 1025, 4097, 70001])
def test_ragged_sizes(hip, orc, n_obs):
    rng = np.random.default_rng(n_obs)
    y = rng.normal(1.0, 2.0, n_obs)
    m = dsl.gpum_model("~ -((y - mu)^2) / 2 - log(s) + 0 * y", ["mu", "s"], ["y"])
    pts = np.array([[0.5, 2.0], [1.0, 1.5], [-1.0, 3.0]])
    got = hip.loglik_batch(m.loglik.code, m.loglik.consts, 2, y, 1, n_obs, pts)
    ref = orc.loglik_batch(m.loglik.code, m.loglik.consts, 2, y, 1, n_obs, pts)
    assert np.allclose(got, ref, rtol=REL_TOL, atol=1e-9)


def _run_both(hip, orc, w, init, sd, temps, n_iter, seed, mode=0, gamma=0.0,
              de_noise=0.0, L=None, tiling=None):
    m = w.model
    args = (m.loglik.code, m.loglik.consts, m.prior.code, m.prior.consts,
            m.n_params, w.data, w.n_cols, w.n_obs, init, sd, temps, n_iter,
            seed)
    g = hip.run_chains(*args, proposal_mode=mode, gamma=gamma,
                       de_noise=de_noise, proposal_l=L, trace=True,
                       final_state=True, tiling=tiling)
    o = orc.run(*args, proposal_mode=mode, gamma=gamma, de_noise=de_noise,
                proposal_l=L, stream=orc.STREAM_PHILOX, trace=True)
    return g, o


def _init_lp_orc(orc, w, init):
    m = w.model
    ll = orc.loglik_batch(m.loglik.code, m.loglik.consts, m.n_params, w.data,
                          w.n_cols, w.n_obs, init)
    pr = np.array([orc.vm_eval(m.prior.code, m.prior.consts, p, [])
                   for p in np.asarray(init).reshape(-1, m.n_params)])
    return ll + pr


@pytest.mark.parametrize("tiling", [None, (1, 16), (8, 1)])
def test_replay_gaussian_mean_mode0(hip, orc, tiling):
    w = workloads.gaussian_mean(n=20000)
    C, T = 8, 300
    init = np.linspace(3.3, 3.5, C).reshape(C, 1)
    sd = np.full((C, 1), 0.02)
    temps = np.ones(C)
    print(f"[replay] gaussian_mean mode0 tiling={tiling}", flush=True)
    g, o = _run_both(hip, orc, w, init, sd, temps, T, seed=7, tiling=tiling)
    rep = replay_compare(g, o, orc, 7, temps, 0, 1,
                         init_lp_gpu=g.init_log_post,
                         init_lp_orc=_init_lp_orc(orc, w, init))
    print(rep, flush=True)
    assert rep.uncertified_flips == 0
    assert rep.max_plp_rel <= REL_TOL
    assert rep.chains_diverged <= 1
    assert rep.steps_compared >= C * T // 2


def test_replay_gamma_mode2_tempered(hip, orc):
    w = workloads.gamma(n=30000)
    C, T, P = 16, 150, 2
    rng = np.random.default_rng(3)
    init = np.array([np.log(2.5), np.log(0.7)]) + 0.01 * rng.standard_normal((C, P))
    temps = np.geomspace(1.0, 4.0, C)
    Lc = np.array([[0.01, 0.0], [0.008, 0.006]])
    L = np.tile(Lc.ravel(), C)
    sd = np.full((C, P), 0.01)
    print("[replay] gamma mode2", flush=True)
    g, o = _run_both(hip, orc, w, init, sd, temps, T, seed=11, mode=2, L=L)
    rep = replay_compare(g, o, orc, 11, temps, 2, P,
                         init_lp_gpu=g.init_log_post,
                         init_lp_orc=_init_lp_orc(orc, w, init))
    print(rep, flush=True)
    assert rep.uncertified_flips == 0
    assert rep.max_plp_rel <= REL_TOL
    assert rep.steps_compared >= C * T // 2


def test_replay_mixture_mode1_de(hip, orc):
    w = workloads.mixture(n=20000)
    C, T, P = 12, 120, 2
    rng = np.random.default_rng(4)
    init = np.array([-3.0, 3.0]) + 0.05 * rng.standard_normal((C, P))
    temps = np.ones(C)
    sd = np.full((C, P), 0.01)
    print("[replay] mixture mode1", flush=True)
    g, o = _run_both(hip, orc, w, init, sd, temps, T, seed=5, mode=1,
                     gamma=2.38 / math.sqrt(4), de_noise=1e-3)
    rep = replay_compare(g, o, orc, 5, temps, 1, P,
                         init_lp_gpu=g.init_log_post,
                         init_lp_orc=_init_lp_orc(orc, w, init))
    print(rep, flush=True)
    assert rep.uncertified_flips == 0
    assert rep.max_plp_rel <= REL_TOL


def test_gaussian_mean_closed_form(hip):
    """Posterior N(mean(y), 4/n): mean within 3 MCSE, sd within 10%."""
    w = workloads.gaussian_mean(n=20000)
    m = w.model
    C, T = 8, 3000
    post_m, post_s = w.meta["post_mean"], w.meta["post_sd"]
    init = post_m + post_s * np.linspace(-2, 2, C).reshape(C, 1)
    sd = np.full((C, 1), 2.4 * post_s)
    g = hip.run_chains(m.loglik.code, m.loglik.consts, m.prior.code,
                       m.prior.consts, 1, w.data, 1, w.n_obs, init, sd,
                       np.ones(C), T, 99)
    kept = g.draws[T // 2:, :, 0].astype(np.float64)
    e = ess(kept)
    mcse = kept.std(ddof=1) / math.sqrt(e)
    print(f"mean={kept.mean():.6f} want={post_m:.6f} mcse={mcse:.2e} "
          f"sd={kept.std():.5f} want={post_s:.5f} ess={e:.0f} rhat={rhat(kept):.4f}")
    assert abs(kept.mean() - post_m) <= 3 * mcse
    assert abs(kept.std(ddof=1) / post_s - 1) < 0.1
    assert rhat(kept) < 1.01


def test_nonfinite_proposals_rejected(hip):
    """test-numerical.R:4-19: log(mu) - mu with mu <= 0 proposals."""
    m = dsl.gpum_model("~ log(mu) - mu + 0 * y", ["mu"], ["y"])
    C, T = 4, 3000
    g = hip.run_chains(m.loglik.code, m.loglik.consts, [], [], 1, [0.0], 1, 1,
                       np.full((C, 1), 1.5), np.full((C, 1), 0.4), np.ones(C),
                       T, 1)
    d = g.draws[:, :, 0]
    assert np.all(np.isfinite(d)) and d.min() > 0
    assert abs(d[T // 2:].mean() - 2.0) < 0.4


def test_nonfinite_start_completes(hip):
    """test-numerical.R:31-39: a non-finite starting log-density."""
    m = dsl.gpum_model("~ log(mu) - mu + 0 * y", ["mu"], ["y"])
    g = hip.run_chains(m.loglik.code, m.loglik.consts, [], [], 1, [0.0], 1, 1,
                       np.full((2, 1), -1.0), np.full((2, 1), 0.3),
                       np.ones(2), 500, 1)
    assert g.draws.shape == (500, 2, 1)
    assert np.all(np.isfinite(g.draws))


def test_same_seed_identical(hip):
    """test-api.R:30-39: the same seed reproduces the draws exactly."""
    w = workloads.gaussian_mean(n=5000)
    m = w.model
    args = (m.loglik.code, m.loglik.consts, [], [], 1, w.data, 1, w.n_obs,
            np.full((4, 1), 3.0), np.full((4, 1), 0.05), np.ones(4), 200, 3)
    a = hip.run_chains(*args)
    b = hip.run_chains(*args)
    assert np.array_equal(a.draws, b.draws)
    assert np.array_equal(a.last_log_post, b.last_log_post)


def test_errors_are_reported(hip):
    m = dsl.gpum_model("~ -((y - mu)^2) / 2", ["mu"], ["y"])
    with pytest.raises(hip.BackendError, match="proposal_mode 3"):
        hip.run_chains(m.loglik.code, m.loglik.consts, [], [], 1, [1.0], 1, 1,
                       [[0.0]], [[0.1]], [1.0], 10, 1, proposal_mode=3,
                       batches={"swap_every": 5, "swap_group": 1})
    with pytest.raises(hip.BackendError, match="opcode"):
        hip.run_chains([99, 0], [], [], [], 1, [1.0], 1, 1, [[0.0]], [[0.1]],
                       [1.0], 10, 1)
    with pytest.raises(hip.BackendError, match="underflow"):
        hip.run_chains([3, 0], [], [], [], 1, [1.0], 1, 1, [[0.0]], [[0.1]],
                       [1.0], 10, 1)
    with pytest.raises(hip.BackendError, match="n_params"):
        hip.run_chains(m.loglik.code, m.loglik.consts, [], [], 17, [1.0], 1, 1,
                       np.zeros(17), np.ones(17), [1.0], 10, 1)
