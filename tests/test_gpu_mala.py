"""MALA (proposal_mode 3, cpu_native.rs:573-724) on the GPU, through the C
ABI, against the oracle driven by the same Philox uniforms on the same
(FP32-rounded) data.  The log-posterior and its gradient are FP64 on both
sides (reverse mode over the bytecode), so the proposals agree to 1e-10
relative, the proposal log-posterior to 1e-10 relative, and every accept
decision is identical; the posterior mean matches the closed form within 3
MCSE."""
import numpy as np
import pytest

from gpumetropolis_amd import workloads
from gpumetropolis_amd.diagnostics import summarize
from tests.helpers import replay_compare

pytestmark = pytest.mark.gpu


def _chol_factors(rng, C, d, scale):
    """Per-chain lower-triangular factors with a positive diagonal."""
    L = np.zeros((C, d, d))
    for c in range(C):
        A = rng.standard_normal((d, d)) * 0.3
        M = A @ A.T + np.eye(d)
        L[c] = scale * np.linalg.cholesky(M)
    return L.reshape(-1)


@pytest.mark.parametrize("name,centre,sd,chol", [
    ("gaussian_mean", [3.4], 0.05, False),
    ("gamma", [np.log(2.5), np.log(0.7)], 0.01, False),
    ("gamma", [np.log(2.5), np.log(0.7)], 0.01, True),
    ("logistic16", [0.0] * 16, 0.02, True),
])
@pytest.mark.parametrize("compiled", [True, False], ids=["compiled", "interpreter"])
def test_mala_matches_oracle_draw_for_draw(hip, orc, name, centre, sd, chol,
                                           compiled):
    w = workloads.by_name(name, n=3000)
    m = w.model
    d = m.n_params
    # The backend evaluates the model's FP32 columns; both sides get them.
    data = w.data.astype(np.float32).astype(np.float64)
    C, T, seed = 8, 40, 17
    rng = np.random.default_rng(4)
    init = np.asarray(centre) + sd * rng.standard_normal((C, d))
    sdm = np.full((C, d), sd)
    L = _chol_factors(rng, C, d, sd) if chol else None
    temps = np.ones(C)
    print(f"[mala] {name} chol={chol} C={C} T={T}", flush=True)
    args = (m.loglik.code, m.loglik.consts, m.prior.code, m.prior.consts, d,
            data, w.n_cols, w.n_obs, init, sdm, temps, T, seed)
    hip.set_eval_jit(compiled)
    try:
        g = hip.run_chains(*args, proposal_mode=3, proposal_l=L, trace=True)
    finally:
        hip.set_eval_jit(True)
    o = orc.run(*args, proposal_mode=3, proposal_l=L, trace=True)
    rep = replay_compare(g, o, orc, seed, temps, 3, d, prop_rtol=1e-10)
    assert rep.chains_diverged == 0, rep
    assert rep.max_plp_rel <= 1e-10, rep
    assert np.array_equal(g.trace_accept, o.trace_accept)
    assert np.allclose(g.draws, o.draws, rtol=1e-6, atol=0.0)
    assert 0.0 < g.accept_rate.mean() <= 1.0


def test_mala_stepwise_equals_run_chains(hip):
    """The resident model's stepwise entry (langevin) and run_chains with
    proposal_mode 3 are one computation: identical draws."""
    w = workloads.gamma(n=2000)
    m = w.model
    C, T = 6, 30
    init = np.tile([np.log(2.5), np.log(0.7)], (C, 1))
    sd = np.full((C, 2), 0.02)
    a = hip.run_chains(m.loglik.code, m.loglik.consts, m.prior.code,
                       m.prior.consts, 2, w.data, 1, w.n_obs, init, sd,
                       np.ones(C), T, 8, proposal_mode=3)
    dm = hip.DeviceModel(m.loglik.code, m.loglik.consts, m.prior.code,
                         m.prior.consts, 2, w.data, 1, w.n_obs)
    b = dm.run_da(None, init, np.ones(C), T, 8, proposal_sd=sd, langevin=True)
    dm.close()
    assert np.array_equal(a.draws, b.draws)
    assert np.array_equal(a.last_log_post, b.last_log_post)


def test_mala_posterior_closed_form(hip):
    """Gaussian mean with known variance and a flat prior: the posterior
    mean is the sample mean.  Adaptive MALA (gpu_metropolis, method
    "mala") lands within 3 MCSE of it."""
    from gpumetropolis_amd.api import gpu_metropolis
    w = workloads.gaussian_mean(n=4000)
    y = w.data[:, 0]
    fit = gpu_metropolis(w.model, w.data, init=np.linspace(3.0, 3.8, 8),
                         n_iter=3000, warmup=1000, method="mala",
                         proposal_sd=0.05, seed=5, jit=False)
    s = summarize(fit.draws)["params"][0]
    print(f"[mala] mean {s['mean']:.5f} vs {y.mean():.5f} mcse {s['mcse']:.2e} "
          f"accept {fit.accept_rate.mean():.3f}")
    assert fit.method == "mala"
    assert abs(s["mean"] - y.mean()) <= 3 * s["mcse"]
    assert 0.3 < fit.accept_rate.mean() < 0.95
