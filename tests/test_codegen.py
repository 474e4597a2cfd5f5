"""The CPU baseline's per-model compiled log-likelihood (oracle/codegen.py)
against the oracle's interpreter (orc_loglik, cpu_native.rs:110-117).

The generated C evaluates the same operations in the same order in f64,
so the sums must agree bit for bit on every configuration's formula, and
a sampler run through the compiled function must reproduce the
interpreter's draws exactly.
"""
import numpy as np
import pytest

from gpumetropolis_amd import workloads
from oracle import codegen


def _cases():
    yield "gaussian_mean", workloads.gaussian_mean(n=3001)
    yield "gamma", workloads.gamma(n=4099)
    yield "mixture", workloads.mixture(n=2049)
    yield "regression", workloads.regressions(n_models=1, n_min=1500,
                                              n_max=1600)[0]
    yield "logistic16", workloads.logistic(n=1025)


CASES = list(_cases())


def _points(name, n_params, rng):
    if name == "gamma":
        return rng.normal([np.log(2.5), np.log(0.7)], 0.1, size=(5, 2))
    return rng.normal(0.0, 0.7, size=(5, n_params))


@pytest.mark.parametrize("name,w", CASES, ids=[c[0] for c in CASES])
def test_compiled_equals_interpreter_bitwise(orc, name, w):
    m = w.model
    np_ = len(m.params)
    _, fn, _ = codegen.compile_loglik(m.loglik.code, m.loglik.consts)
    rng = np.random.default_rng(1)
    data = np.ascontiguousarray(w.data, dtype=np.float64)
    for p in _points(name, np_, rng):
        p = np.ascontiguousarray(p, dtype=np.float64)
        got = fn(p.ctypes.data_as(codegen.C.POINTER(codegen.C.c_double)),
                 data.ctypes.data_as(codegen.C.POINTER(codegen.C.c_double)),
                 w.n_obs, w.n_cols)
        want = orc.loglik_batch(m.loglik.code, m.loglik.consts, np_, data,
                                w.n_cols, w.n_obs, p)[0]
        assert got == want, (name, p, got, want)


def test_compiled_run_equals_interpreted_run(orc):
    w = workloads.gamma(n=2000)
    m = w.model
    _, _, addr = codegen.compile_loglik(m.loglik.code, m.loglik.consts)
    C = 6
    init = np.tile([np.log(2.5), np.log(0.7)], (C, 1))
    sd = np.full((C, 2), 0.02)
    args = (m.loglik.code, m.loglik.consts, m.prior.code, m.prior.consts, 2,
            w.data, 1, w.n_obs, init, sd, np.ones(C), 40, 9)
    a = orc.run(*args, n_threads=2)
    b = orc.run(*args, n_threads=2, loglik_fn=addr)
    np.testing.assert_array_equal(a.draws, b.draws)
    np.testing.assert_array_equal(a.last_log_post, b.last_log_post)
    assert 0.0 < a.accept_rate.mean() < 1.0
