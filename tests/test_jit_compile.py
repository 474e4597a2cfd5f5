"""The per-model compiled kernels build from the embedded device headers
(hiprtc for gfx950, no device needed): the generated evaluator and the
chain, points and batched kernel bodies of representative programs."""
import pytest

from gpumetropolis_amd import _native, workloads


@pytest.fixture(scope="module")
def lib():
    return _native.lib()


@pytest.mark.parametrize("name,cw,kinds", [
    ("gamma", 2, _native.JIT_KINDS),
    ("gamma", 4, ("chain_stream",)),
    ("logistic16", 1, ("chain_stream", "points_stream", "points_resident")),
    ("mixture", 1, ("chain_stream", "chain_resident", "pointwise64",
                    "grad64")),
    ("logistic16", 1, ("grad64",)),
])
def test_compiled_kernels_build(lib, name, cw, kinds):
    w = workloads.by_name(name, n=64)
    m = w.model
    ms = _native.jit_check(m.loglik.code, m.loglik.consts, m.prior.code,
                           m.prior.consts, m.n_params, w.n_cols, cw, kinds)
    assert ms > 0.0


def test_batched_kernel_builds_for_regressions(lib):
    w = workloads.regressions(n_models=1, n_min=100, n_max=100)[0]
    m = w.model
    _native.jit_check(m.loglik.code, m.loglik.consts, m.prior.code,
                      m.prior.consts, m.n_params, w.n_cols, 1,
                      ("batch", "chain_stream"))


def test_compile_errors_are_reported(lib):
    m = workloads.gamma(n=8).model
    with pytest.raises(_native.BackendError, match="chains_per_wave"):
        _native.jit_check(m.loglik.code, m.loglik.consts, m.prior.code,
                          m.prior.consts, 2, 1, 3)
    with pytest.raises(_native.BackendError, match="prior"):
        _native.jit_check(m.loglik.code, m.loglik.consts, [2, 0], [], 2, 1, 1)
