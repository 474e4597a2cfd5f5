"""FP64 batch evaluation on the GPU against the CPU oracle, through the C ABI:
the per-observation log-likelihood matrix (cpu_native.rs:371-391) and the
reverse-mode gradient of the summed data term (cpu_native.rs:148-329).

Both evaluate every operation in FP64 as the reference does, so the
tolerances are FP64 ones:
- pointwise: |gpu - oracle| <= 1e-12 * max(1, |oracle|) per entry;
- gradient: |gpu - oracle| <= 1e-11 * sum_i |d ll_i / d theta_k|, the sum of
  the absolute per-observation terms (the two paths add the terms in
  different orders).
"""
import numpy as np
import pytest

from gpumetropolis_amd import dsl, workloads

pytestmark = pytest.mark.gpu

PW_TOL = 1e-12
GRAD_TOL = 1e-11

CASES = {
    # name: (workload kwargs, point centre, point scale)
    "gaussian_mean": (dict(n=3000), [3.4], 0.05),
    "gamma": (dict(n=5000), [np.log(2.5), np.log(0.7)], 0.05),
    "mixture": (dict(n=4099), [-3.0, 3.0], 0.2),
    "logistic16": (dict(n=1500), [0.0] * 16, 0.3),
}


@pytest.fixture(params=[True, False], ids=["compiled", "interpreter"])
def eval_mode(request, hip):
    """Run the test on kernels compiled per program and on the interpreter."""
    hip.set_eval_jit(request.param)
    yield request.param
    hip.set_eval_jit(True)


def _points(rng, centre, scale, n):
    centre = np.asarray(centre, float)
    return centre + scale * rng.standard_normal((n, centre.size))


def _grad_scale(orc, code, consts, n_params, data, pts):
    """sum over observations of |d ll_i / d theta| at each point (oracle)."""
    s = np.zeros((len(pts), n_params))
    for row in data:
        s += np.abs(orc.grad_batch(code, consts, n_params, row[None, :],
                                   data.shape[1], 1, pts))
    return s


@pytest.mark.parametrize("name", list(CASES))
def test_pointwise_matches_fp64(hip, orc, eval_mode, name):
    kw, centre, scale = CASES[name]
    w = workloads.by_name(name, **kw)
    m = w.model
    pts = _points(np.random.default_rng(3), centre, scale, 9)
    print(f"[pointwise] {name} n={w.n_obs}", flush=True)
    got = hip.loglik_pointwise(m.loglik.code, m.loglik.consts, m.n_params,
                               w.data, w.n_cols, w.n_obs, pts)
    ref = orc.loglik_pointwise(m.loglik.code, m.loglik.consts, m.n_params,
                               w.data, w.n_cols, w.n_obs, pts)
    assert got.shape == (9, w.n_obs)
    err = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
    assert err.max() <= PW_TOL, (name, err.max())
    assert hip.last_eval_ms() > 0.0


@pytest.mark.parametrize("name", list(CASES))
def test_gradient_matches_fp64(hip, orc, eval_mode, name):
    kw, centre, scale = CASES[name]
    w = workloads.by_name(name, **kw)
    m = w.model
    pts = _points(np.random.default_rng(4), centre, scale, 5)
    print(f"[grad] {name} n={w.n_obs}", flush=True)
    got = hip.grad_batch(m.loglik.code, m.loglik.consts, m.n_params, w.data,
                         w.n_cols, w.n_obs, pts)
    ref = orc.grad_batch(m.loglik.code, m.loglik.consts, m.n_params, w.data,
                         w.n_cols, w.n_obs, pts)
    scale_k = _grad_scale(orc, m.loglik.code, m.loglik.consts, m.n_params,
                          w.data, pts)
    assert got.shape == (5, m.n_params)
    err = np.abs(got - ref) / scale_k
    assert err.max() <= GRAD_TOL, (name, err.max())


def test_gradient_closed_form(hip):
    """Gaussian mean: d/dmu sum -(y - mu)^2 / 8 = sum (y - mu) / 4."""
    w = workloads.gaussian_mean(n=10007)
    m = w.model
    mu = np.array([[3.1], [3.4], [3.9]])
    got = hip.grad_batch(m.loglik.code, m.loglik.consts, 1, w.data, 1,
                         w.n_obs, mu)[:, 0]
    y = w.data[:, 0]
    want = np.array([np.sum(y - u) / 4.0 for u in mu[:, 0]])
    assert np.allclose(got, want, rtol=1e-12, atol=1e-9 * np.abs(y).sum())


# Every opcode and the derivative branches: POW with a parameter exponent and
# with a base that is negative for some observations (at the integer
# exponent 2 the value is finite and the exponent's derivative takes only
# the positive bases, as the reference skips log of a non-positive base),
# DIV, SQRT, LGAMMA (digamma), NEG, EXP and LOG.  The third entry lists the
# point coordinates kept exactly at the centre.
OPCODE_FORMS = [
    ("~ x^a - y * b", [0.7, 1.3], []),
    ("~ (y - a)^b", [0.2, 2.0], [1]),
    ("~ sqrt(x) * a / (b + y * y) - lgamma(a * x)", [1.5, 0.8], []),
    ("~ -exp(a * y) + log(b * x)", [0.3, 1.1], []),
]


@pytest.mark.parametrize("n_obs", [1, 257, 4097])
@pytest.mark.parametrize("form", OPCODE_FORMS, ids=lambda f: f[0])
def test_opcodes_and_ragged_sizes(hip, orc, eval_mode, form, n_obs):
    formula, centre, exact = form
    m = dsl.gpum_model(formula, ["a", "b"], ["y", "x"])
    rng = np.random.default_rng(n_obs)
    data = np.column_stack([rng.normal(0.3, 1.0, n_obs),
                            rng.uniform(0.5, 3.0, n_obs)])
    pts = _points(rng, centre, 0.05, 4)
    pts[:, exact] = np.asarray(centre)[exact]
    print(f"[opcodes] {formula} n={n_obs}", flush=True)
    pw = hip.loglik_pointwise(m.loglik.code, m.loglik.consts, 2, data, 2,
                              n_obs, pts)
    pw_ref = orc.loglik_pointwise(m.loglik.code, m.loglik.consts, 2, data, 2,
                                  n_obs, pts)
    both_nan = np.isnan(pw) & np.isnan(pw_ref)
    err = np.abs(pw - pw_ref) / np.maximum(1.0, np.abs(pw_ref))
    assert np.all(both_nan | (err <= PW_TOL)), (formula, np.nanmax(err))
    g = hip.grad_batch(m.loglik.code, m.loglik.consts, 2, data, 2, n_obs, pts)
    g_ref = orc.grad_batch(m.loglik.code, m.loglik.consts, 2, data, 2, n_obs,
                           pts)
    finite = np.isfinite(g_ref)
    assert np.array_equal(np.isfinite(g), finite), formula
    scale_k = _grad_scale(orc, m.loglik.code, m.loglik.consts, 2, data, pts)
    err = np.abs(g - g_ref)[finite] / scale_k[finite]
    assert err.size == 0 or err.max() <= GRAD_TOL, (formula, err.max())


def test_zero_adjoint_rule(hip, orc):
    """A zero adjoint stops at its operation (cpu_native.rs:218-220): in
    exp(-exp(a x)) with exp(a x) = inf the value is 0 and the reference's
    gradient term is 0, where 0 * inf would give NaN.  Only the second
    8192-observation slice holds such rows, so the compiled gradient takes
    its exact path there alone; it must equal the interpreter bit for bit
    and the oracle within the FP64 tolerance."""
    m = dsl.gpum_model("~ exp(-exp(a * x)) + b * y", ["a", "b"], ["y", "x"])
    n_obs = 3 * 8192 + 5
    rng = np.random.default_rng(11)
    x = rng.uniform(0.1, 1.0, n_obs)
    x[8192 + rng.choice(8192, 7, replace=False)] = 900.0
    data = np.column_stack([rng.normal(0.3, 1.0, n_obs), x])
    pts = np.array([[1.0, 0.5], [1.2, -0.3], [0.9, 0.1]])
    print("[grad] zero-adjoint rule", flush=True)
    args = (m.loglik.code, m.loglik.consts, 2, data, 2, n_obs, pts)
    out = {}
    for jit in (True, False):
        hip.set_eval_jit(jit)
        out[jit] = hip.grad_batch(*args)
        assert hip.last_eval_compiled() == jit
    hip.set_eval_jit(True)
    assert np.all(np.isfinite(out[True]))
    assert np.array_equal(out[True], out[False])
    ref = orc.grad_batch(*args)
    scale_k = _grad_scale(orc, m.loglik.code, m.loglik.consts, 2, data, pts)
    assert np.max(np.abs(out[True] - ref) / scale_k) <= GRAD_TOL


@pytest.mark.parametrize("name", list(CASES))
def test_compiled_equals_interpreter(hip, name):
    """The kernels compiled per program run the interpreter's operations in
    its order: identical bits, for the matrix and the gradient."""
    kw, centre, scale = CASES[name]
    w = workloads.by_name(name, **kw)
    m = w.model
    pts = _points(np.random.default_rng(6), centre, scale, 7)
    args = (m.loglik.code, m.loglik.consts, m.n_params, w.data, w.n_cols,
            w.n_obs, pts)
    out = {}
    for jit in (True, False):
        hip.set_eval_jit(jit)
        pw = hip.loglik_pointwise(*args)
        c1 = hip.last_eval_compiled()
        g = hip.grad_batch(*args)
        c2 = hip.last_eval_compiled()
        assert c1 == c2 == jit
        out[jit] = (pw, g)
    hip.set_eval_jit(True)
    assert np.array_equal(out[True][0], out[False][0])
    assert np.array_equal(out[True][1], out[False][1])


def test_pointwise_over_device_chunks(hip, orc):
    """An output larger than one device chunk (256 MB): 36 points x 1e6
    observations; the first and last rows against the oracle, every row's
    sum against the summed log-likelihood of the same points."""
    w = workloads.gamma(n=1_000_000)
    m = w.model
    pts = _points(np.random.default_rng(8), [np.log(2.5), np.log(0.7)], 0.02, 36)
    print("[pointwise] gamma 36 x 1e6", flush=True)
    got = hip.loglik_pointwise(m.loglik.code, m.loglik.consts, 2, w.data, 1,
                               w.n_obs, pts)
    ends = pts[[0, -1]]
    ref = orc.loglik_pointwise(m.loglik.code, m.loglik.consts, 2, w.data, 1,
                               w.n_obs, ends)
    err = np.abs(got[[0, -1]] - ref) / np.maximum(1.0, np.abs(ref))
    assert err.max() <= PW_TOL
    sums = orc.loglik_batch(m.loglik.code, m.loglik.consts, 2, w.data, 1,
                            w.n_obs, pts)
    rel = np.abs(got.sum(axis=1) - sums) / np.abs(sums)
    assert rel.max() <= 1e-10


def test_empty_and_degenerate_inputs(hip):
    m = workloads.gamma(n=10).model
    data = np.ones((10, 1))
    pts = np.zeros((3, 2))
    # Empty program: zero per observation, zero gradient.
    assert np.array_equal(hip.loglik_pointwise([], [], 2, data, 1, 10, pts),
                          np.zeros((3, 10)))
    assert np.array_equal(hip.grad_batch([], [], 2, data, 1, 10, pts),
                          np.zeros((3, 2)))
    # No observations: an empty matrix and a zero gradient.
    assert hip.loglik_pointwise(m.loglik.code, m.loglik.consts, 2,
                                np.zeros((0, 1)), 1, 0, pts).shape == (3, 0)
    assert np.array_equal(hip.grad_batch(m.loglik.code, m.loglik.consts, 2,
                                         np.zeros((0, 1)), 1, 0, pts),
                          np.zeros((3, 2)))
    # No parameters: no points, as the reference.
    assert hip.loglik_pointwise([2, 0], [], 0, data, 1, 10, []).shape == (0, 10)


def test_invalid_programs_rejected(hip):
    data = np.ones((4, 1))
    pts = np.zeros((1, 1))
    with pytest.raises(hip.BackendError, match="unknown opcode"):
        hip.loglik_pointwise([99, 0], [], 1, data, 1, 4, pts)
    with pytest.raises(hip.BackendError, match="parameter index"):
        hip.grad_batch([1, 3], [], 1, data, 1, 4, pts)
    with pytest.raises(hip.BackendError, match="at most 16 parameters"):
        hip.grad_batch([1, 0], [], 17, data, 1, 4, np.zeros((1, 17)))
