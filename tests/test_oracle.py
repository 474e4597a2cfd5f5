"""Pins the CPU oracle before it is trusted as the checker.

Each oracle function is checked against something that does not retype its
own formula: published known-answer vectors, an independent library
(rocRAND, scipy, math.lgamma), closed forms, invariants, or values read from
the reference's source text (tests/golden/reference_constants.json).
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import special, stats

from gpumetropolis_amd import dsl

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _json(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ------------------------------------------------------------------ RNG ---

def test_philox_random123_kat(orc):
    # Known-answer vectors of Philox4x32-10 published with Random123
    # (kat_vectors: zeros, all-ones, pi digits).
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2,
         [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
         [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(orc.philox(ctr, key)) == want


def test_philox_matches_rocrand(orc):
    for case in _json("philox_rocrand.json"):
        assert list(orc.philox(case["ctr"], case["key"])) == case["out"], case


def test_triple32_matches_reference_constants(orc):
    g = _json("reference_constants.json")
    for seed, outs in g["triple32_outputs"].items():
        s = int(seed)
        got = [orc.triple32((s + k) & 0xFFFFFFFF) for k in range(16)]
        assert got == outs


def test_triple32_is_a_bijection_sample(orc):
    # xorshift and odd multipliers are invertible: no collisions on a block.
    xs = np.arange(0, 1 << 16, dtype=np.uint64) * 65537
    ys = {orc.triple32(int(x) & 0xFFFFFFFF) for x in xs}
    assert len(ys) == len(xs)


def test_uniform_open_interval_and_mean(orc):
    u = np.array([orc.uniform_philox(3, 0, 0, t, 1, s)
                  for t in range(200) for s in range(9)])
    assert u.min() > 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.02
    # the map (h + 0.5) / 2^32 of the reference at the extremes
    assert (0 + 0.5) / 2 ** 32 > 0 and (0xFFFFFFFF + 0.5) / 2 ** 32 < 1


# ------------------------------------------------------------- lgamma ---

@pytest.mark.parametrize("x", [0.5, 1.0, 1.5, 2.0, 2.5, 3.7, 10.0, 55.5,
                               1234.5, 0.1, 0.01, -0.5, -2.3])
def test_lgamma_vs_library(orc, x):
    assert math.isclose(orc.lgamma(x), math.lgamma(x), rel_tol=1e-12,
                        abs_tol=1e-13)


def test_lgamma_coefficients_from_reference(orc):
    g = _json("reference_constants.json")
    c = g["lanczos_coeffs"]
    assert len(c) == 9
    # Independent evaluation of the Lanczos sum with the coefficients read
    # from the reference text.
    for x in [0.75, 2.5, 9.0]:
        z = x - 1.0
        a = c[0] + sum(c[i] / (z + i) for i in range(1, 9))
        t = z + 7.5
        want = 0.5 * math.log(2 * math.pi) + (z + 0.5) * math.log(t) - t + math.log(a)
        assert math.isclose(orc.lgamma(x), want, rel_tol=1e-14)


@pytest.mark.parametrize("x", [0.3, 1.0, 2.5, 7.0, 30.0])
def test_digamma_vs_library(orc, x):
    # The reference's series stops at the x^-6 term after shifting x >= 6;
    # the first omitted term, 1/(240 x^8), bounds the error by 2.5e-9.
    assert math.isclose(orc.digamma(x), special.digamma(x), rel_tol=0,
                        abs_tol=2.6e-9)


# ---------------------------------------------------------- interpreter ---

def test_gamma_opcode_kat(orc):
    """test-catalog.R:1-16: Sum dgamma(y, 2.5, 1.3, log=TRUE) to 1e-10, and
    the gradient against digamma to 1e-6."""
    m = dsl.gpum_model("~ a * log(b) - lgamma(a) + (a - 1) * log(y) - b * y",
                       ["a", "b"], ["y"])
    y = np.array([0.5, 1.2, 2.0, 3.5])
    pt = np.array([2.5, 1.3])
    val = orc.loglik_batch(m.loglik.code, m.loglik.consts, 2, y, 1, 4, pt)[0]
    want = stats.gamma.logpdf(y, a=2.5, scale=1 / 1.3).sum()
    assert abs(val - want) < 1e-10
    g = orc.grad_batch(m.loglik.code, m.loglik.consts, 2, y, 1, 4, pt)[0]
    ga = 4 * math.log(1.3) - 4 * special.digamma(2.5) + np.log(y).sum()
    gb = 4 * 2.5 / 1.3 - y.sum()
    assert np.allclose(g, [ga, gb], rtol=1e-6)


def test_gaussian_kernel_kat(orc):
    """test-kernel.R:4-21: -0.5/sigma^2 sum (d - mu)^2."""
    m = dsl.gpum_model("~ -0.5 * (y - mu)^2", ["mu"], ["y"])
    v = orc.loglik_batch(m.loglik.code, m.loglik.consts, 1, [0, 0, 0], 1, 3,
                         [0.0, 1.0, 2.0])
    assert list(v) == [0.0, -1.5, -6.0]
    m2 = dsl.gpum_model("~ -0.5 * (y - mu)^2 / 4", ["mu"], ["y"])
    assert orc.loglik_batch(m2.loglik.code, m2.loglik.consts, 1, [0, 0], 1, 2,
                            [2.0])[0] == -1.0
    m1 = dsl.gpum_model("~ -0.5 * (y - mu)^2 / 1", ["mu"], ["y"])
    assert orc.loglik_batch(m1.loglik.code, m1.loglik.consts, 1, [0, 0], 1, 2,
                            [2.0])[0] == -4.0


def test_every_opcode_against_numpy(orc):
    """Brute force on a tiny data set: each opcode's arithmetic against numpy
    evaluated independently in float64."""
    y = np.array([0.3, 1.7, 2.2, 4.0])
    x = np.array([1.5, -0.5, 2.0, 0.25])
    data = np.stack([y, x], 1).ravel()
    a, b = 0.7, 1.9
    forms = {
        "~ y + a": y + a, "~ y - a": y - a, "~ y * b": y * b,
        "~ y / b": y / b, "~ -x": -x, "~ exp(x)": np.exp(x),
        "~ log(y)": np.log(y), "~ sqrt(y)": np.sqrt(y), "~ y ^ b": y ** b,
        "~ x ^ 3": x ** 3, "~ lgamma(y + a)": special.gammaln(y + a),
        "~ (y - x) * (y + x) / (a + b)": (y - x) * (y + x) / (a + b),
    }
    for f, want in forms.items():
        m = dsl.gpum_model(f, ["a", "b"], ["y", "x"])
        got = orc.loglik_pointwise(m.loglik.code, m.loglik.consts, 2, data, 2,
                                   4, [a, b])[0]
        assert np.allclose(got, want, rtol=1e-13, atol=1e-14), f
        s = orc.loglik_batch(m.loglik.code, m.loglik.consts, 2, data, 2, 4,
                             [a, b])[0]
        acc = 0.0
        for v in want:  # sequential observation order
            acc += v
        assert math.isclose(s, acc, rel_tol=1e-13, abs_tol=1e-14), f


def test_empty_program_is_zero(orc):
    assert orc.loglik_batch([], [], 1, [1.0, 2.0], 1, 2, [0.5])[0] == 0.0
    assert orc.vm_eval([], [], [0.5], []) == 0.0


def test_gradient_vs_finite_difference(orc):
    """test-mala.R:1-33: tape gradient against central differences."""
    m = dsl.gpum_model("~ log(b) * exp(-a * y) - sqrt(b) / (1 + y^2) + lgamma(b + y)",
                       ["a", "b"], ["y"])
    y = np.array([0.2, 0.9, 1.7])
    p = np.array([0.4, 2.2])
    g = orc.grad_batch(m.loglik.code, m.loglik.consts, 2, y, 1, 3, p)[0]
    h = 1e-6
    for j in range(2):
        e = np.zeros(2)
        e[j] = h
        fp = orc.loglik_batch(m.loglik.code, m.loglik.consts, 2, y, 1, 3, p + e)[0]
        fm = orc.loglik_batch(m.loglik.code, m.loglik.consts, 2, y, 1, 3, p - e)[0]
        assert math.isclose(g[j], (fp - fm) / (2 * h), rel_tol=1e-6, abs_tol=1e-8)


# -------------------------------------------------------------- sampler ---

def _gm(n=2000, seed=1):
    rng = np.random.default_rng(seed)
    y = rng.normal(3.4, 2.0, n)
    return dsl.gpum_model("~ -((y - mu)^2) / 8", ["mu"], ["y"]), y


def test_sampler_same_seed_identical(orc):
    m, y = _gm()
    args = (m.loglik.code, m.loglik.consts, [], [], 1, y, 1, y.size,
            np.full((4, 1), 3.0), np.full((4, 1), 0.1), np.ones(4), 300, 5)
    for stream in (orc.STREAM_T32, orc.STREAM_PHILOX):
        a = orc.run(*args, stream=stream)
        b = orc.run(*args, stream=stream)
        assert np.array_equal(a.draws, b.draws)
        assert np.all((a.accept_rate >= 0) & (a.accept_rate <= 1))


def test_mode2_diagonal_equals_mode0(orc):
    m, y = _gm()
    C = 3
    sd = np.array([[0.05], [0.1], [0.2]])
    L = sd.reshape(C, 1, 1)
    base = (m.loglik.code, m.loglik.consts, [], [], 1, y, 1, y.size,
            np.full((C, 1), 3.0), sd, np.ones(C), 200, 9)
    a = orc.run(*base, proposal_mode=0)
    b = orc.run(*base, proposal_mode=2, proposal_l=L)
    assert np.array_equal(a.draws, b.draws)


def test_trace_consistency(orc):
    """Draws follow the decisions: state changes exactly when accepted, and
    the recorded state equals the accepted proposal."""
    m, y = _gm()
    C, T = 4, 150
    r = orc.run(m.loglik.code, m.loglik.consts, [], [], 1, y, 1, y.size,
                np.full((C, 1), 3.0), np.full((C, 1), 0.05), np.ones(C), T, 2,
                trace=True)
    for c in range(C):
        for t in range(T):
            if r.trace_accept[t, c]:
                assert r.draws[t, c, 0] == np.float32(r.trace_prop[t, c, 0])
            elif t > 0:
                assert r.draws[t, c, 0] == r.draws[t - 1, c, 0]
        assert math.isclose(r.accept_rate[c], r.trace_accept[:, c].mean(),
                            rel_tol=1e-6)


def test_layout_column_major(orc):
    """draws[t + n_iter * (c + n_chains * j)]: a target that never accepts
    keeps every chain at its own init, so the layout is readable."""
    code = [0, 0]  # a constant log-likelihood: every move is accepted
    C, T = 3, 5
    init = np.array([[1.0, 10.0], [2.0, 20.0], [3.0, 30.0]])
    # proposal_sd 0 keeps the state fixed whatever the decision
    r = orc.run(code, [0.0], [], [], 2, [0.0], 1, 1, init, np.zeros((C, 2)),
                np.ones(C), T, 1)
    for c in range(C):
        for j in range(2):
            assert np.all(r.draws[:, c, j] == init[c, j])


def test_closed_form_posterior_oracle(orc):
    """R/diagnostics.R:246-257 analogue: N(mean(y), 4/n) within 3 MCSE."""
    from gpumetropolis_amd.diagnostics import ess
    m, y = _gm(n=4000, seed=3)
    C, T = 4, 4000
    pm, ps = y.mean(), math.sqrt(4 / y.size)
    r = orc.run(m.loglik.code, m.loglik.consts, [], [], 1, y, 1, y.size,
                np.full((C, 1), pm), np.full((C, 1), 2.4 * ps), np.ones(C), T, 4)
    kept = r.draws[T // 2:, :, 0].astype(float)
    e = ess(kept)
    assert abs(kept.mean() - pm) <= 3 * kept.std() / math.sqrt(e)
    assert abs(kept.std() / ps - 1) < 0.1


def test_nonfinite_rejected_oracle(orc):
    m = dsl.gpum_model("~ log(mu) - mu + 0 * y", ["mu"], ["y"])
    r = orc.run(m.loglik.code, m.loglik.consts, [], [], 1, [0.0], 1, 1,
                [[1.5]], [[0.4]], [1.0], 3000, 1)
    d = r.draws[:, 0, 0]
    assert np.all(np.isfinite(d)) and d.min() > 0


def test_de_sync_pair_is_the_two_other_chains(orc):
    """Path B with 3 chains and no jitter (cpu_native.rs:455-471): a != c
    and b not in {c, a} leave exactly the two other chains, so each
    proposal step is +-g (x_o1 - x_o2) of the previous generation, with g
    1 or gamma.  A pair drawn with self, or path A's rule, breaks it."""
    m, y = _gm(n=200)
    C, T, gam = 3, 40, 0.7
    init = np.array([[2.0], [3.1], [4.5]])
    r = orc.run(m.loglik.code, m.loglik.consts, [], [], 1, y, 1, y.size, init,
                np.full((C, 1), 0.1), np.ones(C), T, 9, proposal_mode=1,
                gamma=gam, de_noise=0.0, de_sync=True, trace=True)
    prev = init[:, 0].copy()
    for t in range(T):
        for c in range(C):
            o1, o2 = [k for k in range(C) if k != c]
            step = abs(r.trace_prop[t, c, 0] - prev[c])
            d = abs(prev[o1] - prev[o2])
            assert min(abs(step - d), abs(step - gam * d)) <= 1e-12 * (1 + d), (t, c)
        # the generation's states: accepted proposals, else unchanged
        prev = np.where(r.trace_accept[t], r.trace_prop[t, :, 0], prev)


def test_de_sync_excludes_self(orc):
    m, y = _gm(n=500)
    C = 6
    r = orc.run(m.loglik.code, m.loglik.consts, [], [], 1, y, 1, y.size,
                np.linspace(2, 5, C).reshape(C, 1), np.full((C, 1), 0.1),
                np.ones(C), 400, 3, proposal_mode=1, gamma=1.68, de_noise=1e-3,
                de_sync=True)
    assert np.all(np.isfinite(r.draws))
    assert abs(r.draws[200:].mean() - y.mean()) < 0.2
