"""The DSL compiler reproduces the reference's bytecode (R/dsl.R)."""
import pytest

from gpumetropolis_amd import dsl, workloads

P = {"PC": 0, "PP": 1, "PD": 2, "ADD": 3, "SUB": 4, "MUL": 5, "DIV": 6,
     "NEG": 7, "EXP": 8, "LOG": 9, "SQRT": 10, "POW": 11, "LGAMMA": 12}


def listing(*ins):
    """("PC", 0), "SUB", ... -> flat code."""
    out = []
    for i in ins:
        if isinstance(i, tuple):
            out += [P[i[0]], i[1]]
        else:
            out += [P[i], 0]
    return out


def test_known_expression():
    # test-dsl.R:3-11
    prog = dsl.compile_expr("~ mu + 2", ["mu"], [])
    assert prog.code == [1, 0, 0, 0, 3, 0]
    assert prog.consts == [2.0]


def test_const_pool_dedupe():
    # test-dsl.R:13-19
    assert dsl.compile_expr("~ 2 * mu + 2", ["mu"], []).consts == [2.0]


def test_rejects_unknown_symbols_and_functions():
    # test-dsl.R:21-26
    with pytest.raises(dsl.DSLError, match="unknown symbol"):
        dsl.compile_expr("~ foo + mu", ["mu"], ["y"])
    with pytest.raises(dsl.DSLError, match="unsupported function"):
        dsl.compile_expr("~ cos(mu)", ["mu"], ["y"])


def test_depth_tracked():
    # test-dsl.R:28-34
    assert dsl.compile_expr("~ (y - mu) * (y - mu)", ["mu"], ["y"]).depth >= 2


def test_depth_limit():
    # R/dsl.R:115-119: right-nested sums push every operand first
    deep = "~ " + "(mu + " * 40 + "mu" + ")" * 40
    with pytest.raises(dsl.DSLError, match="too deep"):
        dsl.compile_expr(deep, ["mu"], [])


def test_precedence_unary_minus_and_power():
    # R: -x^2 == -(x^2); -a*b == (-a)*b; a^b^c == a^(b^c); a^-b
    assert dsl.compile_expr("~ -x^2", ["x"]).code == listing(
        ("PP", 0), ("PC", 0), "POW", "NEG")
    assert dsl.compile_expr("~ -a*b", ["a", "b"]).code == listing(
        ("PP", 0), "NEG", ("PP", 1), "MUL")
    assert dsl.compile_expr("~ a^b^c", ["a", "b", "c"]).code == listing(
        ("PP", 0), ("PP", 1), ("PP", 2), "POW", "POW")
    assert dsl.compile_expr("~ a^-b", ["a", "b"]).code == listing(
        ("PP", 0), ("PP", 1), "NEG", "POW")
    assert dsl.compile_expr("~ a - b - c", ["a", "b", "c"]).code == listing(
        ("PP", 0), ("PP", 1), "SUB", ("PP", 2), "SUB")


def test_gaussian_mean_listing():
    p = workloads.gaussian_mean(n=10).model.loglik
    assert p.code == listing(("PD", 0), ("PP", 0), "SUB", ("PC", 0), "POW",
                             "NEG", ("PC", 1), "DIV")
    assert p.consts == [2.0, 8.0] and p.depth == 2


def test_gamma_listing():
    p = workloads.gamma(n=10).model.loglik
    e0 = [("PP", 0), "EXP"]
    want = listing(*e0, ("PP", 1), "MUL", *e0, "LGAMMA", "SUB", *e0,
                   ("PC", 0), "SUB", ("PD", 0), "LOG", "MUL", "ADD",
                   ("PP", 1), "EXP", ("PD", 0), "MUL", "SUB")
    assert p.code == want and p.consts == [1.0]
    assert p.n_instr == 21 and p.depth == 3


def test_mixture_listing():
    p = workloads.mixture(n=10).model.loglik
    comp = lambda j: [("PC", 0), ("PD", 0), ("PP", j), "SUB", ("PC", 1), "POW",
                      "NEG", ("PC", 1), "DIV", "EXP", "MUL"]
    assert p.code == listing(*comp(0), *comp(1), "ADD", "LOG")
    assert p.consts == [0.5, 2.0] and p.n_instr == 24 and p.depth == 4


def test_logistic_listing():
    p = workloads.logistic(n=10).model.loglik
    eta = [("PP", 0), ("PD", 1), "MUL"]
    for j in range(1, 16):
        eta += [("PP", j), ("PD", j + 1), "MUL", "ADD"]
    assert p.code == listing(("PD", 0), *eta, "MUL", ("PC", 0), *eta, "EXP",
                             "ADD", "LOG", "SUB")
    assert p.n_instr == 133 and p.depth == 5 and p.consts == [1.0]


def test_model_validation():
    with pytest.raises(dsl.DSLError):
        dsl.gpum_model("~ mu", [], [])
    with pytest.raises(dsl.DSLError, match="distinct"):
        dsl.gpum_model("~ mu", ["mu"], ["mu"])
    m = dsl.gpum_model("~ mu", ["mu"])
    assert m.prior.code == [] and m.n_cols == 0
    with pytest.raises(dsl.DSLError, match="unknown symbol"):
        dsl.gpum_model("~ mu", ["mu"], ["y"], prior="~ y")
