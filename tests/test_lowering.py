"""The host lowering (no GPU): what the kernels run for each program.

gpum_hip_lower_info runs gpum_hip_model_create's lowering on the host only.
These pin the summation fold's classification (affine programs accumulated
directly, trailing squares, sum tails), the common-subexpression pass, the
square rewrite of parameter programs and the VALU weights the roofline
uses; the GPU parity tests check the resulting values against FP64.
"""
import pytest

from gpumetropolis_amd import _native, dsl, workloads

VALUE, SQUARE, AFFINE = 0, 1, 2


def info(m, **kw):
    return _native.lower_info(m.loglik.code, m.loglik.consts, m.n_params,
                              m.n_cols, **kw)


def test_gamma_is_affine_with_its_constant_in_the_tail():
    i = info(workloads.gamma(n=10).model)
    # c + (a-1) log y - b y: log y precomputed, two products, c in the tail
    assert i["sum_mode"] == AFFINE
    assert i["n_sum_tail"] == 1
    assert i["n_cols_total"] == 2
    assert i["n_vm_instr"] == 3          # the zero load and two FMA_UD
    assert i["w_exec"] == 2.0
    assert i["w_bytecode"] == 56.0


def test_regression_sums_a_square_and_moves_scale_and_constant_out():
    m = workloads.regressions(n_models=1, n_min=10, n_max=20)[0].model
    i = info(m)
    # -(y - b0 - sum bj xj)^2 / (2 exp(2 ls)) - ls: square, tail MUL_U SUB_U
    assert i["sum_mode"] == SQUARE
    assert i["n_sum_tail"] == 2
    assert i["w_exec"] == 10.0           # SUB_U + 8 FMA_UD + the squared sum
    assert i["vm_depth"] == 1


def test_gaussian_mean_is_a_scaled_square():
    i = info(workloads.gaussian_mean(n=10).model)
    assert (i["sum_mode"], i["n_sum_tail"], i["w_exec"]) == (SQUARE, 1, 2.0)


def test_logistic_shares_its_linear_predictor():
    m = workloads.logistic(n=10, d=16).model
    on, off = info(m), info(m, fuse=0)
    # y*eta - log(1 + exp(eta)): eta once (16 products), saved and reused
    assert on["w_exec"] == 27.0
    assert on["vm_depth"] == 3           # the stack's two values + the saved eta
    assert off["w_exec"] > on["w_exec"] + 15
    assert on["sum_mode"] == VALUE and on["n_sum_tail"] == 0


def test_mixture_is_a_log_sum_exp():
    m = workloads.mixture(n=10).model
    on, off = info(m), info(m, fuse=0)
    # log(0.5 exp(A1) + 0.5 exp(A2)): each A = -(y - mu)^2 / 2 is
    # LOAD_D SUB_U SQR MUL_U (3 slots), then one LSE_S (11); the shared
    # log 0.5 moves to the sum tail
    assert (on["sum_mode"], on["n_sum_tail"]) == (VALUE, 1)
    assert on["w_exec"] == 17.0
    assert on["vm_depth"] == 2
    # unfused: two EXPs, the weights' multiplies, the add and the LOG
    assert off["w_exec"] == 23.0


@pytest.mark.parametrize("formula,tail", [
    ("~ log(0.3 * exp(y - a) + exp(x * b) * 0.7)", 0),   # unequal weights
    ("~ log(exp(y - a) + exp(x * b))", 0),               # no weights
    ("~ log(2 * exp(y * a) + 2 * exp(x - b))", 1),       # equal weights
])
def test_log_sum_exp_forms(formula, tail):
    m = dsl.gpum_model(formula, ["a", "b"], ["y", "x"])
    i = info(m)
    assert i["n_sum_tail"] == tail
    # one LSE (11) and the two arguments (one op each), plus an ADD_U of
    # log c per unequal weight
    assert i["w_exec"] == 13.0 + (2.0 if "0.3" in formula else 0.0)


def test_log_of_sum_with_a_negative_weight_is_not_fused():
    m = dsl.gpum_model("~ log(2 * exp(y - a) - exp(x * b))", ["a", "b"],
                       ["y", "x"])
    assert info(m)["w_exec"] == info(m, fuse=0)["w_exec"]


@pytest.mark.parametrize("formula,mode,tail", [
    ("~ y * mu - 3 * x + 2 - s", AFFINE, 2),
    ("~ mu * x - y + s", AFFINE, 1),
    ("~ 2 * (x - mu) - s", AFFINE, 2),
    ("~ -log(s) * 2 + 0 * y", AFFINE, 1),
    ("~ (y - mu)^2 * s + 4 - s", SQUARE, 2),
    ("~ (exp(y - mu) + x) * s", VALUE, 1),
    ("~ log(1 + exp(y * mu)) - x * s", VALUE, 0),
])
def test_fold_forms(formula, mode, tail):
    m = dsl.gpum_model(formula, ["mu", "s"], ["y", "x"])
    i = info(m)
    assert (i["sum_mode"], i["n_sum_tail"]) == (mode, tail)
    j = info(m, fuse=0)
    assert (j["sum_mode"], j["n_sum_tail"]) == (VALUE, 0)


def test_constant_squares_in_the_hoist_become_a_multiply():
    sq = dsl.gpum_model("~ y * (mu / 2)^2", ["mu"], ["y"])
    cube = dsl.gpum_model("~ y * (mu / 2)^3", ["mu"], ["y"])
    # PUSH_PARAM, PUSH_CONST, DIV, SQUARE, STORE_U against
    # PUSH_PARAM, PUSH_CONST, DIV, PUSH_CONST, POW, STORE_U (2 words each)
    assert info(sq)["n_hoist_words"] == 10
    assert info(cube)["n_hoist_words"] == 12


def test_invalid_programs_are_refused():
    with pytest.raises(_native.BackendError, match="stack underflow"):
        _native.lower_info([3, 0], [], 1, 1)            # ADD on an empty stack
    with pytest.raises(_native.BackendError, match="out of range"):
        _native.lower_info([2, 5], [], 1, 1)            # data column 5 of 1
    with pytest.raises(_native.BackendError, match="parameters"):
        _native.lower_info([1, 0], [], 17, 1)           # 17 parameters
