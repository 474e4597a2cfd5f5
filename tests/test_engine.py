"""Host-side choice between the fused chain kernel and the stepwise sampler
(no GPU needed)."""
from gpumetropolis_amd.api import pick_engine


def info(depth, cols):
    return {"vm_depth": depth, "n_cols_total": cols}


def test_gamma_4096_chains_stays_fused():
    assert pick_engine(info(1, 2), 4096, 1_000_000) == "fused"


def test_logistic_1024_chains_wide_rows_goes_stepwise():
    assert pick_engine(info(2, 17), 1024, 4_000_000) == "stepwise"


def test_small_data_stays_fused():
    # 8 chains x 20000 rows: too little work per iteration for launches
    assert pick_engine(info(3, 1), 8, 20_000) == "fused"


def test_narrow_rows_let_few_chains_fill_the_gpu():
    # one column: 16 waves per chain, 256 chains fill 4096 waves
    assert pick_engine(info(1, 1), 256, 2_000_000) == "fused"
    assert pick_engine(info(1, 1), 128, 2_000_000) == "stepwise"
