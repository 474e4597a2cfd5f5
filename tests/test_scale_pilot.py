"""The delayed-acceptance scale pilot (surrogate.tune_scale) on scripted
pilot runs: the score is the whitened jump distance per millisecond, the
smallest factor within `keep` of the best wins, and no factor whose
acceptance is below the floor is picked."""
from types import SimpleNamespace

import numpy as np

from gpumetropolis_amd.surrogate import PILOT_MIN_ACCEPT, tune_scale


class _Net:
    n_in = 2


class _ScriptedModel:
    """run_da returns draws whose every step jumps `jump` (whitened units)
    in the first coordinate, at `ms` per iteration and acceptance `acc`,
    one scripted entry per call."""

    def __init__(self, script, n_chains=4, iters=60):
        self.script = list(script)
        self.n_chains, self.iters = n_chains, iters

    def run_da(self, net, init, temps, iters, seed, **kw):
        jump, ms, acc = self.script.pop(0)
        x = np.zeros((iters, self.n_chains, 2), np.float32)
        x[:, :, 0] = (np.arange(iters) * jump)[:, None]
        return SimpleNamespace(draws=x, kernel_ms=ms * iters,
                               stage1_rate=np.full(self.n_chains, acc),
                               accept_rate=np.full(self.n_chains, acc))


def _pick(script, **kw):
    dm = _ScriptedModel(script)
    return tune_scale(dm, _Net(), np.zeros((4, 2)), np.eye(2), 1,
                      proposal_l=np.eye(2), factors=(1.0, 1.3, 1.6, 2.0), **kw)


def test_best_jump_per_ms_wins():
    # squared jumps 1, 4, 9, 16 at 1 ms each: factor 2.0 scores 16
    f, rows = _pick([(1.0, 1.0, 0.3), (2.0, 1.0, 0.2), (3.0, 1.0, 0.1),
                     (4.0, 1.0, 0.08)])
    assert f == 2.0
    assert [r["esjd_per_ms"] for r in rows] == [1.0, 4.0, 9.0, 16.0]


def test_smallest_factor_within_keep():
    # 1.6 scores 15.5 >= 0.95 * 16: the smaller factor wins
    f, _ = _pick([(1.0, 1.0, 0.3), (2.0, 1.0, 0.2), (np.sqrt(15.5), 1.0, 0.1),
                  (4.0, 1.0, 0.08)])
    assert f == 1.6


def test_acceptance_floor():
    # factor 2.0 has the best score but accepts below the floor
    f, rows = _pick([(1.0, 1.0, 0.3), (2.0, 1.0, 0.2), (3.0, 1.0, 0.08),
                     (4.0, 1.0, PILOT_MIN_ACCEPT / 2)])
    assert rows[3]["esjd_per_ms"] == max(r["esjd_per_ms"] for r in rows)
    assert f == 1.6
    # no factor reaches the floor: the first factor
    f, _ = _pick([(1.0, 1.0, 0.01), (2.0, 1.0, 0.01), (3.0, 1.0, 0.01),
                  (4.0, 1.0, 0.01)])
    assert f == 1.0
