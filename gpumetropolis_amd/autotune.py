"""Tile autotuner for batched launches.

A batched launch gives every workgroup 16 waves, split as chains per block
x waves per chain, with one or two chains per wave.  Many chains per block
share each data tile among more chains (less data traffic per evaluation);
more waves per chain shorten a chain's pass over a long data set and keep
more workgroups busy when the batch has few chains; two chains per wave
(programs without a stack only) read each observation from LDS once for
both.  Which wins depends on the program (its length sets the arithmetic
per tile) and on the observation counts, so the choice is measured: each
candidate runs one short batch of the real workload, the fastest is kept,
and later batches with the same key reuse it.  The key is (longest program
in VM instructions, median and largest observation count rounded to powers
of two).  A tiling is (chains per block, waves per chain[, chains per
wave]).
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, Optional, Sequence, Tuple

CANDIDATES: Tuple[Tuple[int, ...], ...] = ((16, 1), (8, 2), (4, 4))
# Two chains per wave, for compiled batches whose programs all have depth 1.
PAIR_CANDIDATES: Tuple[Tuple[int, ...], ...] = ((32, 1, 2), (16, 2, 2),
                                                (8, 4, 2))
# Four chains per wave, when those programs are also all affine (a program
# that keeps a value per observation holds four chains' values and spills:
# 512 regressions, 2206 ms with two chains per wave, 4240 ms with four).
QUAD_CANDIDATES: Tuple[Tuple[int, ...], ...] = ((16, 4, 4),)

_cache: Dict[str, Tuple[int, ...]] = {}


def tile_key(n_vm_instr: Sequence[int], n_obs: Sequence[int]) -> str:
    obs = sorted(int(n) for n in n_obs)
    med = obs[len(obs) // 2]
    lg = lambda n: int(round(math.log2(max(1, n))))
    return f"vm{max(n_vm_instr)}-n{lg(med)}-max{lg(obs[-1])}"


def load(path: Optional[str]) -> None:
    if path and os.path.exists(path):
        with open(path) as f:
            _cache.update({k: tuple(v) for k, v in json.load(f).items()})


def save(path: Optional[str]) -> None:
    if path:
        with open(path, "w") as f:
            json.dump({k: list(v) for k, v in _cache.items()}, f, indent=1)


def choose(key: str, time_candidate, candidates=CANDIDATES,
           cache_path: Optional[str] = None):
    """The cached tiling for `key`, or the fastest candidate by
    time_candidate(tiling) -> milliseconds.  Returns (tiling, timings)."""
    load(cache_path)
    if key in _cache:
        return _cache[key], {}
    timings = {}
    for t in candidates:
        try:
            timings[t] = float(time_candidate(t))
        except Exception as e:  # a tiling the batch cannot use
            timings[t] = math.inf
            last = e
    best = min(timings, key=timings.get)
    if not math.isfinite(timings[best]):
        raise last
    _cache[key] = best
    save(cache_path)
    return best, timings
