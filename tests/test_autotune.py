"""Host logic of the tile autotuner (no GPU needed)."""
import math

import pytest

from gpumetropolis_amd import autotune


def test_key_buckets_program_length_and_sizes():
    k1 = autotune.tile_key([37, 12], [1000, 5000, 900000])
    k2 = autotune.tile_key([37], [1100, 4800, 1000000])
    assert k1 == k2 == "vm37-n12-max20"
    assert autotune.tile_key([38], [1000, 5000, 900000]) != k1


def test_choose_times_candidates_once_and_caches(tmp_path):
    autotune._cache.clear()
    calls = []

    def timer(t):
        calls.append(t)
        return {(16, 1): 3.0, (8, 2): 2.0, (4, 4): 5.0}[t]

    path = str(tmp_path / "tiles.json")
    best, timings = autotune.choose("k", timer, cache_path=path)
    assert best == (8, 2) and len(calls) == 3 and timings[(4, 4)] == 5.0
    best2, timings2 = autotune.choose("k", timer, cache_path=path)
    assert best2 == (8, 2) and len(calls) == 3 and timings2 == {}
    autotune._cache.clear()
    best3, _ = autotune.choose("k", timer, cache_path=path)  # from the file
    assert best3 == (8, 2) and len(calls) == 3


def test_choose_skips_failing_candidates():
    autotune._cache.clear()

    def timer(t):
        if t == (16, 1):
            raise RuntimeError("does not fit")
        return {(8, 2): 4.0, (4, 4): 1.0}[t]

    best, timings = autotune.choose("k2", timer)
    assert best == (4, 4) and math.isinf(timings[(16, 1)])
    with pytest.raises(RuntimeError):
        autotune._cache.clear()
        autotune.choose("k3", lambda t: (_ for _ in ()).throw(RuntimeError("x")))
