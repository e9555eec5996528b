"""Autotuner host logic: choice, tie rule and the JSON cache (timings injected)."""
from reprobus_amd import autotune
from reprobus_amd.chem import make_grid


def test_choice_and_cache(tmp_path):
    grid = make_grid(144, 91, 60, 60, 14)
    path = str(tmp_path / "tune.json")
    fake = {64: 2.0, 128: 1.5, 256: 1.7}
    calls = []

    def timer(b):
        calls.append(b)
        return fake[b]

    r = autotune.tune(None, grid, {}, mode="plain", cache_path=path, timer=timer)
    assert r.block == 128 and not r.cached and sorted(calls) == [64, 128, 256]
    r2 = autotune.tune(None, grid, {}, mode="plain", cache_path=path, timer=timer)
    assert r2.cached and r2.block == 128 and len(calls) == 3        # served from the cache
    r3 = autotune.tune(None, grid, {}, mode="converged", cache_path=path, timer=timer)
    assert not r3.cached and len(calls) == 6                          # another key
    other = make_grid(72, 46, 60, 60, 14)
    assert not autotune.tune(None, other, {}, cache_path=path, timer=timer).cached


def test_ties_go_to_the_larger_block():
    assert autotune.choose({64: 1.0, 128: 1.0, 256: 1.2}) == 128
    assert autotune.choose({64: 1.0, 128: 1.0, 256: 1.0}) == 256
