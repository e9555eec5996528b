"""Autotuner host logic: choice, tie rules and the JSON cache (timings injected)."""
from reprobus_amd import autotune
from reprobus_amd.chem import make_grid


def test_choice_and_cache(tmp_path):
    grid = make_grid(144, 91, 60, 60, 14)
    path = str(tmp_path / "tune.json")
    fake = {64: 2.0, 128: 1.5, 256: 1.7}
    calls = []

    def timer(b, refill=None):
        calls.append((b, refill))
        return fake[b] + (0.0 if refill in (None, 16) else 0.25)

    r = autotune.tune(None, grid, {}, mode="plain", cache_path=path, timer=timer)
    assert r.block == 128 and r.refill is None and not r.cached and len(calls) == 3
    assert r.ms == {"64": 2.0, "128": 1.5, "256": 1.7}
    r2 = autotune.tune(None, grid, {}, mode="plain", cache_path=path, timer=timer)
    assert r2.cached and r2.block == 128 and len(calls) == 3        # served from the cache
    # the converged mode: every (block, chunk) pair timed, another cache key
    r3 = autotune.tune(None, grid, {}, mode="converged", cache_path=path, timer=timer)
    assert not r3.cached and len(calls) == 3 + 9
    assert (r3.block, r3.refill) == (128, 16) and r3.ms["128/16"] == 1.5
    r4 = autotune.tune(None, grid, {}, mode="converged", cache_path=path, timer=timer)
    assert r4.cached and (r4.block, r4.refill) == (128, 16)
    other = make_grid(72, 46, 60, 60, 14)
    assert not autotune.tune(None, other, {}, cache_path=path, timer=timer).cached
    # a converged-mode choice is cached per Anderson depth and tolerance
    n = len(calls)
    r5 = autotune.tune(None, grid, {}, mode="converged", cache_path=path, timer=timer, accel=0)
    assert not r5.cached and len(calls) == n + 9
    r6 = autotune.tune(None, grid, {}, mode="converged", cache_path=path, timer=timer,
                       tol=1e-10)
    assert not r6.cached
    # a cached chunk outside the requested ones is not served
    r7 = autotune.tune(None, grid, {}, mode="converged", cache_path=path, timer=timer,
                       refills=(8, 32))
    assert not r7.cached and r7.refill in (8, 32)


def test_ties_go_to_the_larger_block_then_chunk():
    assert autotune.choose({64: 1.0, 128: 1.0, 256: 1.2}) == 128
    assert autotune.choose({64: 1.0, 128: 1.0, 256: 1.0}) == 256
    assert autotune.choose({(128, 8): 1.0, (128, 32): 1.0, (64, 32): 1.0}) == (128, 32)


def test_newton_choice_and_cache(tmp_path):
    grid = make_grid(144, 91, 60, 60, 14)
    path = str(tmp_path / "tune.json")
    calls = []

    def timer(sf, w):
        calls.append((sf, w))
        return {20: 24.0, 40: 20.0, 60: 16.0, 80: 16.0, 100: 17.0, 130: 18.0, 160: 19.0}[sf] + \
            (0.5 if w == 2 else 0.0)

    r = autotune.tune_newton(None, grid, {}, cache_path=path, timer=timer)
    assert (r.sweeps_first, r.waves) == (60, 1) and not r.cached   # tie 60/80 -> fewer sweeps
    assert len(calls) == len(autotune.SWEEPS_FIRST) * len(autotune.NEWTON_WAVES)
    r2 = autotune.tune_newton(None, grid, {}, cache_path=path, timer=timer)
    assert r2.cached and (r2.sweeps_first, r2.waves) == (60, 1)
    assert not autotune.tune_newton(None, grid, {}, cache_path=path, timer=timer, tol=1e-10).cached
    assert not autotune.tune_newton(None, grid, {}, cache_path=path, timer=timer,
                                    sweeps=(40, 80)).cached
