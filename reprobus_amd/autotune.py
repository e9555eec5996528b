"""Autotuner for the chemistry kernels' launch shape.

The per-box kernels run one box per lane; what is left to choose per device and grid is
the workgroup size (64, 128 or 256 lanes: how many waves share one LDS block of
start-of-step state, and so how the blocks pack onto the CUs and XCDs) for the
fixed-sweep step and for the persistent converged-mode kernel, whose grid is always the
resident wave set the occupancy calculator gives for that size.

`tune` times each candidate on the caller's own device state (restored between trials,
so the state is unchanged on return), keeps the fastest, and caches the choice in a JSON
file keyed by device name, grid and mode.  Every candidate computes the same numbers:
the launch shape changes no result (checked by the tests).
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass, field

BLOCKS = (64, 128, 256)


@dataclass
class TuneResult:
    mode: str
    block: int
    ms: dict = field(default_factory=dict)       # block -> median ms
    cached: bool = False


def _key(device_name: str, grid, mode: str) -> str:
    return f"{device_name}|{grid.nlon}x{grid.nlat}x{grid.niv}/{grid.nivbas}|{mode}"


def load_cache(path: str | None) -> dict:
    if not path or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_cache(path: str | None, cache: dict) -> None:
    if not path:
        return
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def choose(timings: dict) -> int:
    """The fastest candidate; ties go to the larger block (fewer workgroups)."""
    return min(timings, key=lambda b: (timings[b], -b))


def tune(table, grid, state: dict, mode: str = "plain", candidates=BLOCKS, reps: int = 3,
         cache_path: str | None = None, device: int = 0, tol: float = 1e-10,
         maxiter: int = 2000, timer=None) -> TuneResult:
    """Pick the block size for `mode` ("plain": the fixed-sweep step, "converged": the
    persistent converged-mode kernel) on `state` (device tensors T, p, sza, alt, qj1, hc,
    surfarea, vsed3d, parth2osed, parthno3sed, dt).  `timer(block) -> ms` replaces the
    GPU timing (tests)."""
    import torch

    from .chem import ChemistryStep

    name = torch.cuda.get_device_name(device) if timer is None else "test-device"
    cache = load_cache(cache_path)
    key = _key(name, grid, mode)
    if key in cache and cache[key]["block"] in candidates:
        return TuneResult(mode, cache[key]["block"], cache[key]["ms"], cached=True)

    def measure(block: int) -> float:
        step = ChemistryStep(*table, device=device, block=block)
        try:
            s = torch.cuda.current_stream()
            step.set_stream(s.cuda_stream)
            q0, h0 = state["qj1"].clone(), state["hc"].clone()
            args = [state[k] for k in ("T", "p", "sza", "alt")]
            outs = [state[k] for k in ("surfarea", "vsed3d", "parth2osed", "parthno3sed")]

            def once():
                if mode == "plain":
                    step.step_device(grid, state["dt"], *args, state["qj1"], state["hc"], *outs,
                                     fuse_conserv=True)
                else:
                    step.converge_device(grid, state["dt"], tol, maxiter, *args, state["qj1"],
                                         state["hc"], *outs)

            times = []
            for r in range(reps + 1):                   # first run: warm-up
                state["qj1"].copy_(q0)
                state["hc"].copy_(h0)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(s)
                once()
                e1.record(s)
                torch.cuda.synchronize()
                if r:
                    times.append(e0.elapsed_time(e1))
            state["qj1"].copy_(q0)
            state["hc"].copy_(h0)
            torch.cuda.synchronize()
            return sorted(times)[len(times) // 2]
        finally:
            step.close()

    timer = timer or measure
    ms = {int(b): float(timer(int(b))) for b in candidates}
    best = choose(ms)
    cache[key] = {"block": best, "ms": {str(b): v for b, v in ms.items()},
                  "when": time.strftime("%Y-%m-%dT%H:%M:%S")}
    save_cache(cache_path, cache)
    return TuneResult(mode, best, ms)
