"""Autotuner for the chemistry kernels' launch shape.

The per-box kernels run one box per lane; what is left to choose per device and grid is
the workgroup size (64, 128 or 256 lanes: how many waves share one LDS block of
start-of-step state, and so how the blocks pack onto the CUs and XCDs) for the
fixed-sweep step (its set-up and sweep kernels) and for the persistent converged-mode
kernels, whose grid is always the
resident wave set the occupancy calculator gives for that size; and, for the converged
mode, the queue chunk (how many lanes of a wave must be idle before the wave takes new
boxes: a small chunk keeps lanes busy, a large one runs the set-up and write-back code
less often).  The converged mode is tuned on the accelerated solve.

`tune` times each candidate on the caller's own device state (restored between trials,
so the state is unchanged on return), keeps the fastest, and caches the choice in a JSON
file keyed by device name, grid and mode.  `tune_newton` chooses the Newton solve's launch
shape the same way: how many plain sweeps run before Newton takes the unconverged boxes,
and the Newton kernel's resident waves per SIMD (its occupancy against the register file
its factors need).  Every candidate computes the same numbers:
the launch shape and the chunk change no result (checked by the tests).
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass, field

BLOCKS = (64, 128, 256)
REFILLS = (8, 16, 32)
# Newton solve: plain sweeps before Newton takes the still unconverged boxes, and resident
# waves per SIMD of the Newton kernel (the occupancy it runs at)
SWEEPS_FIRST = (20, 40, 60, 80, 100, 130)
NEWTON_WAVES = (1, 2)


@dataclass
class TuneResult:
    mode: str
    block: int
    refill: int | None = None                    # converged mode: the queue chunk
    ms: dict = field(default_factory=dict)       # "block" or "block/refill" -> median ms
    cached: bool = False
    sweeps_first: int | None = None              # Newton solve: sweeps before Newton
    waves: int | None = None                     # Newton solve: waves per SIMD


def _key(device_name: str, grid, mode: str, accel: int | None = None,
         tol: float | None = None) -> str:
    """Cache key: device, grid and mode; the converged mode's choice also depends on the
    Anderson depth and the tolerance it was timed with."""
    k = f"{device_name}|{grid.nlon}x{grid.nlat}x{grid.niv}/{grid.nivbas}|{mode}"
    if accel is not None:
        k += f"|accel={int(accel)}|tol={float(tol):g}"
    return k


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


def _label(shape: tuple) -> str:
    return "/".join(str(v) for v in shape)


def choose(timings: dict):
    """The fastest candidate (a block, or a (block, refill) pair); ties go to the larger
    block (fewer workgroups), then the larger chunk."""
    def key(c):
        t = c if isinstance(c, tuple) else (c,)
        return (timings[c],) + tuple(-v for v in t)
    return min(timings, key=key)


def tune(table, grid, state: dict, mode: str = "plain", candidates=BLOCKS, reps: int = 3,
         cache_path: str | None = None, device: int = 0, tol: float = 1e-11,
         maxiter: int = 2000, accel: int = 2, refills=REFILLS, timer=None,
         steps: int = 5) -> TuneResult:
    """Pick the launch shape for `mode` ("plain": the fixed-sweep step's block;
    "converged": block and queue chunk of the converged mode with Anderson depth `accel`)
    on `state` (device tensors T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed,
    parthno3sed, dt).  A plain trial times `steps` successive steps (ms per step), as a
    model run issues them: a lone step after the state's restore starts from cold caches
    and an idle queue, and its time is noisier than the differences between blocks.
    `timer(block)` (plain) or `timer(block, refill)` (converged) -> ms replaces the GPU
    timing (tests)."""
    import torch

    from .chem import ChemistryStep

    name = torch.cuda.get_device_name(device) if timer is None else "test-device"
    cache = load_cache(cache_path)
    conv = mode != "plain"
    key = _key(name, grid, mode, accel, tol) if conv else _key(name, grid, mode)
    if key in cache and cache[key]["block"] in candidates and (
            not conv or cache[key].get("refill") in refills):
        c = cache[key]
        return TuneResult(mode, c["block"], c.get("refill"), c["ms"], cached=True)

    def measure(block: int, refill: int = 16) -> float:
        step = ChemistryStep(*table, device=device, block=block, conv_refill=refill)
        try:
            s = torch.cuda.current_stream()
            step.set_stream(s.cuda_stream)
            q0, h0 = state["qj1"].clone(), state["hc"].clone()
            args = [state[k] for k in ("T", "p", "sza", "alt")]
            outs = [state[k] for k in ("surfarea", "vsed3d", "parth2osed", "parthno3sed")]

            def once():
                if not conv:
                    for _ in range(steps):
                        step.step_device(grid, state["dt"], *args, state["qj1"], state["hc"],
                                         *outs, fuse_conserv=True)
                else:
                    step.converge_device(grid, state["dt"], tol, maxiter, *args, state["qj1"],
                                         state["hc"], *outs, accel=accel)

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
                    times.append(e0.elapsed_time(e1) / (1 if conv else steps))
            state["qj1"].copy_(q0)
            state["hc"].copy_(h0)
            torch.cuda.synchronize()
            return sorted(times)[len(times) // 2]
        finally:
            step.close()

    timer = timer or measure
    if conv:
        ms = {(int(b), int(f)): float(timer(int(b), int(f))) for b in candidates for f in refills}
        best = choose(ms)
        block, refill = best
    else:
        ms = {int(b): float(timer(int(b))) for b in candidates}
        block, refill = choose(ms), None
    labels = {_label(k if isinstance(k, tuple) else (k,)): v for k, v in ms.items()}
    cache[key] = {"block": block, "refill": refill, "ms": labels,
                  "when": time.strftime("%Y-%m-%dT%H:%M:%S")}
    save_cache(cache_path, cache)
    return TuneResult(mode, block, refill, labels)


def tune_newton(table, grid, state: dict, block: int = 64, refill: int = 32, reps: int = 3,
                cache_path: str | None = None, device: int = 0, tol: float = 1e-11,
                maxiter: int = 2000, max_newton: int = 20, sweeps=SWEEPS_FIRST,
                waves=NEWTON_WAVES, timer=None) -> TuneResult:
    """Pick the Newton solve's sweeps before Newton and waves per SIMD on `state` (the
    device tensors of tune()), timing every pair (restoring the state between trials), and
    cache the choice keyed by device, grid and tolerance.  Every pair reaches the same
    fixed point (the tests check the answers against the restatement).  `timer(sweeps,
    waves)` -> ms replaces the GPU timing (tests)."""
    import torch

    name = torch.cuda.get_device_name(device) if timer is None else "test-device"
    cache = load_cache(cache_path)
    key = _key(name, grid, "newton") + f"|tol={float(tol):g}|block={block}"
    c = cache.get(key)
    if c and c.get("sweeps_first") in sweeps and c.get("waves") in waves:
        return TuneResult("newton", block, refill, c["ms"], cached=True,
                          sweeps_first=c["sweeps_first"], waves=c["waves"])

    def measure(sf: int, w: int) -> float:
        from .chem import ChemistryStep

        step = ChemistryStep(*table, device=device, block=block, conv_refill=refill,
                             newton_waves=w)
        try:
            s = torch.cuda.current_stream()
            step.set_stream(s.cuda_stream)
            q0, h0 = state["qj1"].clone(), state["hc"].clone()
            args = [state[k] for k in ("T", "p", "sza", "alt")]
            outs = [state[k] for k in ("surfarea", "vsed3d", "parth2osed", "parthno3sed")]
            times = []
            for r in range(reps + 1):                   # first run: warm-up
                state["qj1"].copy_(q0)
                state["hc"].copy_(h0)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(s)
                step.newton_device(grid, state["dt"], tol, max_newton, maxiter, *args,
                                   state["qj1"], state["hc"], *outs, sweeps_first=sf)
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
    ms = {(int(a), int(b)): float(timer(int(a), int(b))) for a in sweeps for b in waves}
    # ties: fewer sweeps first (more boxes to Newton), then fewer waves
    sf, w = min(ms, key=lambda k: (ms[k], k[0], k[1]))
    labels = {_label(k): v for k, v in ms.items()}
    cache[key] = {"sweeps_first": sf, "waves": w, "ms": labels,
                  "when": time.strftime("%Y-%m-%dT%H:%M:%S")}
    save_cache(cache_path, cache)
    return TuneResult("newton", block, refill, labels, sweeps_first=sf, waves=w)
