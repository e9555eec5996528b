"""Benchmark: REPROBUS chemistry step on one MI355X.

Headline (`value`): the full 144 x 91 x 60 grid (786,240 boxes), one 900 s chemistry
step per timed step -- the reference scheme, 5 fixed sweeps in FP64, conserv + floor
fused -- state resident in HBM; successive steps integrate the evolving state.

Beside it on the same line: the converged mode (sweeps to the fixed point: plain sweeps
and the Newton solve, each from the incoming state and from the learned initial guess,
and Anderson-accelerated sweeps from the incoming state) on the 786,240-box c2 grid
(polar vortex), with agreement and sweep histograms; the
surrogate step (the learned network's answer kept where its FP64 residual is under 1e-8,
the FP64 solve elsewhere) over a day of sub-steps beside the FP64 solve; the
semi-Lagrangian transport interpolation of 43 tracers on the model's 180 x 91 x 137 grid
against the HBM roofline; a 96-step day of
(device zenith + step) as one HIP graph against the same steps launched one by one;
the autotuned launch shapes; the roofline of the per-box kernel; and the reference
Fortran timed on the host cores.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-extras]
                  [--extras converged,surrogate,day,transport] [--surrogate-steps 96]
                  [--block B] [--conv-shape BLOCK,CHUNK --newton-shape SWEEPS_FIRST,WAVES]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "chem box-steps/s, plain vs MLP-guess, 786k boxes at 1e-9 rel err"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix)
# Compulsory HBM bytes per box of the per-box kernel: reads T, p, SZA, alt, the 3
# column sums, 43 qj1 and 15 hc; writes 43 qj1, 15 hc, surfarea and the 3
# sedimentation fields = 65 + 62 doubles.
BYTES_PER_BOX = (65 + 62) * 8
# FP64 operations per box-step of the reference algorithm, counted exactly by running the
# restatement with a counting number type (oracle/opcount.cpp): 2,201 add, 3,544 mul,
# 595 div, 13 sqrt; plus 166 calls of exp/log/pow/trig, counted as one operation each.
FLOP_PER_BOX = 2201 + 3544 + 595 + 13 + 166
# Measured HBM bytes per step of the per-box kernels on this workload (k_setup + k_chem),
# per block size: FETCH_SIZE x 2 + WRITE_SIZE in KiB, the factors from a 1 GiB copy of
# known traffic (profiles/k_chem_c1_pmc.md).
TRAFFIC_BYTES = {64: round((2 * (45985.94 + 451234.69) + 365464.75 + 557634.63) * 1024),
                 128: round((2 * (45851.00 + 457743.44) + 365464.38 + 572992.88) * 1024),
                 256: round((2 * (45768.75 + 473824.38) + 365468.50 + 606776.63) * 1024)}
# FP64 VALU instructions per step of the two kernels (SQ_INSTS_VALU_*_F64, same profile,
# equal for every block size): add + mul + transcendental + 2 x fma, per lane
FP64_INSTR_FLOP = ((27.64e6 + 31.27e6) + (14.75e6 + 39.33e6) + (1.74e6 + 5.84e6)
                   + 2 * (25.94e6 + 50.28e6)) * 64


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="CPU baseline processes (0: the cores this process may use)")
    ap.add_argument("--cpu-nrep", type=int, default=8)
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline step (for profiler runs)")
    ap.add_argument("--extras", default="converged,surrogate,day,transport,mlp",
                    help="comma-separated sections reported beside the headline")
    ap.add_argument("--surrogate-steps", type=int, default=96,
                    help="sub-steps of the surrogate run (96 = one day, 960 = ten days)")
    ap.add_argument("--block", type=int, default=0, help="0: autotuned")
    ap.add_argument("--conv-shape", default="",
                    help="BLOCK,CHUNK of the converged-mode kernels (default: autotuned)")
    ap.add_argument("--newton-shape", default="",
                    help="SWEEPS_FIRST,WAVES of the Newton solve (default: autotuned)")
    ap.add_argument("--max-newton", type=int, default=20,
                    help="Newton iterations a box gets before it goes back to plain sweeps")
    ap.add_argument("--conv-order", type=int, default=0,
                    help="converged mode's work order: 0 night boxes then day boxes, 1 grid order")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="after the timed region, one device copy of this many MB (a kernel of "
                         "known traffic, to scale profiler byte counters)")
    return ap.parse_args()


def default_cpu_procs() -> int:
    """Host cores this process may use: its CPU affinity set, capped at the job's CPU
    share where the environment states one (OMP_NUM_THREADS)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(share))) if share.isdigit() and int(share) > 0 else n


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _port_rate(sub, table, nthreads, min_s=2.0):
    """Box-steps/s of the C restatement on `sub` with nthreads OpenMP threads: successive
    steps for at least min_s seconds."""
    import cport

    g = sub.grid
    nb = g.nlon * g.nlat * g.nivbas
    x = sub.copy()
    n, t0 = 0, time.perf_counter()
    while True:
        out = cport.chem_step(x, *table, nthreads=nthreads, conserv=False)
        x.qj1, x.hc = out["qj1_chem"], out["hc_chem"]
        n += 1
        dt = time.perf_counter() - t0
        if dt >= min_s:
            return n * nb / dt


def cpu_baseline(inp, nproc, nrep):
    """Reference Fortran (flang build) on a bounded sample: 180 columns spread over the
    grid (one reference latitude row), nrep successive steps per process in nproc
    concurrent single-threaded processes.  The reference always computes its 94 levels: the
    sample's 60 levels sit at reference levels 1-60 and level 60 is repeated on 61-94, so
    81 of 94 levels are heterogeneous against 47 of 60 in the c1 grid.  Beside it, the C
    restatement on nproc threads on the same 94-level sample and on the sample's own 60
    levels gives the level-mix factor, and with it the reference's rate on c1's levels."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    from reprobus_amd import jtable, synth

    if not refrun.available():
        return None
    g = inp.grid
    ncol = g.nlon * g.nlat
    cols = np.linspace(0, ncol - 1, 180).astype(int)
    sub = inp.copy()
    sub.grid = synth.Grid(180, 1, g.niv, g.nivbas, g.nivhau)
    flat = lambda a: np.ascontiguousarray(a.reshape(a.shape[:-2] + (ncol,))[..., cols][..., None, :])
    sub.T, sub.p, sub.sza, sub.alt = flat(inp.T), flat(inp.p), flat(inp.sza), flat(inp.alt)
    sub.qj1, sub.hc = flat(inp.qj1), flat(inp.hc)
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    r = refrun.time_reference(sub, nrep=nrep, nproc=nproc)
    # the 94-level embedding the reference runs, as the restatement's input
    ext = 94 - g.nivbas
    rep = lambda a, ax: np.ascontiguousarray(np.concatenate(
        [a, np.repeat(np.take(a, [g.nivbas - 1], axis=ax), ext, axis=ax)], axis=ax))
    s94 = sub.copy()
    s94.grid = synth.Grid(180, 1, 94, 94, g.nivhau)
    s94.T, s94.p, s94.sza, s94.alt = (rep(a, 0) for a in (sub.T, sub.p, sub.sza, sub.alt))
    s94.qj1, s94.hc = rep(sub.qj1, 1), rep(sub.hc, 1)
    table = jtable.default_table()
    port60 = _port_rate(sub, table, nproc)
    port94 = _port_rate(s94, table, nproc)
    factor = port60 / port94
    return {"value": r["box_steps_per_s"], "unit": "box-steps/s", "cores": nproc,
            "kind": "reference",
            "cpu_model": cpu_model(), "host_affinity_cores": affinity,
            "levels": (f"the reference's 94 levels: the sample's {g.nivbas} at reference levels "
                       f"1-{g.nivbas}, level {g.nivbas} repeated on {g.nivbas + 1}-94 "
                       f"(heterogeneous chemistry on {94 - g.nivhau + 1} of 94 levels; c1: "
                       f"{g.nivbas - g.nivhau + 1} of {g.nivbas})"),
            "sample": (f"flang build of the reference chem3d (src/reprobus_l137_20230723.f + "
                       f"jno.f90), {nproc} concurrent single-threaded processes, each "
                       f"{nrep} successive steps of 180 columns sampled across the 144x91 grid "
                       f"({r['calls']} row calls, {r['chem3d_s']:.1f} CPU-s); rate summed over "
                       f"processes, timed around chem3d"),
            "per_core": r["per_core_box_steps_per_s"],
            "c1_levels": {"value": r["box_steps_per_s"] * factor,
                          "how": (f"the reference's rate x the C restatement's ratio on {nproc} "
                                  f"threads between the sample's own {g.nivbas} levels "
                                  f"({port60:.3g} box-steps/s) and the 94-level embedding "
                                  f"({port94:.3g}): {factor:.3f}")}}


def _device_state(inp, dev):
    import torch

    g = inp.grid
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    st = {"T": t(inp.T), "p": t(inp.p), "sza": t(inp.sza), "alt": t(inp.alt), "qj1": t(inp.qj1),
          "hc": t(inp.hc), "surfarea": torch.zeros(g.shape3, dtype=torch.float64, device=dev),
          "dt": inp.dt}
    shb = (g.nivbas, g.nlat, g.nlon)
    for k in ("vsed3d", "parth2osed", "parthno3sed"):
        st[k] = torch.zeros(shb, dtype=torch.float64, device=dev)
    return st


def _timed(fn, reps, stream):
    import torch

    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    return sorted(ms)[len(ms) // 2]


def converged_section(table, block_conv, refill, dev, newton_shape=(60, 1), tol=1e-11,
                      maxiter=2000, max_newton=20, order=0):
    """Converged mode on the c2 grid (seed 1, the guess network's evaluation seed): plain
    sweeps, Anderson-accelerated sweeps and the Newton solve (generated sparse LU, after
    newton_shape[0] plain sweeps, newton_shape[1] waves per SIMD), each from the incoming
    state and from the learned guess (the guess launch inside the timed solve); agreement
    with the plain solve over boxes that converged on both paths."""
    import torch

    from reprobus_amd import guess, synth
    from reprobus_amd.chem import ChemistryStep, make_grid
    from reprobus_amd.mlp import GuessMLP

    inp = synth.make_c2(1)
    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    st = _device_state(inp, dev)
    step = ChemistryStep(*table, device=0, block=block_conv, conv_refill=refill,
                         newton_waves=newton_shape[1], conv_order=order)
    net = GuessMLP(0)
    if not guess.load(net):
        return {"error": "no trained guess weights (reprobus_amd/data/guess_mlp.npz)"}
    stream = torch.cuda.current_stream()
    step.set_stream(stream.cuda_stream)
    net.set_stream(stream.cuda_stream)
    q0, h0 = st["qj1"].clone(), st["hc"].clone()
    shb = (g.nivbas, g.nlat, g.nlon)
    it = torch.zeros(shb, dtype=torch.int32, device=dev)
    gq, gh = torch.zeros_like(q0), torch.zeros_like(h0)
    args = [st[k] for k in ("T", "p", "sza", "alt")]
    aux = [st[k] for k in ("surfarea", "vsed3d", "parth2osed", "parthno3sed")]

    nb = g.nivbas
    rel = lambda a, b: ((a - b).abs() / torch.maximum(a.abs(), b.abs()).clamp(min=1e-30)).amax(0)
    nbox = g.nbox
    guess_fn = lambda: net.guess(grid, inp.dt, st["T"], st["p"], st["sza"], st["qj1"], st["hc"], gq, gh)
    guess_ms = _timed(guess_fn, 5, stream)
    res, ref = {}, None
    # (name, solver: Anderson depth or "newton", learned guess)
    for name, accel, use_guess in (("plain", 0, False), ("mlp_guess", 0, True),
                                   ("accelerated", 2, False),
                                   ("newton", "newton", False), ("mlp_guess_newton", "newton", True)):
        def solve():
            st["qj1"].copy_(q0)
            st["hc"].copy_(h0)
            if use_guess:
                guess_fn()
            g_ = {"guess_qj1": gq if use_guess else None, "guess_hc": gh if use_guess else None}
            if accel == "newton":
                step.newton_device(grid, inp.dt, tol, max_newton, maxiter, *args, st["qj1"],
                                   st["hc"], *aux, iters=it, sweeps_first=newton_shape[0], **g_)
            else:
                step.converge_device(grid, inp.dt, tol, maxiter, *args, st["qj1"], st["hc"], *aux,
                                     iters=it, accel=accel, **g_)

        solve()
        torch.cuda.synchronize()
        ms = _timed(solve, 3, stream)
        r = {"ms": ms, "box_steps_per_s": nbox / (ms / 1e3),
             "fallback_boxes": int((it < 0).sum())}
        if accel == "newton":
            sw = (it > 0) & (it < 10000)
            nw = (it >= 10000) & (it < 20000)
            r.update({"boxes_converged_by_first_sweeps": int(sw.sum()),
                      "boxes_solved_by_newton": int(nw.sum()),
                      "newton_iterations_mean": float((it[nw] - 10000).float().mean()),
                      "newton_iterations_max": int((it[nw] - 10000).max()),
                      "newton_iteration_histogram": {
                          str(k): int(v) for k, v in
                          enumerate(torch.bincount((it[nw] - 10000).reshape(-1).long()).tolist())
                          if v},
                      "boxes_handed_back_to_sweeps": int((it >= 20000).sum())})
        else:
            r.update({"sweeps_mean": float(it.abs().float().mean()),
                      "sweeps_max": int(it.abs().max()), "sweep_histogram": _histogram(it.abs())})
        if ref is None:
            ref = (st["qj1"].clone(), st["hc"].clone(), it.clone())
        else:
            d = torch.maximum(rel(st["qj1"][:42, :nb], ref[0][:42, :nb]),
                              rel(st["hc"][:, :nb], ref[1][:, :nb]))
            live = (ref[2] > 0) & (it > 0)
            dl = d[live]
            r["agreement_with_plain"] = {
                "max_rel": float(dl.max()), "p99_rel": float(torch.quantile(dl[:2_000_000].float(), 0.99)),
                "boxes_over_1e-9": int((dl > 1e-9).sum()), "boxes_compared": int(live.sum())}
            r["speedup_over_plain"] = res["plain"]["ms"] / ms
        res[name] = r
    # the next sub-step (15 min later, from the plain solve's state, conserv fused): the
    # solvers on a relaxed state, where the radicals are no longer random
    st["qj1"].copy_(ref[0])
    st["hc"].copy_(ref[1])
    st["sza"].copy_(torch.from_numpy(synth.c2_sza(inp.meta, 1)).to(dev))
    q0.copy_(st["qj1"])
    h0.copy_(st["hc"])
    nxt = {}
    for name, accel in (("plain", 0), ("accelerated", 2), ("newton", "newton")):
        def solve():
            st["qj1"].copy_(q0)
            st["hc"].copy_(h0)
            if accel == "newton":
                step.newton_device(grid, inp.dt, tol, max_newton, maxiter, *args, st["qj1"],
                                   st["hc"], *aux, iters=it, sweeps_first=newton_shape[0])
            else:
                step.converge_device(grid, inp.dt, tol, maxiter, *args, st["qj1"], st["hc"],
                                     *aux, iters=it, accel=accel)

        ms = _timed(solve, 3, stream)
        nxt[name] = {"ms": ms, "box_steps_per_s": nbox / (ms / 1e3),
                     "fallback_boxes": int((it < 0).sum())}
        if accel != "newton":
            nxt[name].update({"sweeps_mean": float(it.abs().float().mean()),
                              "sweeps_max": int(it.abs().max())})
    nxt["speedup_accelerated_over_plain"] = nxt["plain"]["ms"] / nxt["accelerated"]["ms"]
    nxt["speedup_newton_over_plain"] = nxt["plain"]["ms"] / nxt["newton"]["ms"]
    out = {"workload": "c2 seed 1 (786,240 boxes, polar vortex), incoming state, dt 900 s",
           "tol": tol, "maxiter": maxiter, "block": block_conv, "refill": refill,
           "work_order": "night boxes, then day boxes" if order == 0 else "grid order",
           "anderson": {"depth": 2, "from_sweep": step.conv_aa_start,
                        "budget": step.conv_aa_budget},
           "newton_shape": {"sweeps_first": newton_shape[0], "waves_per_simd": newton_shape[1],
                      "max_newton": max_newton, "unknowns": 55,
                      "jacobian_nonzeros_day_night": [399, 278],
                      "lu_nonzeros_day_night": [514, 365]},
           "guess_ms": guess_ms, **res,
           "speedup_mlp_over_plain": res["plain"]["ms"] / res["mlp_guess"]["ms"],
           "speedup_mlp_newton_over_plain": res["plain"]["ms"] / res["mlp_guess_newton"]["ms"],
           "speedup_newton_over_plain": res["plain"]["ms"] / res["newton"]["ms"],
           "mlp_gain_in_newton": res["newton"]["ms"] / res["mlp_guess_newton"]["ms"],
           "next_step": nxt}
    step.close()
    net.close()
    return out


def mlp_section(dev, batch=8192, nbig=786_240, reps=20):
    """The learned guess's own kernels: one Adam training step at the training batch (features
    -> forward -> loss -> backward -> Adam, all BF16 MFMA GEMMs) and the GEMM forward chain
    over a c1-sized batch; achieved BF16 rates on the GEMMs' algorithmic flops (the 60-, 55-
    wide layers counted padded to 64, as computed)."""
    import torch

    from reprobus_amd.mlp import NIN, NOUT, GuessMLP, init_params

    net = GuessMLP(0)
    stream = torch.cuda.current_stream()
    net.set_stream(stream.cuda_stream)
    net.set_params(init_params(0))
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    X = torch.randn((nbig, NIN), generator=g, device=dev, dtype=torch.float32)
    T = torch.randn((nbig, NOUT), generator=g, device=dev, dtype=torch.float32)
    net.set_norm(np.zeros(NIN), np.ones(NIN), np.zeros(NOUT), np.ones(NOUT))
    Y = torch.empty((nbig, NOUT), device=dev, dtype=torch.float32)
    fwd_flop_row = 2 * (64 * 256 + 2 * 256 * 256 + 256 * 64)
    # backward: the dX products of layers 2-4 and the dW products of all four
    train_flop_row = fwd_flop_row + (fwd_flop_row - 2 * 64 * 256) + fwd_flop_row
    Xb, Tb = X[:batch].contiguous(), T[:batch].contiguous()
    net.train_step(Xb, Tb, 1e-3)
    torch.cuda.synchronize()
    t_ms = _timed(lambda: net.train_step(Xb, Tb, 1e-3), reps, stream)
    net.forward(X, Y)
    torch.cuda.synchronize()
    f_ms = _timed(lambda: net.forward(X, Y), 5, stream)
    net.close()
    return {"train_step": {"batch": batch, "ms": t_ms,
                           "tflops": batch * train_flop_row / (t_ms / 1e3) / 1e12,
                           "note": "whole Adam step: GEMMs, loss, bias sums, Adam, BF16 copies"},
            "forward_gemm_chain": {"rows": nbig, "ms": f_ms,
                                   "tflops": nbig * fwd_flop_row / (f_ms / 1e3) / 1e12,
                                   "peak_tflops": 2500.0},
            "flop_per_row": {"forward": fwd_flop_row, "train": train_flop_row}}


def _histogram(iters):
    """Boxes per sweep-count bin (edges 1, 5, 10, 20, 40, 60, 80, 120, 200, 500, 1000,
    2000, above)."""
    import torch

    edges = [1, 5, 10, 20, 40, 60, 80, 120, 200, 500, 1000, 2000]
    x = iters.reshape(-1).long()
    out = {}
    for lo, hi in zip(edges, edges[1:] + [1 << 30]):
        out[f"{lo}-{hi - 1}" if hi < (1 << 30) else f">={lo}"] = int(((x >= lo) & (x < hi)).sum())
    return out


def day_section(table, block, dev, nsteps=96):
    """96 sub-steps (device zenith at each step's time + the 5-sweep step, conserv fused) of
    the c2 day on the 786,240-box grid: one HIP graph against the launches one by one."""
    import torch

    from reprobus_amd import synth
    from reprobus_amd.chem import ChemistryStep, make_grid

    inp = synth.make_c2(0)
    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    st = _device_state(inp, dev)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    lat, lon = inp.meta["lat"], inp.meta["lon"]
    tlon = t(np.broadcast_to(lon[None, None, :], g.shape3))
    tlat = t(np.broadcast_to(lat[None, :, None], g.shape3))
    sza = torch.zeros((g.nivbas, g.nlat, g.nlon), dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream()
    step = ChemistryStep(*table, device=0, block=block)
    step.set_stream(stream.cuda_stream)
    q0, h0 = st["qj1"].clone(), st["hc"].clone()
    aux = [st[k] for k in ("surfarea", "vsed3d", "parth2osed", "parthno3sed")]
    res = {}
    with torch.cuda.stream(stream):
        for graph in (False, True, True):
            st["qj1"].copy_(q0)
            st["hc"].copy_(h0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            step.run_day_device(grid, nsteps, 265.0, inp.dt, tlon, tlat, st["T"], st["p"],
                                st["alt"], sza, st["qj1"], st["hc"], *aux, use_graph=graph)
            e1.record(stream)
            torch.cuda.synchronize()
            res["graph" if graph else "eager"] = e0.elapsed_time(e1)   # graph: the replay
    step.close()
    return {"workload": f"c2 seed 0 day: {nsteps} x (zenith + 5-sweep step), 786,240 boxes",
            "ms_eager": res["eager"], "ms_graph": res["graph"],
            "box_steps_per_s_graph": g.nbox * nsteps / (res["graph"] / 1e3)}


def surrogate_section(table, block_conv, refill, dev, nsteps=96, accept_tol=1e-8, tol=1e-11,
                      maxiter=2000):
    """Surrogate step with fallback on the c2 grid (seed 0, vortex and liquid PSC), nsteps
    sub-steps of 15 min with diurnal SZA (96 = one day).  Two states advance side by side
    from the same start: the surrogate path (the learned network's answer for the step,
    kept where one FP64 sweep from it changes no species by more than accept_tol, the
    plain sweeps to tol from it elsewhere) and the FP64 path (plain sweeps to tol from the
    incoming state); conserv + floor after each step on both.  Reported: accept rate by
    latitude band and by day/night, the network's one-step error per species against the
    FP64 step, the end-of-run difference between the two states, the drift of total
    chlorine and bromine (Cly + its source gases, Bry + its source gases: what the FP64
    solve conserves at its fixed point), and box-steps/s of both paths."""
    import torch

    from reprobus_amd import guess, synth
    from reprobus_amd.chem import ChemistryStep, make_grid
    from reprobus_amd.mlp import GuessMLP

    inp = synth.make_c2(0)
    g = inp.grid
    nb = g.nivbas
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    net = GuessMLP(0)
    if not guess.load(net):
        return {"error": "no trained guess weights (reprobus_amd/data/guess_mlp.npz)"}
    sp = _device_state(inp, dev)                    # surrogate path
    fp = _device_state(inp, dev)                    # FP64 path
    step = ChemistryStep(*table, device=0, block=block_conv, conv_refill=refill)
    stream = torch.cuda.current_stream()
    step.set_stream(stream.cuda_stream)
    net.set_stream(stream.cuda_stream)
    shb = (nb, g.nlat, g.nlon)
    it_s = torch.zeros(shb, dtype=torch.int32, device=dev)
    it_f = torch.zeros(shb, dtype=torch.int32, device=dev)
    gq, gh = torch.zeros_like(sp["qj1"]), torch.zeros_like(sp["hc"])
    # the diurnal SZA of the day's 96 steps (one surface field each, the same on every level)
    s2 = torch.from_numpy(np.stack([synth.c2_sza(inp.meta, k)[0] for k in range(96)])).to(dev)
    sza = sp["sza"]                                  # shared by both paths
    fp["sza"] = sza
    lat = torch.from_numpy(np.asarray(inp.meta["lat"], dtype=np.float64)).to(dev)
    bands = {"vortex, lat <= 60S": lat <= -60.0, "60S-30S": (lat > -60.0) & (lat < -30.0),
             "30S-30N": (lat >= -30.0) & (lat <= 30.0), "30N-60N": (lat > 30.0) & (lat < 60.0),
             "60N-90N": lat >= 60.0}
    band_acc = {k: torch.zeros((), dtype=torch.float64, device=dev) for k in bands}
    band_n = {k: 0 for k in bands}
    dn_acc = {"day": torch.zeros((), dtype=torch.float64, device=dev),
              "night": torch.zeros((), dtype=torch.float64, device=dev)}
    dn_n = {"day": torch.zeros((), dtype=torch.float64, device=dev),
            "night": torch.zeros((), dtype=torch.float64, device=dev)}
    L, S = synth.L, synth.S
    cl_w = {"cly": 1, "cfc11": 3, "cfc12": 2, "cfc113": 3, "ccl4": 4, "ch3ccl3": 3, "ch3cl": 1,
            "hcfc22": 1, "h1211": 1}
    br_w = {"bry": 1, "ch3br": 1, "h1211": 1, "h1301": 1, "ch2br2": 2}
    total = lambda q, w: sum(c * q[L[n], :nb] for n, c in w.items())
    cl0, br0 = total(sp["qj1"], cl_w), total(sp["qj1"], br_w)
    args = lambda s: [s[k] for k in ("T", "p", "sza", "alt")]
    aux = lambda s: [s[k] for k in ("surfarea", "vsed3d", "parth2osed", "parthno3sed")]

    def rel_stats(a, b):
        """per species (max, p99) of |a - b| / |b| over boxes where either is above 1e-30"""
        a, b = a.reshape(a.shape[0], -1), b.reshape(b.shape[0], -1)
        m = torch.maximum(a.abs(), b.abs()) > 1e-30
        d = ((a - b).abs() / b.abs().clamp(min=1e-30)).where(m, torch.zeros_like(a))
        q = torch.quantile(d[:, ::7].float(), 0.99, dim=1)   # p99 on a 1-in-7 subsample
        return d.amax(1).tolist(), q.tolist()

    def named(mx, p99, names):
        return {n: [float(f"{x:.3g}"), float(f"{y:.3g}")] for n, x, y in zip(names, mx, p99)}

    names_long = [n for n in synth.LONG[:42] if n not in ("ox_passive", "h2so4")]
    idx_long = [L[n] for n in names_long]
    ev = []
    one_step = None
    t_wall = time.time()
    for k in range(nsteps):
        sza.copy_(s2[k % 96].expand(shb))
        day = sza[0] <= 95.0
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        e[0].record(stream)
        net.guess(grid, inp.dt, sp["T"], sp["p"], sza, sp["qj1"], sp["hc"], gq, gh)
        step.surrogate_device(grid, inp.dt, accept_tol, tol, maxiter, *args(sp), sp["qj1"],
                              sp["hc"], *aux(sp), gq, gh, iters=it_s, fuse_conserv=True)
        e[1].record(stream)
        e[2].record(stream)
        step.converge_device(grid, inp.dt, tol, maxiter, *args(fp), fp["qj1"], fp["hc"],
                             *aux(fp), iters=it_f, fuse_conserv=True)
        e[3].record(stream)
        ev.append(e)
        acc = (it_s == 0).double()
        for name, m in bands.items():
            band_acc[name] += acc[:, m, :].sum()
            band_n[name] += nb * int(m.sum()) * g.nlon
        dmask = day.double().expand(shb)
        dn_acc["day"] += (acc * dmask).sum()
        dn_acc["night"] += (acc * (1.0 - dmask)).sum()
        dn_n["day"] += dmask.sum()
        dn_n["night"] += (1.0 - dmask).sum()
        if k == 0:
            # the network's answer for the first step against the FP64 step from the same state
            mq, pq = rel_stats(gq[idx_long, :nb], fp["qj1"][idx_long, :nb])
            mh, ph = rel_stats(gh[:, :nb], fp["hc"][:, :nb])
            one_step = {"long": named(mq, pq, names_long), "short": named(mh, ph, synth.SHORT)}
        if k % 96 == 95:
            print(f"[surrogate] step {k + 1}/{nsteps}, {time.time() - t_wall:.0f} s",
                  file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ms_sur = sum(e[0].elapsed_time(e[1]) for e in ev)
    ms_fp = sum(e[2].elapsed_time(e[3]) for e in ev)
    mq, pq = rel_stats(sp["qj1"][idx_long, :nb], fp["qj1"][idx_long, :nb])
    mh, ph = rel_stats(sp["hc"][:, :nb], fp["hc"][:, :nb])
    drift = {}
    for name, w, x0 in (("chlorine", cl_w, cl0), ("bromine", br_w, br0)):
        for path, st in (("surrogate", sp), ("fp64", fp)):
            r = (total(st["qj1"], w) / x0 - 1.0).abs()
            drift[f"{name}_{path}"] = {"max": float(r.max()), "mean": float(r.mean())}
    nbox = g.nbox
    accepted = sum(float(v) for v in band_acc.values())
    out = {"workload": f"c2 seed 0 (786,240 boxes, vortex and liquid PSC), {nsteps} x 15 min "
                       "sub-steps with diurnal SZA; surrogate path and FP64 path from the same "
                       "start, conserv + floor after each step",
           "steps": nsteps, "accept_tol": accept_tol, "tol": tol, "maxiter": maxiter,
           "solver": "plain sweeps", "block": block_conv, "refill": refill,
           "accept_rate": accepted / (nbox * nsteps),
           "accept_rate_by_band": {k: float(band_acc[k]) / max(band_n[k], 1) for k in bands},
           "accept_rate_day_night": {k: float(dn_acc[k]) / max(float(dn_n[k]), 1.0) for k in dn_acc},
           "ms_surrogate_path": ms_sur, "ms_fp64_path": ms_fp,
           "box_steps_per_s_surrogate": nbox * nsteps / (ms_sur / 1e3),
           "box_steps_per_s_fp64": nbox * nsteps / (ms_fp / 1e3),
           "speedup_surrogate_over_fp64": ms_fp / ms_sur,
           "one_step_error_max_p99": one_step,
           "end_state_difference_max_p99": {"long": named(mq, pq, names_long),
                                             "short": named(mh, ph, synth.SHORT)},
           "element_drift": drift}
    step.close()
    net.close()
    return out


def transport_section(table, dev, ntr=43, reps=5):
    """Semi-Lagrangian transport interpolation on the model's 180 x 91 x 137 grid
    (2,244,060 points): departure-point cells and weights, then all ntr tracers (the
    model's 43) interpolated out of place; inputs resident in HBM.  Fields: 3 synthetic
    tracers, scaled into 43 (the interpolation's cost does not depend on the values)."""
    import torch

    from reprobus_amd import synth
    from reprobus_amd.chem import ChemistryStep, make_grid

    tr = synth.make_transport(0)
    grid = make_grid(tr.nlon, tr.nlat, tr.niv, tr.niv, 1)
    n = tr.nlon * tr.nlat * tr.niv
    step = ChemistryStep(*table, device=0)
    stream = torch.cuda.current_stream()
    step.set_stream(stream.cuda_stream)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    pos = [t(a) for a in (tr.aaa, tr.bbb, tr.x0, tr.y0, tr.z0, tr.p0)]
    idx = torch.zeros((3, n), dtype=torch.int32, device=dev)
    w = torch.zeros((12, n), dtype=torch.float64, device=dev)
    q3 = t(tr.q)
    q_in = torch.stack([q3[k % 3] * (1.0 + 1e-3 * k) for k in range(ntr)])
    q_out = torch.empty_like(q_in)
    w_ms = _timed(lambda: step.transport_weights_device(grid, *pos, idx, w), reps, stream)
    i_ms = _timed(lambda: step.transport_interp_device(grid, idx, w, q_in, q_out), reps, stream)
    # the back-trajectories of one dynamical step (1 h in 4 sub-steps, as semilag calls
    # backtra) from two synthetic forcing records, with the trajectory slots
    ti = synth.make_trajectory_inputs(0)
    nrapp = 4
    rec = lambda r: {"u": t(r.u), "v": t(r.v), "w": t(r.w), "t": t(r.t), "lnps": t(r.lnps),
                     "day": r.day}
    m1, p1 = rec(ti.m1), rec(ti.p1)
    dep = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(4)]
    trj = [torch.empty((nrapp, tr.niv, tr.nlat, tr.nlon), dtype=torch.float64, device=dev)
           for _ in range(4)]
    bt = lambda: step.backtra_device(grid, pos[0], pos[1], m1, p1, ti.daynum, ti.dtdyn, nrapp,
                                     ti.ntime, ti.nintecmwf, *dep, idx, w, traj=trj)
    bt()
    b_ms = _timed(bt, reps, stream)
    del m1, p1, dep, trj
    step.close()
    # per point and sub-step: the fields pass reads 8 forcing values and writes 6 fields; the
    # trajectory pass reads the 4 position values (after the first sub-step) and writes 4
    # positions, 3 indices, 12 weights and 4 trajectory values
    b_backtra = n * nrapp * ((8 + 6) * 8 + (4 + 4 + 12 + 4) * 8 + 3 * 4)
    b_interp = n * (3 * 4 + 12 * 8 + ntr * 2 * 8)
    b_weights = n * (4 * 8 + 3 * 4 + 12 * 8)
    gbs = lambda b, ms: b / (ms / 1e3) / 1e9
    return {"workload": f"180 x 91 x 137 grid ({n:,} points), {ntr} tracers, departure points of "
                        "a synthetic 1 h wind; inputs resident in HBM",
            "weights_ms": w_ms, "interp_ms": i_ms, "ms": w_ms + i_ms,
            "point_tracers_per_s": n * ntr / ((w_ms + i_ms) / 1e3),
            "backtra": {"workload": f"back-trajectories of every node over 1 h in {nrapp} "
                                    "sub-steps between two synthetic 6-hourly forcing records, "
                                    "with the trajectory slots",
                        "ms": b_ms, "nrapp": nrapp, "point_substeps_per_s": n * nrapp / (b_ms / 1e3),
                        "bytes": b_backtra, "achieved_gbs": gbs(b_backtra, b_ms),
                        "frac": gbs(b_backtra, b_ms) / HBM_PEAK_GBS},
            "semilag_ms": b_ms + i_ms,
            "roofline": {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "interp_achieved": gbs(b_interp, i_ms),
                         "interp_frac": gbs(b_interp, i_ms) / HBM_PEAK_GBS,
                         "interp_bytes": b_interp,
                         "weights_achieved": gbs(b_weights, w_ms),
                         "weights_frac": gbs(b_weights, w_ms) / HBM_PEAK_GBS,
                         "weights_bytes": b_weights}}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the step shards by whole columns with no communication)", file=sys.stderr)
        sys.exit(2)
    import tempfile

    import torch

    from reprobus_amd import autotune, jtable, synth
    from reprobus_amd.chem import ChemistryStep, make_grid

    t_gen = time.time()
    inp = synth.make_c1(seed=args.seed)
    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    dev = torch.device("cuda:0")
    st = _device_state(inp, dev)
    gen_s = time.time() - t_gen
    table = jtable.default_table()

    tuned = {}
    block = args.block
    if block == 0:
        cache = os.path.join(tempfile.gettempdir(), "reprobus_amd_tune.json")
        r = autotune.tune(table, grid, st, mode="plain", cache_path=cache, reps=5, steps=10)
        block = r.block
        tuned["plain"] = {"block": r.block, "ms": r.ms}

    T, p, sza, alt, q, h = (st[k] for k in ("T", "p", "sza", "alt", "qj1", "hc"))
    surf, vsed, p2o, pno3 = (st[k] for k in ("surfarea", "vsed3d", "parth2osed", "parthno3sed"))
    stream = torch.cuda.current_stream()

    def stepper(profile):
        s = ChemistryStep(*table, device=0, block=block, profile=profile)
        s.set_stream(stream.cuda_stream)
        return s, lambda: s.step_device(grid, inp.dt, T, p, sza, alt, q, h, surf, vsed, p2o,
                                        pno3, fuse_conserv=True)

    # the timed steps without the library's per-launch events
    step, run = stepper(False)
    for _ in range(args.warmup):
        run()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        run()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_total = e0.elapsed_time(e1)
    step.close()
    # the per-box kernels' own time (HIP events around each set-up + sweeps pair) over the
    # next steps, for the roofline
    step, run = stepper(True)
    run()
    torch.cuda.synchronize()
    step.stats()                                    # reset the per-launch accumulators
    for _ in range(max(3, min(args.steps, 10))):
        run()
    torch.cuda.synchronize()
    st_ = step.stats()
    finite = bool(torch.isfinite(q).all().item() and torch.isfinite(h).all().item())
    step.close()
    if args.calib_copy_mb > 0:
        n = args.calib_copy_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    nbox = g.nbox
    ms_per_step = ms_total / args.steps
    value = nbox * args.steps / (ms_total / 1e3)
    chem_ms = st_["chem_ms_total"] / max(st_["chem_launches"], 1)
    achieved_gbs = nbox * BYTES_PER_BOX / (chem_ms / 1e3) / 1e9
    traffic = TRAFFIC_BYTES.get(block)
    roofline = {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": f"k_setup<{block}> + k_chem<{block}> (the step's per-box kernels)",
                "kernel_ms": chem_ms, "bytes_per_box": BYTES_PER_BOX,
                "traffic_gbs": (traffic / (chem_ms / 1e3) / 1e9) if traffic else None,
                "fp64": {"algorithmic_tflops": nbox * FLOP_PER_BOX / (chem_ms / 1e3) / 1e12,
                         "instruction_tflops": FP64_INSTR_FLOP / (chem_ms / 1e3) / 1e12,
                         "peak_tflops": FP64_PEAK_TFLOPS, "flop_per_box": FLOP_PER_BOX,
                         "note": "instruction_tflops from the SQ_INSTS_VALU_*_F64 counters "
                                 "(profiles/k_chem_c1_pmc.md)"}}

    extras = {}
    want = set() if args.no_extras else set(s for s in args.extras.split(",") if s)
    conv_shape = None
    newton_shape = (60, 1)
    if args.conv_shape and args.newton_shape:
        conv_shape = tuple(int(v) for v in args.conv_shape.split(","))
        newton_shape = tuple(int(v) for v in args.newton_shape.split(","))
        tuned["converged"] = {"block": conv_shape[0], "refill": conv_shape[1], "given": True}
        tuned["newton"] = {"sweeps_first": newton_shape[0], "waves": newton_shape[1],
                           "given": True}
    elif want & {"converged", "surrogate"}:
        try:
            cache = os.path.join(tempfile.gettempdir(), "reprobus_amd_tune.json")
            c2 = _device_state(synth.make_c2(1), dev)
            r = autotune.tune(table, grid, c2, mode="converged", reps=2, cache_path=cache)
            tuned["converged"] = {"block": r.block, "refill": r.refill, "ms": r.ms}
            conv_shape = (r.block, r.refill)
            rn = autotune.tune_newton(table, grid, c2, block=r.block, refill=r.refill, reps=2,
                                      cache_path=cache)
            tuned["newton"] = {"sweeps_first": rn.sweeps_first, "waves": rn.waves, "ms": rn.ms}
            newton_shape = (rn.sweeps_first, rn.waves)
            del c2
        except Exception as exc:  # reported beside the headline, never instead of it
            extras["converged"] = {"error": repr(exc)}
    if "converged" in want and conv_shape:
        try:
            extras["converged"] = converged_section(table, *conv_shape, dev, newton_shape,
                                                    max_newton=args.max_newton,
                                                    order=args.conv_order)
        except Exception as exc:
            extras["converged"] = {"error": repr(exc)}
    if "surrogate" in want and conv_shape:
        try:
            extras["surrogate"] = surrogate_section(table, *conv_shape, dev,
                                                    nsteps=args.surrogate_steps)
        except Exception as exc:
            extras["surrogate"] = {"error": repr(exc)}
    if "day" in want:
        try:
            extras["day"] = day_section(table, block, dev)
        except Exception as exc:
            extras["day"] = {"error": repr(exc)}
    if "transport" in want:
        try:
            extras["transport"] = transport_section(table, dev)
        except Exception as exc:
            extras["transport"] = {"error": repr(exc)}
    if "mlp" in want:
        try:
            extras["mlp"] = mlp_section(dev)
        except Exception as exc:
            extras["mlp"] = {"error": repr(exc)}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(inp, args.cpu_procs or default_cpu_procs(), args.cpu_nrep)
        except Exception as exc:  # the baseline is reported, never the target
            cpu = {"error": str(exc)}

    line = {"metric": METRIC, "value": value, "unit": "box-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "c1 plain FP64 chemistry step: 144 lon x 91 lat x 60 levels "
                                   "= 786,240 boxes, dt 900 s, 5 fixed sweeps (the reference "
                                   "scheme), conserv+floor fused, state resident in HBM",
                       "boxes": nbox, "solver": "plain", "niter": 5, "seed": args.seed,
                       "block": block, "autotune": tuned, "finite": finite,
                       "gen_s": round(gen_s, 1)},
            "roofline": roofline, "cpu_baseline": cpu, **extras}
    if cpu and "value" in cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
