"""Benchmark: REPROBUS chemistry step on one MI355X.

Workload: the full 144 x 91 x 60 grid (786,240 boxes), one 900 s chemistry step
per timed step (fixed 5 sweeps in FP64, conserv + floor fused), state resident
in HBM; successive steps integrate the evolving state.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=16)
    ap.add_argument("--cpu-nrep", type=int, default=8)
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="after the timed region, one device copy of this many MB (a kernel of "
                         "known traffic, to scale profiler byte counters)")
    return ap.parse_args()


def cpu_baseline(inp, nproc, nrep):
    """Reference Fortran (flang build) on a bounded sample: 180 columns spread over the
    grid (one reference latitude row, 94 levels), nrep successive steps per process."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    from reprobus_amd import synth

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
    nproc = max(1, min(nproc, os.cpu_count() or 1))
    r = refrun.time_reference(sub, nrep=nrep, nproc=nproc)
    return {"value": r["box_steps_per_s"], "unit": "box-steps/s", "cores": nproc,
            "kind": "reference",
            "sample": (f"flang build of the reference chem3d (src/reprobus_l137_20230723.f + "
                       f"jno.f90), {nproc} concurrent single-threaded processes, each "
                       f"{nrep} successive steps of 180 columns sampled across the 144x91 grid "
                       f"x 94 reference levels ({r['calls']} row calls, {r['chem3d_s']:.1f} "
                       f"CPU-s); rate summed over processes, timed around chem3d"),
            "per_core": r["per_core_box_steps_per_s"]}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the step shards by whole columns with no communication)", file=sys.stderr)
        sys.exit(2)
    import torch

    from reprobus_amd import jtable, synth
    from reprobus_amd.chem import ChemistryStep, make_grid

    t_gen = time.time()
    inp = synth.make_c1(seed=args.seed)
    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    T, p, sza, alt, q, h = t(inp.T), t(inp.p), t(inp.sza), t(inp.alt), t(inp.qj1), t(inp.hc)
    surf = torch.zeros(g.shape3, dtype=torch.float64, device=dev)
    shb = (g.nivbas, g.nlat, g.nlon)
    vsed, p2o, pno3 = (torch.zeros(shb, dtype=torch.float64, device=dev) for _ in range(3))
    gen_s = time.time() - t_gen

    step = ChemistryStep(*jtable.default_table(), device=0, block=args.block, profile=True)
    stream = torch.cuda.current_stream()
    step.set_stream(stream.cuda_stream)
    run = lambda: step.step_device(grid, inp.dt, T, p, sza, alt, q, h, surf, vsed, p2o, pno3,
                                   fuse_conserv=True)
    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    step.stats()                                    # reset the per-launch accumulators
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        run()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_total = e0.elapsed_time(e1)
    st = step.stats()
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
    chem_ms = st["chem_ms_total"] / max(st["chem_launches"], 1)
    achieved_gbs = nbox * BYTES_PER_BOX / (chem_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": None,
                "kernel": "k_chem", "kernel_ms": chem_ms, "bytes_per_box": BYTES_PER_BOX}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(inp, args.cpu_procs, args.cpu_nrep)
        except Exception as exc:  # the baseline is reported, never the target
            cpu = {"error": str(exc)}

    line = {"metric": METRIC, "value": value, "unit": "box-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "c1 plain FP64 chemistry step: 144 lon x 91 lat x 60 levels "
                                   "= 786,240 boxes, dt 900 s, 5 fixed sweeps, conserv+floor fused, "
                                   "state resident in HBM",
                       "boxes": nbox, "solver": "plain", "niter": 5, "seed": args.seed,
                       "block": args.block, "finite": finite, "gen_s": round(gen_s, 1)},
            "roofline": roofline, "cpu_baseline": cpu}
    if cpu and "value" in cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
