"""Training data, training and use of the learned initial guess (reprobus_amd.mlp).

Run ``python -m reprobus_amd.guess --train`` to regenerate ``data/guess_mlp.npz`` (the
96-step c2 day of seed 0, 42,000 box-steps sampled per step, 120 epochs of Adam at batch
8192 and lr 1e-3, on the library's kernels) and ``--eval`` for the per-species error of
the guess on the seed-1 day; ``--report FILE`` writes both as markdown.

Data: box-steps of the c2 day (a 96-step day over the 786,240-box grid with a polar
vortex and diurnal SZA) solved by the converged-mode GPU solver from the incoming state.
Each pair is (features of the incoming state, log increments to the raw fixed point of
the sweeps, before conserv).  The day advances with conserv applied after each step, as
the model does.

Training: Adam (lr 1e-3, batch 8192) on the MSE of normalised log increments, entirely
in the library's HIP kernels.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import synth
from .chem import ChemistryStep, make_grid
from .mlp import LONG_SLOTS, NIN, NOUT, GuessMLP, init_params, log_increments

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "guess_mlp.npz")


class DeviceDay:
    """Device-resident state of a day on one grid (torch tensors on cuda:0)."""

    def __init__(self, inp: synth.StepInputs):
        import torch

        self.inp = inp
        g = inp.grid
        self.grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
        self.nbox = g.nbox
        dev = torch.device("cuda:0")
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
        self.T, self.p, self.alt = t(inp.T), t(inp.p), t(inp.alt)
        self.sza = t(inp.sza)
        self.q, self.h = t(inp.qj1), t(inp.hc)
        self.surf = torch.zeros(g.shape3, dtype=torch.float64, device=dev)
        shb = (g.nivbas, g.nlat, g.nlon)
        self.aux = [torch.zeros(shb, dtype=torch.float64, device=dev) for _ in range(3)]
        self.iters = torch.zeros(shb, dtype=torch.int32, device=dev)
        self.gq, self.gh = torch.zeros_like(self.q), torch.zeros_like(self.h)
        self.X = torch.empty((self.nbox, NIN), dtype=torch.float32, device=dev)

    def set_step(self, k: int) -> None:
        import torch

        s = synth.c2_sza(self.inp.meta, k)
        self.sza.copy_(torch.from_numpy(s))

    def converge(self, chem: ChemistryStep, tol: float, maxiter: int, guess: bool = False,
                 fuse_conserv: bool = False) -> None:
        chem.converge_device(self.grid, self.inp.dt, tol, maxiter, self.T, self.p, self.sza,
                             self.alt, self.q, self.h, self.surf, *self.aux,
                             guess_qj1=self.gq if guess else None,
                             guess_hc=self.gh if guess else None, iters=self.iters,
                             fuse_conserv=fuse_conserv)

    def predict(self, net: GuessMLP) -> None:
        net.guess(self.grid, self.inp.dt, self.T, self.p, self.sza, self.q, self.h, self.gq, self.gh)


def generate_pairs(chem: ChemistryStep, net: GuessMLP, inp: synth.StepInputs, nsteps: int,
                   per_step: int, tol: float, maxiter: int, seed: int = 0, log=print):
    """(X, Y) device tensors of nsteps * per_step box-steps sampled along the day."""
    import torch

    day = DeviceDay(inp)
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(seed)
    Xs, Ys = [], []
    nb = day.grid.nivbas
    for k in range(nsteps):
        day.set_step(k)
        net.features(day.grid, inp.dt, day.T, day.p, day.sza, day.q, day.h, day.X)
        q0, h0 = day.q.clone(), day.h.clone()
        day.converge(chem, tol, maxiter)
        Y = log_increments(q0, h0, day.q, day.h, nb)
        # boxes that converged (negative counts hold the fixed-sweep fallback, no fixed point)
        perm = torch.randperm(day.nbox, generator=gen, device="cuda:0")
        idx = perm[day.iters.reshape(-1)[perm] > 0][:per_step]
        Xs.append(day.X[idx].clone())
        Ys.append(Y[idx].clone())
        chem.conserv_device(day.grid, day.q, day.h)
        if log and (k % 16 == 0 or k == nsteps - 1):
            torch.cuda.synchronize()
            it = day.iters.float()
            log(f"[guess] day step {k}: sweeps mean {it.mean().item():.1f} max {int(it.max().item())}")
    return torch.cat(Xs), torch.cat(Ys)


def normalisation(X, Y):
    """Per-feature mean and spread (spread floored so constant features map to 0)."""
    mu_x, sd_x = X.mean(0), X.std(0)
    mu_y, sd_y = Y.mean(0), Y.std(0)
    sd_x = sd_x.clamp(min=1e-3)
    sd_y = sd_y.clamp(min=1e-3)
    return [v.cpu().numpy().astype(np.float32) for v in (mu_x, sd_x, mu_y, sd_y)]


def train(net: GuessMLP, X, Y, epochs: int, batch: int = 8192, lr: float = 1e-3, seed: int = 0,
          log=print) -> list:
    import torch

    n = X.shape[0]
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(seed)
    history = []
    nb = n // batch
    for ep in range(epochs):
        perm = torch.randperm(n, generator=gen, device="cuda:0")
        Xp, Yp = X[perm], Y[perm]
        t0 = time.time()
        loss = None
        for b in range(nb):
            s = slice(b * batch, (b + 1) * batch)
            loss = net.train_step(Xp[s], Yp[s], lr, want_loss=(b == nb - 1))
        torch.cuda.synchronize()
        history.append(loss)
        if log:
            log(f"[guess] epoch {ep}: loss {loss:.5f} ({nb} steps, {time.time() - t0:.2f} s)")
    return history


def save(net: GuessMLP, path: str = WEIGHTS) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=net.get_params(), **dict(zip(("in_mu", "in_sd", "out_mu", "out_sd"),
                                                       net._norm)))


def load(net: GuessMLP, path: str = WEIGHTS) -> bool:
    if not os.path.exists(path):
        return False
    with np.load(path, allow_pickle=False) as z:
        net.set_params(z["params"])
        net.set_norm(z["in_mu"], z["in_sd"], z["out_mu"], z["out_sd"])
    return True


def evaluate(chem: ChemistryStep, net: GuessMLP, seed: int = 1, steps=(0, 48), tol: float = 1e-11,
             maxiter: int = 2000) -> dict:
    """Per-species error of the guess against the converged answer on the c2 day of `seed`
    at the given sub-steps (the day advanced with conserv in between): the median and p99 of
    |log(guess / answer)| over the boxes that converged, long- and short-lived."""
    import torch

    inp = synth.make_c2(seed)
    day = DeviceDay(inp)
    nb = day.grid.nivbas
    out = {}
    for k in range(max(steps) + 1):
        day.set_step(k)
        if k in steps:
            day.predict(net)
            gq, gh = day.gq.clone(), day.gh.clone()
        day.converge(chem, tol, maxiter)
        if k in steps:
            ok = (day.iters > 0).reshape(-1)
            lg = lambda a: torch.log(torch.clamp(a[:, :nb], min=1e-30)).reshape(a.shape[0], -1)
            idx = torch.tensor(LONG_SLOTS, device=gq.device)
            err = torch.cat([(lg(gq.index_select(0, idx)) - lg(day.q.index_select(0, idx))),
                             (lg(gh) - lg(day.h))], 0).abs()[:, ok]
            names = [synth.LONG[s] for s in LONG_SLOTS] + [f"hc:{n}" for n in synth.SHORT]
            med = torch.quantile(err[:, :200000].float(), 0.5, dim=1).cpu().numpy()
            p99 = torch.quantile(err[:, :200000].float(), 0.99, dim=1).cpu().numpy()
            out[k] = {n: (float(np.exp(m)), float(np.exp(q))) for n, m, q in zip(names, med, p99)}
        chem.conserv_device(day.grid, day.q, day.h)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m reprobus_amd.guess",
                                 description="train / evaluate the learned initial guess")
    ap.add_argument("--train", action="store_true", help="generate the data and train")
    ap.add_argument("--eval", action="store_true", help="per-species error on the seed-1 day")
    ap.add_argument("--seed", type=int, default=0, help="c2 day of the training data")
    ap.add_argument("--steps", type=int, default=96)
    ap.add_argument("--per-step", type=int, default=42000)
    ap.add_argument("--epochs", type=int, default=120)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-11)
    ap.add_argument("--weights", default=WEIGHTS, help="where the weights are read / written")
    ap.add_argument("--report", default=None, help="markdown report of the loss and errors")
    args = ap.parse_args(argv)
    if not (args.train or args.eval):
        ap.error("nothing to do: give --train and/or --eval")
    import torch

    from . import jtable

    lines = []
    log = lambda m: (print(m, flush=True), lines.append(m))
    chem = ChemistryStep(*jtable.default_table(), device=0)
    net = GuessMLP(0)
    stream = torch.cuda.current_stream()
    chem.set_stream(stream.cuda_stream)
    net.set_stream(stream.cuda_stream)
    report = {}
    if args.train:
        t0 = time.time()
        net.set_params(init_params(args.seed))
        X, Y = generate_pairs(chem, net, synth.make_c2(args.seed), args.steps, args.per_step,
                              args.tol, 2000, seed=args.seed, log=log)
        log(f"[guess] {X.shape[0]:,} pairs in {time.time() - t0:.1f} s")
        net.set_norm(*normalisation(X, Y))
        hist = train(net, X, Y, args.epochs, args.batch, args.lr, args.seed, log=log)
        save(net, args.weights)
        log(f"[guess] weights written to {args.weights} ({time.time() - t0:.1f} s in all)")
        report["loss"] = hist
        report["pairs"] = int(X.shape[0])
        del X, Y
    elif not load(net, args.weights):
        raise SystemExit(f"no weights at {args.weights}")
    if args.eval:
        report["eval"] = evaluate(chem, net, tol=args.tol)
        for k, tab in report["eval"].items():
            worst = sorted(tab.items(), key=lambda kv: -kv[1][1])[:5]
            log(f"[guess] seed 1, step {k}: worst p99 factors " +
                ", ".join(f"{n} {v[1]:.3g}" for n, v in worst))
    if args.report:
        with open(args.report, "w") as f:
            f.write("# Learned initial guess: training and per-species error\n\n")
            cmd = " ".join(argv if argv is not None else sys.argv[1:])
            f.write(f"Command: `python -m reprobus_amd.guess {cmd}`\n\n")
            if "loss" in report:
                f.write(f"Training pairs: {report['pairs']:,} (c2 day, seed {args.seed}, "
                        f"{args.steps} steps)\n\n| epoch | normalised MSE |\n|---|---|\n")
                for i, v in enumerate(report["loss"]):
                    f.write(f"| {i} | {v:.5f} |\n")
            for k, tab in report.get("eval", {}).items():
                f.write(f"\n## seed 1, sub-step {k}: |guess / converged answer| per species\n\n"
                        "| species | median factor | p99 factor |\n|---|---|---|\n")
                for n, (m, q) in tab.items():
                    f.write(f"| {n} | {m:.4g} | {q:.4g} |\n")
    chem.close()
    net.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
