"""Training data, training and use of the learned initial guess (reprobus_amd.mlp).

Data: box-steps of the c2 day (a 96-step day over the 786,240-box grid with a polar
vortex and diurnal SZA) solved by the converged-mode GPU solver from the incoming state.
Each pair is (features of the incoming state, log increments to the raw fixed point of
the sweeps, before conserv).  The day advances with conserv applied after each step, as
the model does.

Training: Adam (lr 1e-3, batch 8192) on the MSE of normalised log increments, entirely
in the library's HIP kernels.
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import synth
from .chem import ChemistryStep, make_grid
from .mlp import NIN, NOUT, GuessMLP, log_increments

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
