"""What the trained network gains over the method without it: the converged mode (plain
sweeps to tol) from the incoming state and from the trained network's guess, on a c2
subgrid of the evaluation seed (seed 1, polar vortex).  Both must reach the same fixed point
(1e-9 per species above the 1e-30 floor on boxes that converge on both paths, as the
converged-mode parity bar), and the guess must take fewer sweeps on average.  The gain is
printed: it is small, because the sweep iteration converges linearly (each sweep divides the
remaining error by a roughly constant factor), so a start that is k times closer saves only
about log(k) / log(that factor) sweeps.
"""
import numpy as np
import pytest

from conftest import PARITY_TOL, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-11


def test_trained_guess_gain(table):
    import torch
    from reprobus_amd import guess, synth
    from reprobus_amd.chem import ChemistryStep, make_grid
    from reprobus_amd.mlp import GuessMLP

    inp = synth.make_c2(1, nlon=48, nlat=26)
    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    T, p, sza, alt = t(inp.T), t(inp.p), t(inp.sza), t(inp.alt)
    shb = (g.nivbas, g.nlat, g.nlon)
    step = ChemistryStep(*table, device=0)
    net = GuessMLP(0)
    try:
        assert guess.load(net), "no trained weights"
        gq, gh = torch.zeros_like(t(inp.qj1)), torch.zeros_like(t(inp.hc))
        net.guess(grid, inp.dt, T, p, sza, t(inp.qj1), t(inp.hc), gq, gh)
        out = {}
        for name, use in (("incoming", False), ("guess", True)):
            q, h = t(inp.qj1), t(inp.hc)
            surf = torch.zeros(g.shape3, dtype=torch.float64, device=dev)
            aux = [torch.zeros(shb, dtype=torch.float64, device=dev) for _ in range(3)]
            it = torch.zeros(shb, dtype=torch.int32, device=dev)
            print(f"[guess gain] c2 seed 1 {g.nlon}x{g.nlat}: from {name}", flush=True)
            step.converge_device(grid, inp.dt, TOL, 2000, T, p, sza, alt, q, h, surf, *aux,
                                 guess_qj1=gq if use else None, guess_hc=gh if use else None,
                                 iters=it, fuse_conserv=False)
            torch.cuda.synchronize()
            out[name] = (q.cpu().numpy(), h.cpu().numpy(), it.cpu().numpy())
        guess_q, guess_h = gq.cpu().numpy(), gh.cpu().numpy()
    finally:
        net.close()
        step.close()
    (q0, h0, i0), (q1, h1, i1) = out["incoming"], out["guess"]
    live = (i0 > 0) & (i1 > 0)
    d = np.maximum(rel_err(q1[:42], q0[:42]).max(axis=0), rel_err(h1, h0).max(axis=0))
    # how far each starting point is from the fixed point (short-lived species, median)
    start_in = np.median(rel_err(inp.hc, h0)[:, live])
    start_nn = np.median(rel_err(guess_h, h0)[:, live])
    m0, m1 = float(i0[live].mean()), float(i1[live].mean())
    print(f"[guess gain] sweeps {m0:.2f} -> {m1:.2f} ({(1 - m1 / m0):.1%} fewer); fallback "
          f"boxes {int((i0 < 0).sum())} -> {int((i1 < 0).sum())}; boxes over 1e-9 "
          f"{int((d[live] > PARITY_TOL).sum())} of {int(live.sum())} (max {d[live].max():.2e}); "
          f"median start error of the short-lived species: incoming {start_in:.2e}, "
          f"guess {start_nn:.2e}", flush=True)
    assert live.mean() > 0.99
    assert m1 < m0
    # the same fixed point: all but a few boxes (night chemistry has more than one fixed point
    # and a box can stop within tol of the fixed point from either side) agree to 1e-9
    assert (d[live] > PARITY_TOL).mean() < 1e-3
    assert np.percentile(d[live], 99) < PARITY_TOL
