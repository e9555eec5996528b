"""GPU surrogate step: a surrogate's answer kept where one FP64 sweep from it changes no
species by more than accept_tol (its implicit-step residual), the converged solve from it
everywhere else -- against the restatement's surrogate step on the same answer.

Bar: the same boxes accepted; 1e-9 relative per species above the 1e-30 floor, as for the
converged mode.
"""
import numpy as np
import pytest

from conftest import PARITY_TOL, box_rel, max_rel_by_species, rel_err

pytestmark = pytest.mark.gpu

ACCEPT = 1e-8


@pytest.fixture(scope="module")
def step(table):
    from reprobus_amd.chem import ChemistryStep

    s = ChemistryStep(*table, device=0)
    yield s
    s.close()


def _dev(inp):
    import torch

    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    g = inp.grid
    d = {"T": t(inp.T), "p": t(inp.p), "sza": t(inp.sza), "alt": t(inp.alt), "qj1": t(inp.qj1),
         "hc": t(inp.hc), "surfarea": torch.zeros(g.shape3, dtype=torch.float64, device=dev)}
    shb = (g.nivbas, g.nlat, g.nlon)
    for k in ("vsed3d", "parth2osed", "parthno3sed"):
        d[k] = torch.zeros(shb, dtype=torch.float64, device=dev)
    d["iters"] = torch.zeros(shb, dtype=torch.int32, device=dev)
    return d


def surrogate_gpu(step, inp, sur, tol, maxiter=5000, accel=0, accept_tol=ACCEPT):
    import torch
    from reprobus_amd.chem import make_grid

    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    d = _dev(inp)
    sq = torch.from_numpy(np.ascontiguousarray(sur[0])).to("cuda:0")
    sh = torch.from_numpy(np.ascontiguousarray(sur[1])).to("cuda:0")
    step.set_stream(torch.cuda.current_stream().cuda_stream)
    print(f"[surrogate] grid {g.nlon}x{g.nlat}x{g.nivbas} accel {accel} accept {accept_tol:g}",
          flush=True)
    step.surrogate_device(grid, inp.dt, accept_tol, tol, maxiter, d["T"], d["p"], d["sza"],
                          d["alt"], d["qj1"], d["hc"], d["surfarea"], d["vsed3d"],
                          d["parth2osed"], d["parthno3sed"], sq, sh, iters=d["iters"],
                          fuse_conserv=False, accel=accel)
    torch.cuda.synchronize()
    step.set_stream(None)
    return {k: v.cpu().numpy() for k, v in d.items()}


@pytest.mark.parametrize("accel", [0])
def test_surrogate_keeps_fixed_points_and_solves_the_rest(step, table, accel):
    """Half the boxes get their fixed point (to 1e-12) as the surrogate's answer, the rest
    the same 1e-4 off: the GPU keeps exactly the first half, as the restatement does, and
    solves the rest to the restatement's answer."""
    import cport
    from reprobus_amd import synth

    inp = synth.make_c1(seed=5, nlon=16, nlat=9)
    base = cport.chem_converge(inp, *table, tol=1e-12, maxiter=20000, conserv=False,
                               cr_psc=True)
    live = base["iters"] > 0
    rng = np.random.default_rng(4)
    keep = (rng.random(live.shape) < 0.5) & live
    gq, gh = base["qj1_chem"].copy(), base["hc_chem"].copy()
    gq[:42, ~keep] *= np.exp(rng.normal(0, 1e-4, gq[:42, ~keep].shape))
    gh[:, ~keep] *= np.exp(rng.normal(0, 1e-4, gh[:, ~keep].shape))
    out = surrogate_gpu(step, inp, (gq, gh), tol=1e-12, accel=accel)
    ref = cport.chem_surrogate(inp, *table, (gq, gh), accept_tol=ACCEPT, tol=1e-12,
                               maxiter=5000, conserv=False, cr_psc=True, accel=accel,
                               aa_start=step.conv_aa_start)
    it, rit = out["iters"], ref["iters"]
    both = (it >= 0) & (rit >= 0)
    w = float(box_rel(out["qj1"], out["hc"], ref["qj1_chem"], ref["hc_chem"])[both].max())
    print(f"[surrogate] accel {accel}: kept {int((it == 0).sum())}/{it.size} (restatement "
          f"{int((rit == 0).sum())}), same count {float((it == rit).mean()):.4%}, worst rel "
          f"{w:.2e}", flush=True)
    assert np.array_equal(it == 0, keep)
    assert np.array_equal(rit == 0, keep)
    assert rel_err(out["qj1"][:42], gq[:42])[:, keep].max() < 1e-15
    assert rel_err(out["hc"], gh)[:, keep].max() < 1e-15
    assert float((it == rit).mean()) > 0.99
    assert both.mean() > 0.99
    assert w < PARITY_TOL


def test_surrogate_with_the_learned_network(step, table):
    """The trained network's answer on a c2 subgrid (vortex): the GPU and the restatement
    accept the same boxes and agree everywhere.  The accept rate is printed (a BF16
    network's answer is far from a 1e-8 residual)."""
    import cport
    import torch
    from reprobus_amd import guess, synth
    from reprobus_amd.chem import make_grid
    from reprobus_amd.mlp import GuessMLP

    inp = synth.make_c2(1, nlon=24, nlat=13)
    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    net = GuessMLP(0)
    try:
        assert guess.load(net), "no trained weights"
        d = _dev(inp)
        gq, gh = torch.zeros_like(d["qj1"]), torch.zeros_like(d["hc"])
        net.guess(grid, inp.dt, d["T"], d["p"], d["sza"], d["qj1"], d["hc"], gq, gh)
        torch.cuda.synchronize()
        sur = (gq.cpu().numpy(), gh.cpu().numpy())
    finally:
        net.close()
    out = surrogate_gpu(step, inp, sur, tol=1e-10, maxiter=2000)
    ref = cport.chem_surrogate(inp, *table, sur, accept_tol=ACCEPT, tol=1e-10, maxiter=2000,
                               conserv=False, cr_psc=True)
    it, rit = out["iters"], ref["iters"]
    both = (it >= 0) & (rit >= 0)
    w = float(box_rel(out["qj1"], out["hc"], ref["qj1_chem"], ref["hc_chem"])[both].max())
    print(f"[surrogate] learned network on c2 {g.nlon}x{g.nlat}: accepted {float((it == 0).mean()):.3%}"
          f" (restatement {float((rit == 0).mean()):.3%}), fallback boxes {int((it < 0).sum())}, "
          f"same count {float((it == rit).mean()):.4%}, worst rel {w:.2e}", flush=True)
    assert np.array_equal(it == 0, rit == 0)
    assert (np.asarray(it < 0) != np.asarray(rit < 0)).sum() <= 2
    assert float((it == rit).mean()) > 0.99
    assert w < PARITY_TOL


def test_surrogate_rejects_bad_arguments(step):
    from reprobus_amd import synth
    from reprobus_amd._lib import RbError

    inp = synth.make_c1(seed=3, nlon=4, nlat=3)
    sur = (inp.qj1, inp.hc)
    for bad in (0.0, -1e-8, float("inf")):
        with pytest.raises(RbError):
            surrogate_gpu(step, inp, sur, tol=1e-10, accept_tol=bad)
    # the fallback sweeps start from the surrogate's answer, a guess: no Anderson mixing
    with pytest.raises(RbError, match="accel with a guess"):
        surrogate_gpu(step, inp, sur, tol=1e-10, accel=2)
