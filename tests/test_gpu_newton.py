"""GPU Newton solve of the converged mode's fixed point (rb_chem_newton_dev: Newton
iterations with the generated sparse LU, a plain sweep as the check, plain sweeps for the
boxes Newton hands over).

Its answer is the fixed point of the reference's sweep map (src/reprobus_l137_20230723.f:
1350-2372 swept to convergence), so it is checked against the restatement's converged mode
run to 1e-14 and against the GPU's own plain converged solve, on the boxes each path
converges (a box whose sweeps never settle -- a kink cycle, a degenerate night NO2 -- gets
the fixed-sweep answer on that path), at the 1e-9 bar.
"""
import numpy as np
import pytest

from conftest import PARITY_TOL, box_rel

pytestmark = pytest.mark.gpu

TOL = 1e-11


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


def solve_gpu(step, inp, how, tol=TOL, guess=None, max_newton=20, maxiter=2000, fuse=False,
              sweeps_first=0):
    import torch
    from reprobus_amd.chem import make_grid

    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    d = _dev(inp)
    gq = gh = None
    if guess is not None:
        gq = torch.from_numpy(np.ascontiguousarray(guess[0])).to("cuda:0")
        gh = torch.from_numpy(np.ascontiguousarray(guess[1])).to("cuda:0")
    args = [d[k] for k in ("T", "p", "sza", "alt", "qj1", "hc", "surfarea", "vsed3d",
                           "parth2osed", "parthno3sed")]
    print(f"[newton] {how} on {g.nlon}x{g.nlat}x{g.nivbas}", flush=True)
    step.set_stream(torch.cuda.current_stream().cuda_stream)
    if how == "newton":
        step.newton_device(grid, inp.dt, tol, max_newton, maxiter, *args, guess_qj1=gq,
                           guess_hc=gh, iters=d["iters"], fuse_conserv=fuse,
                           sweeps_first=sweeps_first)
    else:
        step.converge_device(grid, inp.dt, tol, maxiter, *args, guess_qj1=gq, guess_hc=gh,
                             iters=d["iters"], fuse_conserv=fuse)
    torch.cuda.synchronize()
    step.set_stream(None)
    return {k: v.cpu().numpy() for k, v in d.items()}


SOLVED, RESUMED = 10000, 20000


def _summary(tag, it):
    """Boxes the first sweeps converged, Newton solved, the last sweeps finished."""
    sw = (it > 0) & (it < SOLVED)
    nw = (it >= SOLVED) & (it < RESUMED)
    rs = it >= RESUMED
    n = it[nw] - SOLVED
    h = np.bincount(n.ravel(), minlength=8)[1:8] if nw.any() else np.zeros(7, int)
    print(f"[newton] {tag}: sweeps {int(sw.sum())}, Newton {int(nw.sum())} (iterations 1..7: "
          f"{h.tolist()}, max {int(n.max()) if nw.any() else 0}), handed back to sweeps "
          f"{int(rs.sum())}, fixed-sweep fallback {int((it < 0).sum())} of {it.size}",
          flush=True)
    return nw, rs


@pytest.fixture(scope="module")
def c2(table):
    import cport
    from reprobus_amd import synth

    inp = synth.make_c2(1, nlon=24, nlat=13)
    tight = cport.chem_converge(inp, *table, tol=1e-14, maxiter=50000, conserv=False,
                                cr_psc=True)
    return inp, tight


@pytest.mark.parametrize("sweeps_first", [0, 30])
def test_newton_reaches_the_fixed_point(step, c2, sweeps_first):
    inp, tight = c2
    out = solve_gpu(step, inp, "newton", sweeps_first=sweeps_first)
    it = out["iters"]
    nw, sw = _summary(f"c2 24x13 from the incoming state, {sweeps_first} sweeps first", it)
    both = (it > 0) & (tight["iters"] > 0)
    d = box_rel(out["qj1"], out["hc"], tight["qj1_chem"], tight["hc_chem"])
    print(f"[newton] vs the restatement at 1e-14: max {d[both].max():.2e}, "
          f"{int((d[both] > PARITY_TOL).sum())} boxes above {PARITY_TOL:g}; Newton boxes max "
          f"{d[nw & both].max():.2e}", flush=True)
    assert (it < 0).mean() < 1e-3 and sw.mean() < 1e-2
    assert d[nw & both].max() < PARITY_TOL
    assert d[both].max() < PARITY_TOL
    assert np.isfinite(out["qj1"]).all() and np.isfinite(out["hc"]).all()


@pytest.mark.parametrize("sweeps_first", [0, 30])
def test_newton_agrees_with_the_plain_converged_solve(step, c2, sweeps_first):
    inp, _ = c2
    new = solve_gpu(step, inp, "newton", sweeps_first=sweeps_first)
    plain = solve_gpu(step, inp, "plain")
    both = (new["iters"] > 0) & (plain["iters"] > 0)
    d = box_rel(new["qj1"], new["hc"], plain["qj1"], plain["hc"])
    print(f"[newton] vs GPU plain sweeps (tol {TOL:g}): {int(both.sum())} boxes converged on "
          f"both, max {d[both].max():.2e}, {int((d[both] > PARITY_TOL).sum())} above "
          f"{PARITY_TOL:g}; plain sweeps mean {np.abs(plain['iters']).mean():.1f}", flush=True)
    assert both.mean() > 0.99
    assert d[both].max() < PARITY_TOL


def test_newton_from_a_guess(step, c2):
    """From a guess 5 % off the fixed point (a learned guess's role): the same fixed point in
    fewer iterations than from the incoming state."""
    inp, tight = c2
    rng = np.random.default_rng(0)
    gq = tight["qj1_chem"] * np.exp(rng.normal(0, 0.05, tight["qj1_chem"].shape))
    gh = tight["hc_chem"] * np.exp(rng.normal(0, 0.05, tight["hc_chem"].shape))
    base = solve_gpu(step, inp, "newton")
    out = solve_gpu(step, inp, "newton", guess=(gq, gh))
    nw, _ = _summary("c2 24x13 from a guess 5 % off", out["iters"])
    both = (out["iters"] > 0) & (tight["iters"] > 0)
    d = box_rel(out["qj1"], out["hc"], tight["qj1_chem"], tight["hc_chem"])
    bn = (base["iters"] >= SOLVED) & (base["iters"] < RESUMED) & nw
    print(f"[newton] from the guess: vs the restatement max {d[both].max():.2e}; mean Newton "
          f"iterations {base['iters'][bn].mean() - SOLVED:.2f} -> "
          f"{out['iters'][bn].mean() - SOLVED:.2f}", flush=True)
    assert d[both].max() < PARITY_TOL
    assert out["iters"][bn].mean() < base["iters"][bn].mean()


def test_newton_c1_subgrid(step, table):
    """The c1 grid's chemistry (no vortex, declination 20 deg): against the restatement's
    converged mode, with conserv and the floor fused."""
    import cport
    from reprobus_amd import synth

    inp = synth.make_c1(seed=1, nlon=24, nlat=13)
    out = solve_gpu(step, inp, "newton", fuse=True)
    ref = cport.chem_converge(inp, *table, tol=1e-14, maxiter=50000, cr_psc=True)
    _summary("c1 24x13", out["iters"])
    both = (out["iters"] > 0) & (ref["iters"] > 0)
    d = box_rel(out["qj1"], out["hc"], ref["qj1"], ref["hc"])
    print(f"[newton] c1 vs the restatement (conserv fused): max {d[both].max():.2e}", flush=True)
    assert d[both].max() < PARITY_TOL


def test_newton_rejects_bad_arguments(step):
    from reprobus_amd import synth
    from reprobus_amd._lib import RbError

    inp = synth.make_c1(seed=3, nlon=4, nlat=3)
    for kw in ({"tol": 0.0}, {"max_newton": 0}, {"maxiter": 0}, {"sweeps_first": -1},
               {"sweeps_first": 2000}):
        with pytest.raises(RbError):
            solve_gpu(step, inp, "newton", **kw)


def test_launch_shapes_agree(table, c2):
    """The Newton kernel at 1, 2 and 4 resident waves per SIMD runs the same operations in
    the same order: the same bits and counts, from the incoming state and after sweeps
    first."""
    from reprobus_amd.chem import ChemistryStep

    inp, _ = c2
    outs = {}
    for waves in (1, 2, 4):
        s = ChemistryStep(*table, device=0, newton_waves=waves)
        try:
            for sf in (0, 30):
                outs[waves, sf] = solve_gpu(s, inp, "newton", sweeps_first=sf)
        finally:
            s.close()
    for sf in (0, 30):
        for waves in (2, 4):
            for k in ("qj1", "hc", "iters"):
                assert np.array_equal(outs[1, sf][k], outs[waves, sf][k]), (waves, sf, k)
