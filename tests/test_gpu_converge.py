"""GPU converged mode (the same per-box algorithm swept to its fixed point) against the
restatement's converged mode, and the initial-guess entry.

Both sides stop a box at the first sweep whose largest relative change of any species
(values under the 1e-30 mole-fraction floor compared at the floor) is below tol, so they
take the same sweep counts; where last-bit differences move the stop by one sweep, the
extra sweep changes no species by more than tol.  Bar: 1e-9 relative per species above
the floor, as for the 5-sweep step.
"""
import numpy as np
import pytest

from conftest import PARITY_TOL, box_rel, max_rel_by_species, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10
MAXITER = 5000


AA_START = 40          # the library's default: plain sweeps before, Anderson mixing after


@pytest.fixture(scope="module")
def step(table):
    from reprobus_amd.chem import ChemistryStep

    s = ChemistryStep(*table, device=0, conv_aa_start=AA_START)
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


def converge_gpu(step, inp, tol=TOL, maxiter=MAXITER, guess=None, fuse=True, accel=0):
    import torch
    from reprobus_amd.chem import make_grid

    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    d = _dev(inp)
    gq = gh = None
    if guess is not None:
        gq = torch.from_numpy(np.ascontiguousarray(guess[0])).to("cuda:0")
        gh = torch.from_numpy(np.ascontiguousarray(guess[1])).to("cuda:0")
    step.set_stream(torch.cuda.current_stream().cuda_stream)
    step.converge_device(grid, inp.dt, tol, maxiter, d["T"], d["p"], d["sza"], d["alt"], d["qj1"],
                         d["hc"], d["surfarea"], d["vsed3d"], d["parth2osed"], d["parthno3sed"],
                         guess_qj1=gq, guess_hc=gh, iters=d["iters"], fuse_conserv=fuse,
                         accel=accel)
    torch.cuda.synchronize()
    step.set_stream(None)
    return {k: v.cpu().numpy() for k, v in d.items()}


def _worst(out, ref):
    return max(float(max_rel_by_species(out[k], ref[k]).max()) for k in ("qj1", "hc"))


@pytest.mark.parametrize("cr_psc", [True, False])
def test_converged_matches_restatement(step, table, cr_psc):
    import cport
    from reprobus_amd import synth

    inp = synth.make_c1(seed=1, nlon=24, nlat=13)
    out = converge_gpu(step, inp)
    ref = cport.chem_converge(inp, *table, tol=TOL, maxiter=MAXITER, cr_psc=cr_psc)
    it, rit = out["iters"], ref["iters"]
    same = float((it == rit).mean())
    w = _worst(out, ref)
    print(f"[converge] cr_psc={cr_psc}: sweeps median {np.median(it):.0f} p99 "
          f"{np.percentile(it, 99):.0f} max {it.max()}, same count {same:.4%}, "
          f"max |dcount| {int(np.abs(it - rit).max())}, worst rel {w:.2e}", flush=True)
    assert np.all((it >= 1) & (it <= MAXITER) | (it < 0) & (it >= -MAXITER))
    assert same > 0.999
    if cr_psc:
        assert w < PARITY_TOL
    else:
        # the C-library variant: only the ill-conditioned ternary PSC boxes may differ
        d = np.maximum(rel_err(out["qj1"], ref["qj1"]).max(axis=0),
                       rel_err(out["hc"], ref["hc"]).max(axis=0))
        over = d > PARITY_TOL
        print(f"[converge] C-library variant: {int(over.sum())} of {d.size} boxes above "
              f"{PARITY_TOL:g}, max {d.max():.2e}", flush=True)
        assert over.sum() <= 5 and d.max() < 1e-8 and np.all(inp.T[over] < 215.0)


@pytest.mark.parametrize("seed", [1, 0])
def test_accelerated_matches_restatement(step, table, seed):
    """Anderson depth 2 on the GPU against the restatement's depth-2 iteration (the same
    arithmetic): the same sweep counts on almost every box and agreement at the 1e-9 bar;
    and both reach the plain iteration's fixed point."""
    import cport
    from reprobus_amd import synth

    inp = synth.make_c2(seed, nlon=24, nlat=13)
    out = converge_gpu(step, inp, maxiter=2000, accel=2)
    ref = cport.chem_converge(inp, *table, tol=TOL, maxiter=2000, cr_psc=True, accel=2,
                              aa_start=AA_START)
    it, rit = out["iters"], ref["iters"]
    same = float((it == rit).mean())
    both = (it > 0) & (rit > 0)
    dr = box_rel(out["qj1"], out["hc"], ref["qj1"], ref["hc"])
    plain = cport.chem_converge(inp, *table, tol=1e-13, maxiter=20000, cr_psc=True)
    live = (plain["iters"] > 0) & (it > 0)
    d = np.maximum(rel_err(out["qj1"], plain["qj1"]).max(axis=0),
                   rel_err(out["hc"], plain["hc"]).max(axis=0))
    print(f"[converge] AA(2) from sweep {AA_START}, seed {seed}: sweeps mean {np.abs(it).mean():.1f} max {np.abs(it).max()} "
          f"(plain {np.abs(plain['iters']).mean():.1f}), same count {same:.4%}, vs restatement "
          f"max {dr[both].max():.2e} ({int((dr[both] > PARITY_TOL).sum())} boxes above "
          f"{PARITY_TOL:g}), vs tight plain fixed point p99 {np.percentile(d[live], 99):.2e}",
          flush=True)
    assert same > 0.99
    # a box at a kink of the reference's min() terms (the odd-oxygen loss, F:1900-1917) has
    # no unique limit under mixing: the two sides may stop a few 1e-9 apart there
    assert (dr[both] > PARITY_TOL).sum() <= 2 and dr[both].max() < 2e-8
    assert np.percentile(d[live], 99) < PARITY_TOL


def test_converged_c2_with_fallback(step, table):
    """The c2 vortex state, where some boxes' iteration diverges past the reference's 5
    sweeps: the GPU flags the same boxes (negative counts), gives them the 5-sweep answer,
    and matches the restatement everywhere."""
    import cport
    from reprobus_amd import synth

    inp = synth.make_c2(0, nlon=24, nlat=13)
    out = converge_gpu(step, inp, maxiter=2000)
    ref = cport.chem_converge(inp, *table, tol=TOL, maxiter=2000, cr_psc=True)
    it, rit = out["iters"], ref["iters"]
    nfail = int((rit < 0).sum())
    w = _worst(out, ref)
    print(f"[converge] c2: {nfail} fallback boxes (GPU {int((it < 0).sum())}), same count "
          f"{float((it == rit).mean()):.4%}, worst rel {w:.2e}", flush=True)
    assert nfail > 0
    assert np.array_equal(it < 0, rit < 0)
    assert float((it == rit).mean()) > 0.999
    assert np.isfinite(out["qj1"]).all() and np.isfinite(out["hc"]).all()
    assert w < PARITY_TOL


def test_converged_from_guess(step, table):
    """A perturbed fixed point as the initial iterate (the learned-guess entry): the GPU
    matches the restatement started from the same guess, converges back to the fixed point
    reached from the incoming state, and needs fewer sweeps.  Raw chemistry output (no
    conserv), since conserv moves the state off the fixed point."""
    import cport
    from reprobus_amd import synth

    inp = synth.make_c1(seed=3, nlon=16, nlat=9)
    tol = 1e-12
    base = converge_gpu(step, inp, tol=tol, fuse=False)
    rng = np.random.default_rng(0)
    gq = base["qj1"] * np.exp(rng.normal(0, 1e-3, base["qj1"].shape))
    gh = base["hc"] * np.exp(rng.normal(0, 1e-3, base["hc"].shape))
    out = converge_gpu(step, inp, tol=tol, guess=(gq, gh), fuse=False)
    ref = cport.chem_converge(inp, *table, tol=tol, maxiter=MAXITER, conserv=False, cr_psc=True,
                              guess=(gq, gh))
    same = float((out["iters"] == ref["iters"]).mean())
    # boxes converged on every path (a box at a kink of the min() terms can settle into a
    # two-sweep cycle above tol from one start and converge from another)
    live = (base["iters"] > 0) & (out["iters"] > 0) & (ref["iters"] > 0)
    assert live.mean() > 0.99
    w_ref = float(box_rel(out["qj1"], out["hc"], ref["qj1_chem"], ref["hc_chem"])[live].max())
    d = np.maximum(rel_err(out["qj1"], base["qj1"]).max(axis=0),
                   rel_err(out["hc"], base["hc"]).max(axis=0))
    print(f"[converge] guess: vs restatement {w_ref:.2e} (same count {same:.4%}); vs unguessed "
          f"fixed point {float(d[live].max()):.2e}; mean sweeps {base['iters'][live].mean():.1f} -> "
          f"{out['iters'][live].mean():.1f}", flush=True)
    assert w_ref < PARITY_TOL and same > 0.99
    assert d[live].max() < PARITY_TOL
    assert out["iters"][live].mean() < base["iters"][live].mean()


def test_converged_rejects_bad_arguments(step):
    from reprobus_amd._lib import RbError
    from reprobus_amd import synth

    inp = synth.make_c1(seed=3, nlon=4, nlat=3)
    with pytest.raises(RbError):
        converge_gpu(step, inp, tol=0.0)
    with pytest.raises(RbError):
        converge_gpu(step, inp, maxiter=0)
    with pytest.raises(RbError):
        converge_gpu(step, inp, accel=1)
    # Anderson mixing from a guess is refused (it can leave boxes short of the fixed point
    # under the 1e-30 floor); from the incoming state it runs
    with pytest.raises(RbError, match="accel with a guess"):
        converge_gpu(step, inp, guess=(inp.qj1, inp.hc), accel=2)
    converge_gpu(step, inp, maxiter=50, accel=2)


def test_work_order_changes_nothing(table):
    """The converged kernels take boxes night first, then day (rb_config.conv_order 0), so
    that a wave's lanes run one branch of the sweep.  A box's answer and sweep count do not
    depend on the lane or the time it is swept in: grid order (conv_order 1) gives the same
    bits, for plain and accelerated sweeps and for the Newton solve."""
    import torch
    from reprobus_amd import synth
    from reprobus_amd.chem import ChemistryStep, make_grid

    inp = synth.make_c2(1, nlon=24, nlat=13)
    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    night = float((inp.sza[:g.nivbas] > 95.).mean())
    print(f"[order] 24x13x{g.nivbas} c2 subgrid, night fraction {night:.2f}", flush=True)
    assert 0.1 < night < 0.9
    outs = {}
    for order in (0, 1):
        s = ChemistryStep(*table, device=0, conv_aa_start=AA_START, conv_order=order)
        try:
            for how in ("plain", "accel", "newton"):
                d = _dev(inp)
                args = [d[k] for k in ("T", "p", "sza", "alt", "qj1", "hc", "surfarea", "vsed3d",
                                       "parth2osed", "parthno3sed")]
                s.set_stream(torch.cuda.current_stream().cuda_stream)
                if how == "newton":
                    s.newton_device(grid, inp.dt, 1e-11, 20, 2000, *args, iters=d["iters"],
                                    sweeps_first=30)
                else:
                    s.converge_device(grid, inp.dt, TOL, 2000, *args, iters=d["iters"],
                                      accel=2 if how == "accel" else 0)
                torch.cuda.synchronize()
                s.set_stream(None)
                outs[order, how] = {k: d[k].cpu().numpy() for k in ("qj1", "hc", "iters")}
        finally:
            s.close()
    for how in ("plain", "accel", "newton"):
        a, b = outs[0, how], outs[1, how]
        for k in ("qj1", "hc", "iters"):
            assert np.array_equal(a[k], b[k]), (how, k)


def test_gpu_step_matches_reference_with_niter_raised(table):
    """The fixed-sweep step at rb_config.niter = 60 against the reference built with niter
    (F:630) = 60, on the 16 c0 columns (960 boxes), at the 1e-9 bar."""
    from conftest import load_fixture
    from reprobus_amd.chem import ChemistryStep

    inp, ref = load_fixture("c0_seed0_16col_n60")
    s = ChemistryStep(*table, device=0, niter=60)
    try:
        out = s.run_inputs(inp)
    finally:
        s.close()
    w = max(float(max_rel_by_species(out[k], ref[k]).max())
            for k in ("qj1_chem", "hc_chem", "qj1", "hc"))
    print(f"[niter] GPU 60 sweeps vs reference niter=60: max {w:.2e}", flush=True)
    assert w < PARITY_TOL


@pytest.mark.parametrize("how", ["plain", "newton"])
def test_gpu_converged_matches_reference_swept_to_its_fixed_point(table, how):
    """The converged solve (plain sweeps; Newton after 30 sweeps) against the reference
    built with niter = 400 on 16 c2 columns (8 in the polar vortex), on the boxes where the
    400 sweeps have reached the fixed point -- those the restatement's converged mode
    (tol 1e-14) converges within 300 sweeps -- and the GPU solve converged: the same fixed
    point at the 1e-9 bar.  (A box the sweeps converge more slowly than that is not at its
    fixed point after 400 of them.)"""
    import cport
    from conftest import load_fixture
    from reprobus_amd.chem import ChemistryStep
    from test_gpu_newton import solve_gpu

    inp, ref = load_fixture("c2_seed1_16col_n400")
    rit = cport.chem_converge(inp, *table, tol=1e-14, maxiter=50000, conserv=True)["iters"]
    s = ChemistryStep(*table, device=0)
    try:
        if how == "plain":
            out = converge_gpu(s, inp, tol=1e-12, maxiter=2000)
            it = out["iters"]
            ok = (it > 0) & (rit > 0) & (rit < 300)
        else:
            out = solve_gpu(s, inp, "newton", tol=1e-12, sweeps_first=30, fuse=True)
            it = out["iters"]
            ok = (it > 0) & (it < 20000) & (rit > 0) & (rit < 300)
    finally:
        s.close()
    d = np.maximum(rel_err(out["qj1"][:42], ref["qj1"][:42]).max(axis=0),
                   rel_err(out["hc"], ref["hc"]).max(axis=0))
    print(f"[niter] {how}: {int(ok.sum())} of {ok.size} boxes converged, vs reference "
          f"niter=400 max {d[ok].max():.2e}", flush=True)
    assert ok.sum() >= 940
    assert d[ok].max() < PARITY_TOL


def test_block_sizes_give_the_same_bits(table):
    """A box's converged answer and sweep count do not depend on the launch shape: plain,
    accelerated and Newton solves at blocks 64, 128 and 256 (the shapes the autotuner
    chooses from) give the same bits on a c2 subgrid."""
    import torch
    from reprobus_amd import synth
    from reprobus_amd.chem import ChemistryStep, make_grid

    inp = synth.make_c2(1, nlon=24, nlat=13)
    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    outs = {}
    for block in (64, 128, 256):
        print(f"[blocks] block {block}", flush=True)
        s = ChemistryStep(*table, device=0, block=block, conv_refill=32, newton_waves=1,
                          conv_aa_start=AA_START)
        try:
            for how in ("plain", "accel", "newton"):
                d = _dev(inp)
                args = [d[k] for k in ("T", "p", "sza", "alt", "qj1", "hc", "surfarea", "vsed3d",
                                       "parth2osed", "parthno3sed")]
                s.set_stream(torch.cuda.current_stream().cuda_stream)
                if how == "newton":
                    s.newton_device(grid, inp.dt, 1e-11, 20, 2000, *args, iters=d["iters"],
                                    sweeps_first=30)
                else:
                    s.converge_device(grid, inp.dt, TOL, 2000, *args, iters=d["iters"],
                                      accel=2 if how == "accel" else 0)
                torch.cuda.synchronize()
                s.set_stream(None)
                outs[block, how] = {k: d[k].cpu().numpy() for k in ("qj1", "hc", "iters")}
        finally:
            s.close()
    for how in ("plain", "accel", "newton"):
        for block in (128, 256):
            for k in ("qj1", "hc", "iters"):
                assert np.array_equal(outs[64, how][k], outs[block, how][k]), (how, block, k)
