"""The autotuner on the device: every launch shape gives the same numbers, and tuning
leaves the caller's state unchanged."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(inp):
    import torch

    g = inp.grid
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    st = {"T": t(inp.T), "p": t(inp.p), "sza": t(inp.sza), "alt": t(inp.alt), "qj1": t(inp.qj1),
          "hc": t(inp.hc), "surfarea": torch.zeros(g.shape3, dtype=torch.float64, device=dev),
          "dt": inp.dt}
    shb = (g.nivbas, g.nlat, g.nlon)
    for k in ("vsed3d", "parth2osed", "parthno3sed"):
        st[k] = torch.zeros(shb, dtype=torch.float64, device=dev)
    return st


@pytest.mark.parametrize("mode", ["plain", "converged"])
def test_tune_and_shape_independence(table, mode, tmp_path):
    import torch
    from reprobus_amd import autotune, synth
    from reprobus_amd.chem import ChemistryStep, make_grid

    inp = synth.make_c2(2, nlon=24, nlat=13)
    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    st = _state(inp)
    q_before = st["qj1"].cpu().numpy().copy()
    r = autotune.tune(table, grid, st, mode=mode, reps=2, cache_path=str(tmp_path / "t.json"))
    print(f"[autotune] {mode}: block {r.block}, refill {r.refill}, ms {r.ms}", flush=True)
    assert r.block in autotune.BLOCKS
    assert (r.refill is None) == (mode == "plain")
    assert np.array_equal(st["qj1"].cpu().numpy(), q_before)
    outs = []
    shapes = [(b, 16) for b in autotune.BLOCKS]
    if mode == "converged":
        shapes += [(128, f) for f in autotune.REFILLS if f != 16]
    for b, f in shapes:
        s = ChemistryStep(*table, device=0, block=b, conv_refill=f)
        s.set_stream(torch.cuda.current_stream().cuda_stream)
        q, h = st["qj1"].clone(), st["hc"].clone()
        args = [st[k] for k in ("T", "p", "sza", "alt")]
        aux = [st[k] for k in ("surfarea", "vsed3d", "parth2osed", "parthno3sed")]
        if mode == "plain":
            s.step_device(grid, inp.dt, *args, q, h, *aux, fuse_conserv=True)
        else:
            s.converge_device(grid, inp.dt, 1e-10, 500, *args, q, h, *aux, accel=2)
        torch.cuda.synchronize()
        outs.append((q.cpu().numpy(), h.cpu().numpy()))
        s.close()
    for q, h in outs[1:]:
        assert np.array_equal(q, outs[0][0]) and np.array_equal(h, outs[0][1])
