"""The allocation-free step (rb_chem_step_noalloc_dev) against the allocating one
(rb_chem_step_dev), through the C ABI, on two calls in a row -- the second on a smaller grid,
so that it runs on what the first left in the shared scratch -- and its size limit.  Both
run the same kernels on the same inputs: the results must be bit-identical."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _device(inp):
    import torch

    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    g = inp.grid
    shb = (g.nivbas, g.nlat, g.nlon)
    return [t(inp.T), t(inp.p), t(inp.sza), t(inp.alt), t(inp.qj1), t(inp.hc),
            torch.zeros(g.shape3, dtype=torch.float64, device=dev)] + \
        [torch.zeros(shb, dtype=torch.float64, device=dev) for _ in range(3)]


def _run(L, ctx, inp, noalloc):
    import torch
    from reprobus_amd.chem import make_grid

    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    d = _device(inp)
    fn = L.rb_chem_step_noalloc_dev if noalloc else L.rb_chem_step_dev
    print(f"[noalloc] {'noalloc' if noalloc else 'allocating'} step on "
          f"{g.nlon}x{g.nlat}x{g.nivbas}", flush=True)
    rc = fn(ctx, C.byref(grid), float(inp.dt), *[x.data_ptr() for x in d], 1)
    torch.cuda.synchronize()
    return rc, [x.cpu().numpy() for x in d[4:]]


def test_noalloc_step_matches_the_allocating_step(table):
    from reprobus_amd import _lib, synth
    from reprobus_amd.chem import ChemistryStep

    L = _lib.load()
    big = synth.make_c1(seed=4, nlon=24, nlat=13)       # 18,720 boxes
    small = synth.make_c1(seed=5, nlon=8, nlat=5)       # 2,400 boxes
    assert big.grid.nlon * big.grid.nlat * big.grid.nivbas <= _lib.RB_NOALLOC_MAX_BOXES
    ref = ChemistryStep(*table, device=0)
    na = ChemistryStep(*table, device=0)
    try:
        expect = [_run(L, ref._ctx, inp, False) for inp in (big, small)]
        got = [_run(L, na._ctx, inp, True) for inp in (big, small)]
    finally:
        ref.close()
        na.close()
    for (rc_e, out_e), (rc_g, out_g) in zip(expect, got):
        assert rc_e == 0 and rc_g == 0
        for a, b in zip(out_e, out_g):
            assert np.array_equal(a, b)


def test_noalloc_step_rejects_a_grid_beyond_its_scratch(table):
    from reprobus_amd import _lib
    from reprobus_amd.chem import ChemistryStep, make_grid

    L = _lib.load()
    step = ChemistryStep(*table, device=0)
    try:
        g = make_grid(144, 91, 60, 60, 14)              # 786,240 boxes
        dummy = C.c_void_p(8)                           # never dereferenced: nothing launches
        rc = L.rb_chem_step_noalloc_dev(step._ctx, C.byref(g), 900.0, *([dummy] * 10), 1)
        assert rc == _lib.RB_EINVAL
        assert b"allocation-free" in L.rb_last_error(step._ctx)
    finally:
        step.close()
