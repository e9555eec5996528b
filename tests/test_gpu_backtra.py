"""Back-trajectories on the GPU (rb_transport_backtra_dev) against the restatement, which is
pinned bit for bit to the reference's own backtra (test_transport_oracle.py), on the
model's 180 x 91 x 137 grid; the closed forms of still air and solid-body rotation; the
departure points feeding the interpolation, as semilag chains them (F:6142-6150).

The device evaluates sin, cos, asin, acos and exp correctly rounded.  The reference's
glibc misrounds ~0.1 % of such calls, so a few thousand of the 2.2 M points differ from the
restatement by a last-bit change propagated through the sub-steps.  Near the rotated frame's
acos(~1) that grows to ~1e-8 rad of longitude.  The bar: most values bit-identical, the rest
within 1e-9 relative (positions, pressures, temperatures, interpolated tracers) and the same
cells except where a point lies within rounding of a cell edge."""
import numpy as np
import pytest

import cport
from reprobus_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def step(table):
    from reprobus_amd.chem import ChemistryStep

    s = ChemistryStep(*table, device=0)
    yield s
    s.close()


def backtra_gpu(step, ti, nrapp, traj=True):
    import torch
    from reprobus_amd.chem import make_grid

    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    grid = make_grid(ti.nlon, ti.nlat, ti.niv, ti.niv, 1)
    n = ti.nlon * ti.nlat * ti.niv
    rec = lambda r: {"u": t(r.u), "v": t(r.v), "w": t(r.w), "t": t(r.t), "lnps": t(r.lnps),
                     "day": r.day}
    out = {k: torch.zeros(n, dtype=torch.float64, device=dev) for k in ("x0", "y0", "z0", "p0")}
    idx = torch.zeros((3, n), dtype=torch.int32, device=dev)
    w = torch.zeros((12, n), dtype=torch.float64, device=dev)
    tr = None
    if traj:
        tr = [torch.zeros((nrapp, ti.niv, ti.nlat, ti.nlon), dtype=torch.float64, device=dev)
              for _ in range(4)]
    print(f"[backtra] {ti.nlon}x{ti.nlat}x{ti.niv}, nrapp {nrapp}", flush=True)
    step.set_stream(torch.cuda.current_stream().cuda_stream)
    step.backtra_device(grid, t(ti.aaa), t(ti.bbb), rec(ti.m1), rec(ti.p1), ti.daynum, ti.dtdyn,
                        nrapp, ti.ntime, ti.nintecmwf, out["x0"], out["y0"], out["z0"],
                        out["p0"], idx, w, traj=tr)
    torch.cuda.synchronize()
    step.set_stream(None)
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["idx"], res["w"] = idx.cpu().numpy(), w.cpu().numpy()
    if traj:
        for k, a in zip(("trajlon", "trajlat", "trajpre", "trajtem"), tr):
            res[k] = a.cpu().numpy()
    return res


@pytest.fixture(scope="module")
def full(step):
    ti = synth.make_trajectory_inputs(0)
    return ti, backtra_gpu(step, ti, 4), cport.backtra(ti, 4)


def _angle(a, b):
    return np.abs(np.angle(np.exp(1j * (a - b))))


def test_backtra_matches_restatement(full):
    ti, g, r = full
    same = {k: float((g[k] == r[k]).mean()) for k in ("x0", "y0", "z0", "p0", "trajtem")}
    cells = float((g["idx"] == r["idx"]).all(axis=0).mean())
    dx, dy = _angle(g["x0"], r["x0"]), np.abs(g["y0"] - r["y0"])
    rel = lambda a, b: np.abs(a - b) / np.abs(b)
    print(f"[backtra] bit-identical fractions {same}, same cells {cells:.6f}; max |dx| "
          f"{dx.max():.2e} rad, |dy| {dy.max():.2e} rad, z0 {rel(g['z0'], r['z0']).max():.2e}, "
          f"p0 {rel(g['p0'], r['p0']).max():.2e}, T {rel(g['trajtem'], r['trajtem']).max():.2e}",
          flush=True)
    for k in ("x0", "y0", "z0", "p0"):
        assert same[k] > 0.98, k
    assert cells > 0.9999
    assert dx.max() < 1e-7 and dy.max() < 1e-9
    for k in ("z0", "p0", "trajpre", "trajtem"):
        assert rel(g[k], r[k]).max() < 1e-9, k
    both = (g["idx"] == r["idx"]).all(axis=0)
    assert np.abs(g["w"] - r["w"])[:, both].max() < 1e-6
    assert np.abs(g["trajlat"] - r["trajlat"]).max() < 1e-7
    assert _angle(np.radians(g["trajlon"]), np.radians(r["trajlon"])).max() < 1e-7


def test_departure_points_feed_the_interpolation(step, full):
    """semilag (F:6142-6150): the GPU's departure cells and weights into
    rb_transport_interp_dev against the restatement's into rbo_internew, for two tracers."""
    import torch
    from reprobus_amd.chem import make_grid

    ti, g, r = full
    tr = synth.make_transport(1)
    grid = make_grid(ti.nlon, ti.nlat, ti.niv, ti.niv, 1)
    dev = torch.device("cuda:0")
    q_in = torch.from_numpy(tr.q[:2].copy()).to(dev)
    q_out = torch.zeros_like(q_in)
    step.transport_interp_device(grid, torch.from_numpy(g["idx"]).to(dev),
                                 torch.from_numpy(g["w"]).to(dev), q_in, q_out)
    torch.cuda.synchronize()
    out = q_out.cpu().numpy()
    for t in range(2):
        ref, bad = cport.internew(tr, r["idx"], r["w"], tr.q[t])
        assert bad == 0
        d = np.abs(out[t].ravel() - ref.ravel()) / np.abs(ref.ravel())
        print(f"[backtra] tracer {t} after semilag: max rel {d.max():.2e}, bit-identical "
              f"{float((out[t].ravel() == ref.ravel()).mean()):.5f}", flush=True)
        assert d.max() < 1e-9


def test_backtra_closed_forms(step):
    from test_transport_oracle import backtra_closed_forms

    backtra_closed_forms(lambda ti, nr: backtra_gpu(step, ti, nr))


def test_backtra_rejects_bad_arguments(step):
    from reprobus_amd._lib import RbError

    ti = synth.make_trajectory_inputs(2, nlon=12, nlat=7, niv=6)
    with pytest.raises(RbError):
        backtra_gpu(step, ti, 0)
    small = synth.make_trajectory_inputs(2, nlon=12, nlat=4, niv=6)
    with pytest.raises(RbError, match="nlat >= 5"):
        backtra_gpu(step, small, 2)
    bad = synth.make_trajectory_inputs(2, nlon=12, nlat=7, niv=6)
    bad.dtdyn = -1.0
    with pytest.raises(RbError):
        backtra_gpu(step, bad, 2)
    out = backtra_gpu(step, ti, 2, traj=False)          # the trajectory arrays are optional
    assert np.isfinite(out["x0"]).all()


def test_host_entry_equals_device_entry(step):
    """rb_transport_backtra (host arrays, the Fortran model's entry) gives the device entry's
    departure points and trajectory bit for bit."""
    import ctypes as C

    from reprobus_amd import _lib
    from reprobus_amd.chem import make_grid

    ti = synth.make_trajectory_inputs(3, nlon=36, nlat=19, niv=12)
    nrapp = 3
    dev = backtra_gpu(step, ti, nrapp)
    n = ti.nlon * ti.nlat * ti.niv
    keep = []

    def rec(r):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (r.u, r.v, r.w, r.t, r.lnps)]
        keep.extend(arrs)
        return _lib.Forcing(*[a.ctypes.data for a in arrs], float(r.day))

    m1, p1 = rec(ti.m1), rec(ti.p1)
    out = {k: np.zeros(n) for k in ("x0", "y0", "z0", "p0")}
    tr = {k: np.zeros((nrapp, ti.niv, ti.nlat, ti.nlon)) for k in ("trajlon", "trajlat", "trajpre", "trajtem")}
    grid = make_grid(ti.nlon, ti.nlat, ti.niv, ti.niv, 1)
    rc = step._L.rb_transport_backtra(step._ctx, C.byref(grid), ti.aaa.ctypes.data,
                                      ti.bbb.ctypes.data, C.byref(m1), C.byref(p1), ti.daynum,
                                      ti.dtdyn, nrapp, ti.ntime, ti.nintecmwf,
                                      *[out[k].ctypes.data for k in ("x0", "y0", "z0", "p0")],
                                      *[tr[k].ctypes.data for k in ("trajlon", "trajlat", "trajpre", "trajtem")])
    assert rc == 0
    for k in ("x0", "y0", "z0", "p0"):
        assert np.array_equal(out[k], dev[k]), k
    for k in tr:
        assert np.array_equal(tr[k], dev[k]), k
