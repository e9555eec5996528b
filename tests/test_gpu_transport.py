"""The transport interpolation on the GPU (rb_transport_weights_dev, rb_transport_interp_dev,
rb_transport) against the restatement (pinned against the reference's build in
test_transport_oracle.py) and the golden fixture: bit for bit, FP64 throughout."""
import os

import numpy as np
import pytest

import cport
from reprobus_amd import synth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "transport_seed0.npz")


def _dev(a, dtype=None):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to("cuda:0")


def _run(step, tr):
    import torch
    from reprobus_amd.chem import make_grid

    grid = make_grid(tr.nlon, tr.nlat, tr.niv, tr.niv, 1)
    n = tr.nlon * tr.nlat * tr.niv
    idx = torch.zeros((3, n), dtype=torch.int32, device="cuda:0")
    w = torch.zeros((12, n), dtype=torch.float64, device="cuda:0")
    step.transport_weights_device(grid, *[_dev(a) for a in (tr.aaa, tr.bbb, tr.x0, tr.y0, tr.z0, tr.p0)],
                                  idx, w)
    q_in = _dev(tr.q)
    q_out = torch.zeros_like(q_in)
    step.transport_interp_device(grid, idx, w, q_in, q_out)
    torch.cuda.synchronize()
    return idx.cpu().numpy(), w.cpu().numpy(), q_out.cpu().numpy()


@pytest.fixture(scope="module")
def step(table):
    from reprobus_amd.chem import ChemistryStep

    s = ChemistryStep(*table, device=0)
    yield s
    s.close()


def test_full_grid_matches_restatement_and_golden(step):
    """2,244,060 points (180 x 91 x 137), 3 tracers."""
    tr = synth.make_transport(0)
    print("[transport] full grid, seed 0", flush=True)
    idx, w, q = _run(step, tr)
    idx_c, w_c = cport.transport_weights(tr)
    assert np.array_equal(idx, idx_c)
    assert np.array_equal(w, w_c)
    g = np.load(GOLDEN)
    pts = g["pts"]
    for t in range(len(tr.jc)):
        ref, bad = cport.internew(tr, idx_c, w_c, tr.q[t])
        assert bad == 0
        assert np.array_equal(q[t], ref)
        assert np.array_equal(q[t].reshape(-1)[pts], g["q"][t])


def test_many_tracers_small_grid(step):
    """43 tracers (the model's nbcon) on a 24 x 13 x 11 grid, with departure points in the
    polar rows and across the date line."""
    tr = synth.make_transport(3, nlon=24, nlat=13, niv=11, jc=tuple(range(1, 44)))
    rng = np.random.default_rng(3)
    xpi = 2.0 * np.arcsin(1.0)
    x0, y0 = tr.x0.reshape(-1).copy(), tr.y0.reshape(-1).copy()
    m = x0.size // 5
    x0[:m] = rng.uniform(2 * xpi - 0.5, 2 * xpi - 1e-9, m)
    y0[m:2 * m] = rng.uniform(xpi / 2 - 0.4, xpi / 2 - 1e-9, m)
    y0[2 * m:3 * m] = rng.uniform(-xpi / 2 + 1e-9, -xpi / 2 + 0.4, m)
    tr.x0, tr.y0 = x0.reshape(tr.x0.shape), y0.reshape(tr.y0.shape)
    print("[transport] 24x13x11, 43 tracers", flush=True)
    idx, w, q = _run(step, tr)
    idx_c, w_c = cport.transport_weights(tr)
    assert np.array_equal(idx, idx_c) and np.array_equal(w, w_c)
    for t in range(43):
        ref, _ = cport.internew(tr, idx_c, w_c, tr.q[t])
        assert np.array_equal(q[t], ref), t


def test_host_entry_in_place(step):
    from reprobus_amd.chem import make_grid

    tr = synth.make_transport(4, nlon=20, nlat=11, niv=9)
    print("[transport] host entry 20x11x9", flush=True)
    _, _, q_dev = _run(step, tr)
    q = tr.q.copy()
    step.transport(make_grid(20, 11, 9, 9, 1), tr.aaa, tr.bbb, tr.x0, tr.y0, tr.z0, tr.p0, q)
    assert np.array_equal(q, q_dev)


def test_departure_points_outside_the_grid_are_rejected(step):
    """A departure point on the eastern edge x0 = 2 pi (its cell would be nlon + 1) makes
    the weights entry fail; corrupted indices make the interpolation fail; no kernel
    touches memory outside its arrays meanwhile (the kernels skip such points)."""
    import torch
    from reprobus_amd._lib import RbError
    from reprobus_amd.chem import make_grid

    tr = synth.make_transport(7, nlon=16, nlat=9, niv=7)
    grid = make_grid(16, 9, 7, 7, 1)
    n = 16 * 9 * 7
    x0 = tr.x0.copy()
    x0.reshape(-1)[11] = 2.0 * (2.0 * np.arcsin(1.0))
    idx = torch.zeros((3, n), dtype=torch.int32, device="cuda:0")
    w = torch.zeros((12, n), dtype=torch.float64, device="cuda:0")
    print("[transport] point outside the grid", flush=True)
    with pytest.raises(RbError):
        step.transport_weights_device(grid, *[_dev(a) for a in (tr.aaa, tr.bbb, x0, tr.y0, tr.z0, tr.p0)],
                                      idx, w)
    idx_c, w_c = cport.transport_weights(tr)
    idx_c[1, 7] = 10                    # kg0 = nlat + 1: the stencil would leave the grid
    q_in = _dev(tr.q)
    q_out = torch.zeros_like(q_in)
    with pytest.raises(RbError):
        step.transport_interp_device(grid, _dev(idx_c), _dev(w_c), q_in, q_out)
    with pytest.raises(RbError):
        step.transport_interp_device(grid, _dev(idx_c), _dev(w_c), q_in, q_in)   # overlap


def test_tiles_without_near_points(step):
    """The interpolation works on tiles of points: a tile whose every departure cell is far
    from its arrival point gathers each stencil from the fields, and a tile with no point
    inside the grid writes nothing and only counts its points; the other tiles are
    unaffected (bit for bit against the restatement)."""
    import torch
    from reprobus_amd._lib import RbError
    from reprobus_amd.chem import make_grid

    nlon, nlat, niv = 16, 9, 7                     # one tile per level
    tr = synth.make_transport(5, nlon=nlon, nlat=nlat, niv=niv, jc=(1, 2, 3))
    grid = make_grid(nlon, nlat, niv, niv, 1)
    idx_c, w_c = cport.transport_weights(tr)
    idx_c = idx_c.reshape(3, niv, nlat, nlon).copy()
    # level 2: every departure cell 3 to 8 longitudes from its arrival point, east or west,
    # at any latitude and level: the stencils span more than a tile's window can hold, and
    # no point is near enough for the window of the near points
    rng = np.random.default_rng(5)
    i = np.arange(nlon)[None, :]
    shift = rng.integers(3, 9, (nlat, nlon)) * rng.choice([-1, 1], (nlat, nlon))
    idx_c[0, 1] = (i + shift) % nlon + 1
    idx_c[1, 1] = rng.integers(1, nlat + 1, (nlat, nlon))
    idx_c[2, 1] = rng.integers(1, niv, (nlat, nlon))
    idx_c[1, 1, 0, :2] = (1, nlat)
    idx_c[2, 1, 0, :2] = (1, niv - 1)
    # level 4: every departure cell outside the grid
    idx_c[1, 3] = nlat + 3
    idx_c = idx_c.reshape(3, -1)
    q_in = _dev(tr.q)
    q_out = torch.full_like(q_in, -7.0)
    print("[transport] 16x9x7: a level of far cells, a level outside the grid", flush=True)
    with pytest.raises(RbError):
        step.transport_interp_device(grid, _dev(idx_c), _dev(w_c), q_in, q_out)
    torch.cuda.synchronize()
    q = q_out.cpu().numpy().reshape(len(tr.jc), niv, nlat, nlon)
    assert np.all(q[:, 3] == -7.0)                 # the points outside the grid: untouched
    good = np.ones((niv, nlat, nlon), bool)
    good[3] = False
    for t in range(len(tr.jc)):
        ref, _ = cport.internew(tr, idx_c, w_c, tr.q[t])
        ref = ref.reshape(niv, nlat, nlon)
        assert np.array_equal(q[t][good], ref[good]), t
