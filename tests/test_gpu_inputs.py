"""Device input producers (zenith F:5308-5406, altitude src/altitude.f90) against the
reference's own output (golden fixture from the flang build) and the C restatement."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, PARITY_TOL, max_rel_by_species

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def step(table):
    from reprobus_amd.chem import ChemistryStep

    s = ChemistryStep(*table, device=0)
    yield s
    s.close()


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(GOLDEN, "inputs_seed11.npz"), allow_pickle=False)


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


def test_zenith_matches_reference(step, fx):
    import torch
    from reprobus_amd.chem import make_grid

    nz = fx["zlon"].shape[0]
    grid = make_grid(180, 1, nz, nz, 2)
    worst = 0.0
    for d, day in enumerate(fx["days"]):
        sza = np.empty((nz, 180))
        step.zenith(grid, day, np.ascontiguousarray(fx["zlon"]), np.ascontiguousarray(fx["zlat"]), sza)
        worst = max(worst, float(np.abs(sza - fx["sza"][d]).max()))
        ds = torch.empty((nz, 180), dtype=torch.float64, device="cuda:0")
        step.set_stream(torch.cuda.current_stream().cuda_stream)
        step.zenith_device(grid, day, _dev(fx["zlon"]), _dev(fx["zlat"]), ds)
        torch.cuda.synchronize()
        step.set_stream(None)
        assert np.array_equal(ds.cpu().numpy(), sza)
    print(f"[inputs] zenith: max |dSZA| {worst:.2e} deg", flush=True)
    assert worst < 1e-10


def _alt_args(fx):
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    return [c(fx[k]) for k in ("relief", "psol", "aaa", "bbb", "t", "h2o", "tlon", "tlat", "tpre")]


def test_altitude_matches_reference(step, fx):
    import torch
    from reprobus_amd.chem import make_grid

    nl, m, _ = fx["t"].shape
    grid = make_grid(m, m, nl, nl, 2)
    alt = np.empty((nl, m, m))
    args = _alt_args(fx)
    step.altitude(grid, *args, alt)
    rel = float(np.max(np.abs(alt - fx["alt"]) / np.abs(fx["alt"])))
    dalt = torch.empty((nl, m, m), dtype=torch.float64, device="cuda:0")
    step.altitude_device(grid, *[_dev(a) for a in args], dalt)
    torch.cuda.synchronize()
    print(f"[inputs] altitude: max rel {rel:.2e}", flush=True)
    assert rel < 1e-13
    assert np.array_equal(dalt.cpu().numpy(), alt)


def test_altitude_model_grid_against_restatement(step):
    """The model's own regular grid, 180 x 91 x 137 (2.24 M points)."""
    import cport
    from reprobus_amd.chem import make_grid

    rng = np.random.default_rng(3)
    m1, m2, nl = 180, 91, 137
    eta = (np.arange(nl) + 0.5) / nl
    bbb, aaa = eta ** 3, 0.01 + 200.0 * eta * (1 - eta) ** 2
    psol = rng.uniform(500.0, 1040.0, (m2, m1))
    relief = rng.uniform(0.0, 5.0, (m2, m1))
    t = 200.0 + 90.0 * eta[:, None, None] + rng.normal(0, 3, (nl, m2, m1))
    h2o = np.exp(rng.uniform(np.log(2e-6), np.log(2e-2), (nl, m2, m1)))
    lon = np.arange(m1) * 2.0
    lat = 90.0 - np.arange(m2) * 2.0
    tlon = (lon[None, None, :] + rng.uniform(-3, 3, (nl, m2, m1))) % 360.0
    tlat = np.clip(lat[None, :, None] + rng.uniform(-3, 3, (nl, m2, m1)), -90, 90)
    p = aaa[:, None, None] + bbb[:, None, None] * psol[None]
    tpre = p * rng.uniform(0.9, 1.1, p.shape)
    grid = make_grid(m1, m2, nl, nl, 2)
    args = [np.ascontiguousarray(a) for a in (relief, psol, aaa, bbb, t, h2o, tlon, tlat, tpre)]
    alt = np.empty((nl, m2, m1))
    step.altitude(grid, *args, alt)
    ref = cport.altitude(*args)
    rel = float(np.max(np.abs(alt - ref) / np.abs(ref)))
    print(f"[inputs] altitude 180x91x137: max rel vs restatement {rel:.2e}", flush=True)
    assert rel < 1e-13


def test_chemistry_fed_by_device_inputs(step, table):
    """Driver path with only lon, lat, p, T and the day number: device zenith and altitude
    feed the device step; the restatement fed by its own zenith/altitude agrees at 1e-9."""
    import torch
    import cport
    from reprobus_amd import synth
    from reprobus_amd.chem import make_grid

    inp = synth.make_c1(seed=7, nlon=24, nlat=13)
    g = inp.grid
    grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
    lat, lon = inp.meta["lat"], inp.meta["lon"]
    tlon = np.ascontiguousarray(np.broadcast_to(lon[None, None, :], g.shape3))
    tlat = np.ascontiguousarray(np.broadcast_to(lat[None, :, None], g.shape3))
    day = 172.25
    nl = g.niv
    eta = np.linspace(0.0, 1.0, nl)
    bbb = np.zeros(nl)
    aaa = inp.p[:, 0, 0].copy()                   # pure pressure levels: p = aaa
    psol = np.full((g.nlat, g.nlon), 1013.25)
    relief = np.zeros((g.nlat, g.nlon))
    args = [np.ascontiguousarray(a) for a in (relief, psol, aaa, bbb, inp.T, inp.qj1[synth.L["h2o"]],
                                              tlon, tlat, inp.p)]
    dlon, dlat = 360.0 / g.nlon, 180.0 / (g.nlat - 1)
    # device producers -> device step
    d = {k: _dev(v) for k, v in (("T", inp.T), ("p", inp.p), ("q", inp.qj1), ("h", inp.hc))}
    sza = torch.empty(g.shape3, dtype=torch.float64, device="cuda:0")
    alt = torch.empty(g.shape3, dtype=torch.float64, device="cuda:0")
    step.set_stream(torch.cuda.current_stream().cuda_stream)
    step.zenith_device(grid, day, _dev(tlon), _dev(tlat), sza)
    step.altitude_device(grid, *[_dev(a) for a in args], alt, dlon=dlon, dlat=dlat)
    surf = torch.zeros(g.shape3, dtype=torch.float64, device="cuda:0")
    aux = [torch.zeros((g.nivbas, g.nlat, g.nlon), dtype=torch.float64, device="cuda:0")
           for _ in range(3)]
    step.step_device(grid, inp.dt, d["T"], d["p"], sza, alt, d["q"], d["h"], surf, *aux,
                     fuse_conserv=True)
    torch.cuda.synchronize()
    step.set_stream(None)
    # restatement with its own producers
    ref_in = inp.copy()
    ref_in.sza = cport.zenith(day, tlon, tlat).reshape(g.shape3)
    ref_in.alt = cport.altitude(*args, dlon=dlon, dlat=dlat)
    assert np.max(np.abs(sza.cpu().numpy() - ref_in.sza)) < 1e-10
    assert np.max(np.abs(alt.cpu().numpy() - ref_in.alt) / np.abs(ref_in.alt)) < 1e-13
    ref = cport.chem_step(ref_in, *table, cr_psc=True)
    worst = max(float(max_rel_by_species(d["q"].cpu().numpy(), ref["qj1"]).max()),
                float(max_rel_by_species(d["h"].cpu().numpy(), ref["hc"]).max()))
    print(f"[inputs] chemistry on device-produced SZA/alt vs restatement: {worst:.2e}", flush=True)
    assert worst < PARITY_TOL
