"""Write the golden fixtures: reference REPROBUS chemistry (flang build of
src/reprobus_l137_20230723.f + src/jno.f90, driven by oracle/ref_harness.f90)
on small synthetic grids.

Run from the repository root where /root/reference is present:
    make -C oracle ref && python tests/golden/make_golden.py
(``python tests/golden/make_golden.py wet``, ``transport`` or ``niter`` writes that fixture
only; ``niter`` needs the raised-sweep builds, ``make -C oracle ref``.)

Each chemistry fixture stores the step inputs (with the SZA the reference realised,
``sza3d``) and the reference outputs after chem3d and after conserv + floor; the
sedimentation fixture also holds qj1 after the reference's sedimentation.  The input
producers, ECMWF reader and transport interpolation have fixtures of their own.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from reprobus_amd import synth  # noqa: E402
import refrun  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def subset_columns(inp, cols):
    """Keep the listed columns (nlat == 1 grids)."""
    out = inp.copy()
    out.grid = synth.Grid(len(cols), 1, inp.grid.niv, inp.grid.nivbas, inp.grid.nivhau)
    for name in ("T", "p", "sza", "alt", "sad_target"):
        a = getattr(inp, name)
        if a is not None:
            setattr(out, name, np.ascontiguousarray(a[..., cols]))
    out.qj1 = np.ascontiguousarray(inp.qj1[..., cols])
    out.hc = np.ascontiguousarray(inp.hc[..., cols])
    return out


def psc_grid(seed: int = 7):
    """Cold polar boxes: T 183-196 K, high water and HNO3, liquid/NAT/ice PSC."""
    inp = synth.make_c0(seed)
    rng = np.random.default_rng(seed + 100)
    shape = inp.grid.shape3
    inp.T = rng.uniform(183.0, 196.0, shape)
    inp.qj1[synth.L["h2o"]] = rng.uniform(3e-6, 8e-6, shape)
    scale = rng.uniform(1.0, 10.0, shape)
    for n in ("noy", "hno3", "n2o5", "nox", "ho2no2", "clono2", "clno2", "brono2", "hno3_gas"):
        inp.qj1[synth.L[n]] *= scale
    inp.qj1[synth.L["h2so4"]] = synth.h2so4_for_sad(inp.T, inp.p, inp.qj1[synth.L["h2o"]],
                                                     inp.sad_target * 5.0)
    inp.meta["name"] = "psc"
    return inp


def sedi_grid(seed: int = 9, ncol: int = 8, nlev: int = 94):
    """The reference's own level count (nivbas = 94, heterogeneous chemistry from level
    14, H2O sedimented above level nivh2oecmwf = 59): cold, wet, HNO3-rich columns so that
    NAT/ice particles settle."""
    rng = np.random.default_rng(seed)
    g = synth.Grid(nlon=ncol, nlat=1, niv=nlev, nivbas=nlev)
    shape = g.shape3
    p = np.broadcast_to(synth.level_pressures(nlev)[:, None, None], shape).copy()
    T = rng.uniform(183.0, 200.0, shape)
    sza = np.degrees(np.arccos(rng.uniform(-0.3, 1.0, shape)))
    q, h = synth.draw_species(rng, shape)
    q[synth.L["h2o"]] = rng.uniform(3e-6, 8e-6, shape)
    scale = rng.uniform(1.0, 10.0, shape)
    for n in ("noy", "hno3", "n2o5", "nox", "ho2no2", "clono2", "clno2", "brono2", "hno3_gas"):
        q[synth.L[n]] *= scale
    sad = np.exp(rng.uniform(np.log(0.1), np.log(10.0), shape)) * 1e-8
    q[synth.L["h2so4"]] = synth.h2so4_for_sad(T, p, q[synth.L["h2o"]], sad)
    return synth.StepInputs(g, 900.0, T, p, sza, synth.altitude_km(p), q, h, sad,
                            {"name": "sedi", "seed": seed})


def wet_grid(seed: int = 15, ncol: int = 8, nlev: int = 94):
    """Rain-out (F:1315-1338): the reference's 94 levels with tropospheric water on the
    lower ones -- H2O log-uniform 30 ppmv to 3000 ppmv on levels 55..94, so most of those
    boxes sit above the 100 ppmv threshold and some below it -- at 220-260 K there (liquid
    aerosol only: no NAT, ice or ternary solution)."""
    rng = np.random.default_rng(seed)
    g = synth.Grid(nlon=ncol, nlat=1, niv=nlev, nivbas=nlev)
    shape = g.shape3
    p = np.broadcast_to(synth.level_pressures(nlev)[:, None, None], shape).copy()
    T = rng.uniform(200.0, 240.0, shape)
    T[54:] = rng.uniform(220.0, 260.0, (nlev - 54,) + shape[1:])
    sza = np.degrees(np.arccos(rng.uniform(-0.3, 1.0, shape)))
    q, h = synth.draw_species(rng, shape)
    q[synth.L["h2o"], 54:] = np.exp(rng.uniform(np.log(30e-6), np.log(3e-3),
                                                (nlev - 54,) + shape[1:]))
    sad = np.exp(rng.uniform(np.log(0.1), np.log(10.0), shape)) * 1e-8
    q[synth.L["h2so4"]] = synth.h2so4_for_sad(T, p, q[synth.L["h2o"]], sad)
    return synth.StepInputs(g, 900.0, T, p, sza, synth.altitude_km(p), q, h, sad,
                            {"name": "wet", "seed": seed})


DTSEDI = 10800.0       # 12 sub-steps of 900 s
NIVH2O = 59            # nivh2oecmwf, F:18


def input_producers(seed: int = 11) -> dict:
    """Inputs for the reference's zenith and altitude.  zenith: 20 levels x 180
    trajectory points anywhere on the globe, four day numbers.  altitude: a 16 x 16 x 20
    regular grid (the model's 2-degree spacing, the polar cap 60-90 N, 0-30 E) with hybrid
    levels p = a + b psol, and trajectory points in that cap, pressures from above the top
    level to below the bottom one (both clamps)."""
    rng = np.random.default_rng(seed)
    r = lambda a, q: np.round(a / q) * q          # short decimals: the fixture compresses
    d = {"days": np.array([1.0, 80.5, 172.25, 355.9]),
         "zlon": r(rng.uniform(0.0, 360.0, (20, 180)), 1e-4),
         "zlat": r(rng.uniform(-90.0, 90.0, (20, 180)), 1e-4)}
    m, nl = 16, 20
    eta = (np.arange(nl) + 0.5) / nl
    d["bbb"] = eta ** 3
    d["aaa"] = 0.5 + 150.0 * eta * (1.0 - eta)
    d["psol"] = r(rng.uniform(950.0, 1030.0, (m, m)), 1e-3)
    d["relief"] = r(rng.uniform(0.0, 2.0, (m, m)), 1e-4)
    p = d["aaa"][:, None, None] + d["bbb"][:, None, None] * d["psol"][None]
    d["t"] = r(210.0 + 80.0 * eta[:, None, None] + rng.normal(0, 4, (nl, m, m)), 1e-3)
    d["h2o"] = r(np.exp(rng.uniform(np.log(2e-6), np.log(1e-2), (nl, m, m))), 1e-9)
    d["tlon"] = r(rng.uniform(0.0, 30.9, (nl, m, m)), 1e-4)
    d["tlat"] = r(rng.uniform(59.1, 90.0, (nl, m, m)), 1e-4)
    d["tpre"] = r(np.exp(rng.uniform(np.log(0.3), np.log(1100.0), (nl, m, m))), 1e-4)
    return d


def save(name, inp, ref):
    g = inp.grid
    np.savez_compressed(
        os.path.join(OUT, name + ".npz"),
        grid=np.array([g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau]), dt=inp.dt,
        T=inp.T, p=inp.p, sza=ref["sza3d"], alt=inp.alt, qj1_in=inp.qj1, hc_in=inp.hc,
        qj1_chem=ref["qj1_chem"], hc_chem=ref["hc_chem"], surfarea=ref["surfarea"],
        vsed3d=ref["vsed3d"], parth2osed=ref["parth2osed"], parthno3sed=ref["parthno3sed"],
        qj1=ref["qj1"], hc=ref["hc"],
        **({"qj1_sedi": ref["qj1_sedi"], "dtsedi": DTSEDI, "nivh2o": NIVH2O}
           if "qj1_sedi" in ref else {}))
    print("wrote", name, g.nlon * g.nlat * g.nivbas, "boxes")


def main():
    if not refrun.available():
        raise SystemExit("reference harness not built: make -C oracle ref")
    c0 = synth.make_c0(0)
    sub = subset_columns(c0, list(range(16)))
    save("c0_seed0_16col", sub, refrun.run_reference(sub))
    psc = subset_columns(psc_grid(), list(range(8)))
    save("psc_seed7_8col", psc, refrun.run_reference(psc))
    sedi = sedi_grid()
    save("sedi_seed9_8col", sedi, refrun.run_reference(sedi, dtsedi=DTSEDI))
    wet_fixture()
    d = input_producers()
    sza, alt = refrun.run_inputs_reference(d["days"], d["zlon"], d["zlat"], d["relief"],
                                           d["psol"], d["aaa"], d["bbb"], d["t"], d["h2o"],
                                           d["tlon"], d["tlat"], d["tpre"])
    np.savez_compressed(os.path.join(OUT, "inputs_seed11.npz"), sza=sza, alt=alt, **d)
    print("wrote inputs_seed11", sza.size, "SZA values,", alt.size, "altitudes")
    ecmwf_fixture()
    transport_fixture()
    niter_fixtures()


def columns_of(inp, cols):
    """The listed columns (flat index lat * nlon + lon) of any grid, as a 1 x ncol grid."""
    g = inp.grid
    ncol = g.nlat * g.nlon
    out = inp.copy()
    out.grid = synth.Grid(len(cols), 1, g.niv, g.nivbas, g.nivhau)
    flat = lambda a: np.ascontiguousarray(a.reshape(a.shape[:-2] + (ncol,))[..., cols][..., None, :])
    for name in ("T", "p", "sza", "alt", "sad_target"):
        a = getattr(inp, name)
        if a is not None:
            setattr(out, name, flat(a))
    out.qj1, out.hc = flat(inp.qj1), flat(inp.hc)
    return out


def niter_fixtures():
    """The reference built with niter (F:630) raised: 60 sweeps on the c0 columns, and 400
    sweeps on 16 columns of the c2 grid (8 in the polar vortex, liquid PSC and cold
    heterogeneous chemistry, 8 across the other latitudes, day and night) -- there the
    sweeps have reached the fixed point the converged mode solves for, except at the few
    boxes that converge slower or not at all."""
    c0 = synth.make_c0(0)
    sub = subset_columns(c0, list(range(16)))
    save("c0_seed0_16col_n60", sub, refrun.run_reference(sub, niter=60))
    c2 = synth.make_c2(1, nlon=24, nlat=13)
    nlon = c2.grid.nlon
    # rows 0-1 are poleward of 60 S (the vortex); 8 columns there, 8 spread over the rest
    cols = [0, 5, 11, 17, nlon + 2, nlon + 8, nlon + 14, nlon + 20]
    cols += [r * nlon + c for r, c in ((3, 1), (4, 7), (6, 13), (7, 19), (9, 4), (10, 10),
                                        (11, 16), (12, 22))]
    vx = columns_of(c2, cols)
    save("c2_seed1_16col_n400", vx, refrun.run_reference(vx, niter=400))


def transport_fixture(seed: int = 0, nsamp: int = 4000):
    """The reference's transport weights and tricubic interpolation (initransp, posx..poidz,
    internew; oracle/ref_transport.f90) on synth.make_transport(seed), 180 x 91 x 137 x 3
    tracers.  The full arrays are hundreds of MB: the fixture keeps the values at sampled
    points (random ones plus points whose stencil crosses a pole or the date line), the sums
    of every output array, and sums of the inputs (the generator must still give them)."""
    tr = synth.make_transport(seed)
    idx, w, q = refrun.run_transport_reference(tr)
    n = idx.shape[1]
    rng = np.random.default_rng(seed)
    polar = np.flatnonzero((idx[1] == 1) | (idx[1] == 90))
    wrap = np.flatnonzero((idx[0] == 1) | (idx[0] >= 179))
    pts = np.unique(np.concatenate([rng.integers(0, n, nsamp), rng.choice(polar, 1000),
                                    rng.choice(wrap, 1000)]))
    qf = q.reshape(q.shape[0], -1)
    np.savez_compressed(os.path.join(OUT, f"transport_seed{seed}.npz"), seed=seed, jc=tr.jc,
                        pts=pts, idx=idx[:, pts], w=w[:, pts], q=qf[:, pts],
                        q_sum=qf.sum(axis=1), w_sum=w.sum(axis=1), idx_sum=idx.astype(np.int64).sum(axis=1),
                        in_sums=np.array([tr.x0.sum(), tr.y0.sum(), tr.z0.sum(), tr.p0.sum(), tr.q.sum()]))
    print(f"wrote transport_seed{seed}: {pts.size} sampled points of {n}")


def ecmwf_fixture(seed: int = 13, nsamp: int = 3000):
    """The reference's ECMWF reader on a forcing file written by the library's writer
    (three analysis times, 0/3/6 h; the model at 03 UT reads the 3 h and 6 h ones)."""
    import tempfile

    sys.path.insert(0, os.path.dirname(OUT))
    import forcing_synth as fs

    y, m, d = fs.DATE
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, f"ecmwf_{y:04d}{m:02d}{d:02d}")
        fs.write_file(path, seed, 3)
        idx = np.random.default_rng(seed).integers(1, 180 * 91 * 137 + 1, nsamp)
        day = (d - 1) + 3.0 / 24.0          # gregtojul(2023, 1, 15, 3, 0)
        ref = refrun.run_ecmwf_reference(path, y, day, idx)
    np.savez_compressed(os.path.join(OUT, f"ecmwf_seed{seed}.npz"), idx=idx, day=day, ian=y,
                        seed=seed, **ref)
    print(f"wrote ecmwf_seed{seed}: days {ref['daym1']} {ref['dayp1']}")


def wet_fixture():
    wet = wet_grid()
    save("wet_seed15_8col", wet, refrun.run_reference(wet))


if __name__ == "__main__":
    if sys.argv[1:] == ["transport"]:
        transport_fixture()
    elif sys.argv[1:] == ["wet"]:
        wet_fixture()
    elif sys.argv[1:] == ["niter"]:
        niter_fixtures()
    else:
        main()
