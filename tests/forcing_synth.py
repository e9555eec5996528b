"""Synthetic ECMWF forcing content (deterministic from a seed), shared by the fixture
generator and the record tests.  Values are rounded to short decimals so the fixture's
samples compress; humidity includes zero and negative values to exercise the floor."""
import numpy as np

# GRIB parameter codes (ECMWF table 128) in the file order T, W, H(q), U, V; surface p
PAR = (130, 135, 133, 131, 132)
PAR_PS = 134
DATE = (2023, 1, 15)


def header(par, lev, hour, year=DATE[0], month=DATE[1], day=DATE[2]):
    """The 18 integers reprobus-extract-ecmwf.sh writes before each field."""
    return np.array([128, 98, 255, 0, 128, par, 109, lev, 0, year, month, day, hour, 0, 1, 0, 0, 0],
                    dtype=np.int32)


def fields(seed, itime, nlon=180, nlat=91, niv=137):
    rng = np.random.default_rng(seed * 100 + itime)
    shp = (niv, nlat, nlon)
    r = lambda a, q: np.round(a / q) * q
    lev = (np.arange(niv) + 0.5)[:, None, None] / niv
    d = {"ps": r(rng.uniform(500.0, 1040.0, (nlat, nlon)), 1e-3),
         "t": r(190.0 + 100.0 * lev + rng.normal(0, 3, shp), 1e-3),
         "w": r(rng.normal(0, 0.05, shp), 1e-5),
         "h": r(rng.uniform(-2e-6, 1e-2, shp) * lev ** 2, 1e-9),
         "u": r(rng.normal(0, 15, shp), 1e-3),
         "v": r(rng.normal(0, 10, shp), 1e-3)}
    hour = 3 * itime
    d["idate_ps"] = header(PAR_PS, 0, hour)
    d["idate_lev"] = np.stack([np.stack([header(p, k + 1, hour) for p in PAR]) for k in range(niv)])
    return d


def write_file(path, seed, ntimes, nlon=180, nlat=91, niv=137):
    from reprobus_amd import records

    for it in range(ntimes):
        d = fields(seed, it, nlon, nlat, niv)
        records.write_ecmwf_time(path, it > 0, d["idate_ps"], d["ps"], d["idate_lev"], d["t"],
                                 d["w"], d["h"], d["u"], d["v"])
