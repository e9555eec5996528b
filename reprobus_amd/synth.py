"""Synthetic chemistry-step inputs for the benchmark and test grids.

Arrays follow the model's Fortran layout (``qj1(nlon,nlat,niv,nbcon)``: lon
fastest, then lat, level, species), held here as C-ordered numpy arrays of shape
``(nspecies, niv, nlat, nlon)`` -- the same bytes.  Level 1 is the model top.

Grids
  * ``c0``: 64 columns x 60 levels (3,840 boxes), one 900 s step; T uniform
    185-270 K per box, cos SZA uniform -0.3..1 per box, sulfate SAD log-uniform
    0.1-10 um2/cm3.
  * ``c1``: 144 lon x 91 lat x 60 levels (786,240 boxes); T from a latitude
    profile (200 K polar winter to 270 K stratopause) plus 5 K Gaussian noise;
    SZA from latitude, longitude and a fixed solar declination.
Both use 60 levels log-spaced from 1 hPa (top) to 100 hPa, ``alt`` from a 7 km
scale height, and species drawn log-uniformly within stratospheric ranges with
the Cly / NOy / Bry family totals fixed per box and partitioned among members.

Surface area is not an input of the reference step: it derives liquid-aerosol
area from H2SO4 (``analytic``, src/reprobus_l137_20230723.f:2919-3096).  The
requested SAD is therefore realised by choosing H2SO4 from the binary
H2SO4/H2O aerosol relation of that routine (aliq proportional to qH2SO4**(2/3) at
fixed T and water vapour); boxes that go ternary (T < 215 K) or form NAT/ice
realise a different area, which the step reports in ``surfarea``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

NBCON, NCM = 43, 15

# Long-lived slots, src/reprobus_l137_20230723.f:720-730 (1-based there).
LONG = ["n2o", "ch4", "h2o", "noy", "hno3", "n2o5", "cly", "ox", "co", "oclo",
        "ox_passive", "h2so4", "hcl", "clono2", "hocl", "cl2", "h2o2", "clno2", "hbr",
        "brono2", "nox", "ho2no2", "clox", "brox", "cl2o2", "hobr", "brcl", "ch2o",
        "ch3o2", "ch3o2h", "cfc11", "cfc12", "cfc113", "ccl4", "ch3ccl3", "ch3cl",
        "hcfc22", "ch3br", "h1211", "h1301", "bry", "ch2br2", "hno3_gas"]
# Short-lived slots, F:743-746.
SHORT = ["o1d", "oh", "cl", "o3p", "o3", "ho2", "no2", "no", "br", "n", "clo", "bro",
         "no3", "h", "ch3"]
L = {n: i for i, n in enumerate(LONG)}
S = {n: i for i, n in enumerate(SHORT)}

NLEV = 60
P_TOP, P_BOT = 1.0, 100.0          # hPa
NIVHAU = 14                        # first level with heterogeneous chemistry (F:2425)

# log-uniform ranges (mixing ratio) for species not tied to a family
_RANGES = {
    "n2o": (2e-8, 3.2e-7), "ch4": (2e-7, 1.8e-6), "h2o": (2e-6, 8e-6),
    "co": (1e-8, 6e-8), "h2o2": (1e-12, 1e-10), "ch2o": (1e-12, 1e-10),
    "ch3o2": (1e-15, 1e-11), "ch3o2h": (1e-12, 1e-10),
    "cfc11": (1e-13, 2.6e-10), "cfc12": (1e-11, 5.3e-10), "cfc113": (1e-13, 8e-11),
    "ccl4": (1e-14, 1e-10), "ch3ccl3": (1e-14, 2e-11), "ch3cl": (1e-12, 5.5e-10),
    "hcfc22": (1e-11, 2.5e-10), "ch3br": (1e-14, 7e-12), "h1211": (1e-14, 4e-12),
    "h1301": (1e-14, 3.5e-12), "ch2br2": (1e-14, 5e-12),
}
O3_RANGE = (1e-7, 1e-5)
NOY_RANGE = (1e-10, 2e-8)
CLY_RANGE = (1e-10, 3.5e-9)
BRY_RANGE = (1e-12, 2.2e-11)
# Short-lived radicals, log-uniform.  The nominal band is 1e-15..1e-10; O(1D), H, CH3
# and N sit below it and O(3P), Cl, Br are narrowed to their stratospheric
# magnitudes, because the first sweep uses the incoming values before it
# recomputes them (F:1358-1370): an O(1D) of 1e-10 (10^6 x physical) makes the
# CH3 -> CH3O2 -> HOx chain run away to mixing ratios of order 100.
RADICAL_RANGES = {
    "o1d": (1e-19, 1e-16), "oh": (1e-14, 1e-11), "cl": (1e-15, 1e-12), "o3p": (1e-12, 1e-9),
    "ho2": (1e-13, 1e-10), "br": (1e-15, 1e-12), "n": (1e-18, 1e-14), "no3": (1e-14, 1e-10),
    "h": (1e-18, 1e-14), "ch3": (1e-20, 1e-16), "no2": (1e-15, 1e-10), "no": (1e-15, 1e-10),
    "clo": (1e-15, 1e-10), "bro": (1e-15, 1e-10), "o3": O3_RANGE,
}


@dataclass
class Grid:
    nlon: int
    nlat: int
    niv: int
    nivbas: int
    nivhau: int = NIVHAU

    @property
    def nbox(self) -> int:
        return self.nlon * self.nlat * self.nivbas

    @property
    def shape3(self) -> tuple:
        return (self.niv, self.nlat, self.nlon)


@dataclass
class StepInputs:
    """One chemistry step's inputs (numpy, Fortran byte order)."""
    grid: Grid
    dt: float
    T: np.ndarray            # (niv, nlat, nlon) K
    p: np.ndarray            # (niv, nlat, nlon) hPa
    sza: np.ndarray          # (niv, nlat, nlon) deg
    alt: np.ndarray          # (niv, nlat, nlon) km
    qj1: np.ndarray          # (43, niv, nlat, nlon) mixing ratio
    hc: np.ndarray           # (15, niv, nlat, nlon) mixing ratio
    sad_target: np.ndarray | None = None   # (niv, nlat, nlon) cm2/cm3
    meta: dict = field(default_factory=dict)

    def copy(self) -> "StepInputs":
        return StepInputs(self.grid, self.dt, self.T.copy(), self.p.copy(), self.sza.copy(),
                          self.alt.copy(), self.qj1.copy(), self.hc.copy(),
                          None if self.sad_target is None else self.sad_target.copy(),
                          dict(self.meta))


def level_pressures(nlev: int = NLEV) -> np.ndarray:
    """Levels log-spaced from P_TOP (level 1) to P_BOT (level nlev), hPa."""
    return np.exp(np.linspace(np.log(P_TOP), np.log(P_BOT), nlev))


def altitude_km(p_hpa: np.ndarray) -> np.ndarray:
    """Geometric altitude from pressure with a 7 km scale height."""
    return 7.0 * np.log(1013.25 / p_hpa)


def _logu(rng, lo, hi, shape):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=shape))


def _partition(rng, total, nmember, shape, lo=0.02):
    """Split ``total`` into ``nmember`` positive parts with log-uniform weights."""
    w = _logu(rng, lo, 1.0, (nmember,) + shape)
    return total[None] * w / w.sum(axis=0, keepdims=True)


def draw_species(rng: np.random.Generator, shape: tuple) -> tuple[np.ndarray, np.ndarray]:
    """Long-lived (43,*shape) and short-lived (15,*shape) mixing ratios."""
    q = np.zeros((NBCON,) + shape)
    h = np.zeros((NCM,) + shape)
    for name, (lo, hi) in _RANGES.items():
        q[L[name]] = _logu(rng, lo, hi, shape)

    # Bry = HBr + BrONO2 + BrOx + HOBr + BrCl   (conserv, F:2591-2598)
    bry = _logu(rng, *BRY_RANGE, shape)
    hbr, brono2, brox, hobr, brcl = _partition(rng, bry, 5, shape)
    q[L["bry"]], q[L["hbr"]], q[L["brono2"]] = bry, hbr, brono2
    q[L["brox"]], q[L["hobr"]], q[L["brcl"]] = brox, hobr, brcl

    # Cly - BrCl = OClO + HCl + ClONO2 + HOCl + 2 Cl2 + ClNO2 + ClOx + 2 Cl2O2  (F:2541-2551)
    cly = _logu(rng, *CLY_RANGE, shape)
    parts = _partition(rng, cly - brcl, 8, shape)
    oclo, hcl, clono2, hocl, cl2x2, clno2, clox, cl2o2x2 = parts

    # NOy - ClONO2 - ClNO2 - BrONO2 = HNO3 + 2 N2O5 + NOx + HNO4  (F:2566-2580);
    # keep the chlorine/bromine nitrates under half of NOy so the NOy members stay positive.
    noy = _logu(rng, *NOY_RANGE, shape)
    nitr = clono2 + clno2 + brono2
    cap = 0.5 * noy
    scale = np.where(nitr > cap, np.maximum(cap - brono2, 0.0) / np.maximum(clono2 + clno2, 1e-300), 1.0)
    scale = np.minimum(scale, 1.0)
    moved = (clono2 + clno2) * (1.0 - scale)
    clono2, clno2, hcl = clono2 * scale, clno2 * scale, hcl + moved
    rest = noy - clono2 - clno2 - brono2
    hno3, n2o5x2, nox, hno4 = _partition(rng, rest, 4, shape)

    q[L["cly"]], q[L["oclo"]], q[L["hcl"]], q[L["clono2"]] = cly, oclo, hcl, clono2
    q[L["hocl"]], q[L["cl2"]], q[L["clno2"]], q[L["clox"]] = hocl, 0.5 * cl2x2, clno2, clox
    q[L["cl2o2"]] = 0.5 * cl2o2x2
    q[L["noy"]], q[L["hno3"]], q[L["n2o5"]] = noy, hno3, 0.5 * n2o5x2
    q[L["nox"]], q[L["ho2no2"]] = nox, hno4
    q[L["hno3_gas"]] = hno3

    # Ox = O3 + O(3P) (F:1400-1408); passive Ox starts equal to Ox.
    o3 = _logu(rng, *O3_RANGE, shape)
    for name in SHORT:
        h[S[name]] = _logu(rng, *RADICAL_RANGES[name], shape)
    h[S["o3"]] = o3
    h[S["no2"]] = nox * rng.uniform(0.3, 0.9, shape)
    h[S["no"]] = np.maximum(nox - h[S["no2"]], 0.0) * rng.uniform(0.5, 1.0, shape)
    h[S["clo"]] = clox * rng.uniform(0.5, 1.0, shape)
    h[S["bro"]] = brox * rng.uniform(0.3, 1.0, shape)
    q[L["ox"]] = o3 + h[S["o3p"]]
    q[L["ox_passive"]] = q[L["ox"]]
    return q, h


# ---- SAD -> H2SO4 (binary liquid aerosol relation of the reference, see module doc) ----
_KS = (-21.661, 2724.2, 51.81, -15732.0, 47.004, -6969.0, -4.6183)   # F:2725, Carslaw 1995
_DENS_X = (2.393284E-02, -4.359335E-05, 7.961181E-08, 0.0, -0.198716351,
           1.39564574E-03, -2.020633E-06, 0.51684706, -3.0539E-03, 4.505475E-06)


def _binary_density(ws, T):
    x = _DENS_X
    v1 = x[0] + x[1] * T + x[2] * T ** 2 + x[3] * T ** 3
    vs = x[4] + x[5] * T + x[6] * T ** 2 + (x[7] + x[8] * T + x[9] * T ** 2) * ws
    # the third density term multiplies w*w with (X11..X13); fold it in:
    vs = vs + (-0.30119511 + 1.840408E-03 * T - 2.7221253742E-06 * T ** 2) * ws * ws
    vmcal = (1.0 - ws) / 18.0160 * v1 + vs * ws / 98.080
    return 1.0 / vmcal * 0.001


def h2so4_for_sad(T, p, qh2o, sad):
    """H2SO4 mixing ratio giving binary liquid-aerosol area ``sad`` (cm2/cm3)."""
    pw0 = np.minimum(qh2o * p / 1013.0, 2e-3 / 1013.0)
    tice = 2668.70 / (10.4310 - (np.log(pw0) + np.log(760.0)) / np.log(10.0))
    tt = np.minimum(T, 273.0)
    ph2o = 10.0 ** (-2668.7 / tt + 10.4310) / 760.0
    parth2o = np.where(tt <= tice, np.clip(1.0 - np.maximum(pw0 - ph2o, 0.0) / pw0, 0, 1), 1.0)
    t = np.maximum(T, np.maximum(tice - 3.0, 185.0))
    pw = np.clip(qh2o * parth2o * p / 1013.0, 2e-5 / 1013.0, 2e-3 / 1013.0)
    ks = _KS
    a = ks[2] + ks[3] / t
    b = ks[0] + ks[1] / t
    c = ks[4] + ks[5] / t + ks[6] * np.log(t) - np.log(pw)
    det = np.maximum(b * b - 4 * a * c, 1e-30)
    xsb = (-b - np.sqrt(det)) / (2 * a)
    msb = 55.51 * xsb / (1.0 - xsb)
    ws = np.maximum(msb * 0.098076 / (1.0 + msb * 0.098076), 0.005)
    dens = _binary_density(ws, T)
    ns_per_q = p * 100.0 / 8.314 / t
    vliq_per_q = ns_per_q * 98.076e-6 / ws / dens
    ls2 = np.log(1.8) ** 2
    ndrop = 10.0
    # aliq = 4 pi ndrop e^{2 ls2} * (3 vliq / (4 pi ndrop) e^{-4.5 ls2})^{2/3}
    k = 4 * np.pi * ndrop * np.exp(2 * ls2) * (3 * vliq_per_q / (4 * np.pi * ndrop) * np.exp(-4.5 * ls2)) ** (2.0 / 3.0)
    return (sad / k) ** 1.5


def make_c0(seed: int = 0) -> StepInputs:
    """Grid c0: 64 columns x 60 levels, one 900 s step."""
    rng = np.random.default_rng(seed)
    g = Grid(nlon=64, nlat=1, niv=NLEV, nivbas=NLEV)
    shape = g.shape3
    p = np.broadcast_to(level_pressures()[:, None, None], shape).copy()
    T = rng.uniform(185.0, 270.0, shape)
    cossza = rng.uniform(-0.3, 1.0, shape)
    sza = np.degrees(np.arccos(cossza))
    q, h = draw_species(rng, shape)
    sad = _logu(rng, 0.1, 10.0, shape) * 1e-8
    q[L["h2so4"]] = h2so4_for_sad(T, p, q[L["h2o"]], sad)
    return StepInputs(g, 900.0, T, p, sza, altitude_km(p), q, h, sad,
                      {"name": "c0", "seed": seed, "cossza": cossza})


def _latlon(nlon, nlat):
    lat = np.linspace(-90.0, 90.0, nlat)
    lon = np.arange(nlon) * (360.0 / nlon)
    return lat, lon


def solar_zenith(lat_deg, lon_deg, decl_deg=20.0, ut_hours=12.0):
    """SZA (deg) from latitude, longitude and a fixed declination (noon at lon 0 at 12 UT)."""
    phi = np.radians(lat_deg)
    dec = np.radians(decl_deg)
    h = np.radians((ut_hours - 12.0) * 15.0 + lon_deg)
    c = np.sin(phi) * np.sin(dec) + np.cos(phi) * np.cos(dec) * np.cos(h)
    return np.degrees(np.arccos(np.clip(c, -1.0, 1.0)))


def temperature_profile(lat_deg, p_hpa):
    """200 K at the winter (south) pole / ~215 K elsewhere at 100 hPa, 270 K at 1 hPa."""
    s = np.log(P_BOT / p_hpa) / np.log(P_BOT / P_TOP)           # 0 at 100 hPa, 1 at 1 hPa
    t_low = 215.0 - 15.0 * np.maximum(-np.sin(np.radians(lat_deg)), 0.0) ** 2 + 5.0 * np.maximum(np.sin(np.radians(lat_deg)), 0.0)
    return t_low + (270.0 - t_low) * s


def make_c1(seed: int = 0, nlon: int = 144, nlat: int = 91, decl_deg: float = 20.0,
            ut_hours: float = 12.0) -> StepInputs:
    """Grid c1: 144 x 91 x 60 boxes, one step (smaller nlon/nlat allowed for tests)."""
    rng = np.random.default_rng(seed)
    g = Grid(nlon=nlon, nlat=nlat, niv=NLEV, nivbas=NLEV)
    shape = g.shape3
    lat, lon = _latlon(nlon, nlat)
    p1 = level_pressures()
    p = np.broadcast_to(p1[:, None, None], shape).copy()
    LAT = np.broadcast_to(lat[None, :, None], shape)
    LON = np.broadcast_to(lon[None, None, :], shape)
    T = temperature_profile(LAT, p) + rng.normal(0.0, 5.0, shape)
    sza = np.broadcast_to(solar_zenith(LAT[0], LON[0], decl_deg, ut_hours)[None], shape).copy()
    q, h = draw_species(rng, shape)
    sad = _logu(rng, 0.1, 10.0, shape) * 1e-8
    q[L["h2so4"]] = h2so4_for_sad(T, p, q[L["h2o"]], sad)
    return StepInputs(g, 900.0, T, p, sza, altitude_km(p), q, h, sad,
                      {"name": "c1", "seed": seed, "decl_deg": decl_deg, "lat": lat, "lon": lon})


C2_STEPS = 96                 # one day of 900 s steps


def c2_sza(meta: dict, step: int) -> np.ndarray:
    """SZA (niv, nlat, nlon) of step `step` of the c2 day: 15 min per step from 0 UT."""
    lat, lon = meta["lat"], meta["lon"]
    LAT = np.broadcast_to(lat[:, None], (lat.size, lon.size))
    LON = np.broadcast_to(lon[None, :], (lat.size, lon.size))
    s2 = solar_zenith(LAT, LON, meta["decl_deg"], step * 0.25)
    return np.broadcast_to(s2[None], (NLEV,) + s2.shape).copy()


def make_c2(seed: int = 0, nlon: int = 144, nlat: int = 91, decl_deg: float = 0.0) -> StepInputs:
    """Grid c2 at step 0 of its day: c1's state with a polar vortex poleward of 60 S on the
    heterogeneous levels -- T ~ U(183, 195) K and liquid-PSC surface area
    logU(1, 50) um2/cm3 (realised through H2SO4) -- and a diurnal SZA (c2_sza) at the
    equinox (decl 0), so the vortex sees both sunlit and dark hours."""
    rng = np.random.default_rng(seed)
    inp = make_c1(seed, nlon, nlat, decl_deg, 0.0)
    g = inp.grid
    lat = inp.meta["lat"]
    vortex = np.zeros(g.shape3, bool)
    vortex[NIVHAU - 1:, lat <= -60.0, :] = True
    nv = int(vortex.sum())
    inp.T[vortex] = rng.uniform(183.0, 195.0, nv)
    sad = inp.sad_target
    sad[vortex] = _logu(rng, 1.0, 50.0, (nv,)) * 1e-8
    inp.qj1[L["h2so4"]] = h2so4_for_sad(inp.T, inp.p, inp.qj1[L["h2o"]], sad)
    inp.sza = c2_sza(inp.meta, 0)
    inp.meta.update({"name": "c2", "vortex": vortex})
    return inp


# ------------------------------------------------------------------ transport inputs
@dataclass
class TransportInputs:
    """Departure points and tracer fields of one semi-Lagrangian step on an nlon x nlat x niv
    grid (arrays numpy order (niv, nlat, nlon) = the model's (nlon, nlat, niv) memory)."""
    nlon: int
    nlat: int
    niv: int
    aaa: np.ndarray      # (niv,) hybrid coefficients, p = aaa + bbb psol (hPa)
    bbb: np.ndarray
    x0: np.ndarray       # departure longitude (rad, [0, 2 pi))
    y0: np.ndarray       # departure latitude (rad, inside (-pi/2, pi/2))
    z0: np.ndarray       # departure pressure (Pa)
    p0: np.ndarray       # surface pressure of the departure column (hPa)
    q: np.ndarray        # (ntr, niv, nlat, nlon) tracer mixing ratios
    jc: np.ndarray       # (ntr,) 1-based tracer slots of the fields


def transport_levels(niv: int) -> tuple:
    """Hybrid coefficients (aaa hPa, bbb) of niv levels, top first: p rises monotonically
    from ~0.01 hPa to the surface for any surface pressure above 500 hPa."""
    eta = (np.arange(1, niv + 1) / (niv + 0.5)) ** 1.5
    return 0.01 + 200.0 * eta * (1 - eta) ** 2, eta ** 3


def make_transport(seed: int = 0, nlon: int = 180, nlat: int = 91, niv: int = 137,
                   jc=(1, 5, 20), dt: float = 3600.0) -> TransportInputs:
    """Departure points of a 1 h step under a smooth synthetic wind (zonal jet with waves,
    meridional wobble, a little noise; vertical velocity in pressure coordinates proportional
    to pressure, ~0.7 % of p per hour with 10 % noise, as stratospheric motion scales), kept
    strictly inside the grid as the model's trajectory code keeps them; tracers are positive
    smooth fields with noise."""
    rng = np.random.default_rng(seed)
    xpi = 2.0 * np.arcsin(1.0)
    ddlon, ddlat = 2 * xpi / nlon, xpi / (nlat - 1)
    lon = ddlon * np.arange(nlon)
    lat = xpi / 2 - ddlat * np.arange(nlat)
    aaa, bbb = transport_levels(niv)
    L, LA, LO = np.meshgrid(np.arange(niv), lat, lon, indexing="ij")
    psol = 1000.0 + 30.0 * np.cos(2 * LA[0]) * np.sin(LO[0]) + rng.normal(0, 2.0, LO[0].shape)
    pres = (aaa[:, None, None] + bbb[:, None, None] * psol[None]) * 100.0
    R = 6.371e6
    u = (30.0 + 20.0 * np.cos(3 * LO)) * np.cos(LA) ** 2 * (1 + 0.5 * np.sin(L / 9.0)) \
        + rng.normal(0, 2.0, LO.shape)
    v = 8.0 * np.sin(2 * LA) * np.cos(2 * LO) + rng.normal(0, 1.0, LO.shape)
    w = pres * 2e-6 * (np.sin(LO + LA) + rng.normal(0, 0.1, LO.shape))
    x0 = np.mod(LO - u * dt / (R * np.maximum(np.cos(LA), 0.05)), 2 * xpi)
    x0 = np.where(x0 >= 2 * xpi, x0 - 2 * xpi, x0)
    y0 = np.clip(LA - v * dt / R, -xpi / 2 + 1e-7, xpi / 2 - 1e-7)
    z0 = np.maximum(pres - w * dt, 1.0)
    p0 = psol[None].repeat(niv, 0) + rng.normal(0, 1.0, LO.shape)
    q = np.empty((len(jc), niv, nlat, nlon))
    for t in range(len(jc)):
        a, b = rng.uniform(-3, 3), rng.uniform(0.5, 2.0)
        q[t] = np.exp(-18.0 + a * np.sin(LO + t) * np.cos(LA) + b * np.log(pres / 1e4 + 1.0)) \
            * (1.0 + 0.05 * rng.standard_normal(LO.shape))
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    return TransportInputs(nlon, nlat, niv, c(aaa), c(bbb), c(x0), c(y0), c(z0), c(p0), c(q),
                           np.asarray(jc, dtype=np.int32))


# ------------------------------------------------------------------ trajectory forcing
@dataclass
class ForcingRecord:
    """One forcing record of the trajectory code: winds u, v (m/s), omega w (Pa/s) and
    temperature t (K), numpy order (niv, nlat, nlon); ln of surface pressure (Pa), (nlat,
    nlon); the record's day number."""
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    t: np.ndarray
    lnps: np.ndarray
    day: float


@dataclass
class TrajectoryInputs:
    """The inputs of backtra: the grid, the hybrid coefficients, two forcing records
    around the step, and the step's timing (daynum at its end, dtdyn seconds in nrapp
    sub-steps; records every nintecmwf / ntime hours, as the model's instant1)."""
    nlon: int
    nlat: int
    niv: int
    aaa: np.ndarray
    bbb: np.ndarray
    m1: ForcingRecord
    p1: ForcingRecord
    daynum: float
    dtdyn: float
    ntime: int
    nintecmwf: int


def make_trajectory_inputs(seed: int = 0, nlon: int = 180, nlat: int = 91, niv: int = 137,
                           daynum: float = 100.1, dtdyn: float = 3600.0) -> TrajectoryInputs:
    """Two forcing records 6 h apart (days 100.0 and 100.25) of a smooth synthetic flow: a
    zonal jet with waves that crosses both poles (so the rotated-frame branch sees strong
    winds), a meridional wobble, omega proportional to pressure, temperature falling with
    height; the second record is the first advected a little, with noise."""
    rng = np.random.default_rng(seed)
    xpi = 2.0 * np.arcsin(1.0)
    lon = 2 * xpi / nlon * np.arange(nlon)
    lat = xpi / 2 - xpi / (nlat - 1) * np.arange(nlat)
    aaa, bbb = transport_levels(niv)
    L, LA, LO = np.meshgrid(np.arange(niv), lat, lon, indexing="ij")
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)

    def record(phase, day):
        psol = 1000.0 + 30.0 * np.cos(2 * LA[0]) * np.sin(LO[0] + phase) \
            + rng.normal(0, 2.0, LO[0].shape)
        pres = (aaa[:, None, None] + bbb[:, None, None] * psol[None]) * 100.0
        u = (30.0 + 20.0 * np.cos(3 * LO + phase)) * np.cos(LA) ** 2 * (1 + 0.5 * np.sin(L / 9.0)) \
            + 15.0 * np.sin(LA) ** 2 * np.cos(LO) + rng.normal(0, 2.0, LO.shape)
        v = 8.0 * np.sin(2 * LA) * np.cos(2 * LO + phase) - 15.0 * np.sin(LA) * np.sin(LO) \
            + rng.normal(0, 1.0, LO.shape)
        w = pres * 2e-6 * (np.sin(LO + LA + phase) + rng.normal(0, 0.1, LO.shape))
        t = 190.0 + 100.0 * (L + 0.5) / niv + 10.0 * np.cos(LA) * np.cos(LO + phase) \
            + rng.normal(0, 1.0, LO.shape)
        return ForcingRecord(c(u), c(v), c(w), c(t), c(np.log(psol * 100.0)), day)

    return TrajectoryInputs(nlon, nlat, niv, c(aaa), c(bbb), record(0.0, 100.0),
                            record(0.2, 100.25), daynum, dtdyn, 1, 6)
