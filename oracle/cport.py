"""Test infrastructure: ctypes wrapper of the C restatement (librbchem_oracle.so)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "librbchem_oracle.so")
# the restatement with correctly rounded math in the PSC block (built with -DRBO_CR_PSC)
LIB_CR = os.path.join(HERE, "librbchem_oracle_cr.so")


class _Grid(C.Structure):
    _fields_ = [("nlon", C.c_int), ("nlat", C.c_int), ("niv", C.c_int), ("nivbas", C.c_int),
                ("nivhau", C.c_int), ("niter", C.c_int)]


_lib = None
_libs = {}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "restatement"], check=True)


def lib(cr_psc: bool = False):
    """The restatement library; cr_psc selects the correctly rounded PSC variant."""
    global _lib
    path = LIB_CR if cr_psc else LIB
    if path == LIB and _lib is not None:
        return _lib
    if path in _libs:
        return _libs[path]
    if True:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.rbo_chem_step.argtypes = [C.POINTER(_Grid), C.c_double, dp, dp, dp, dp, dp, dp, dp, dp,
                                    dp, dp, dp, dp, C.c_int]
        L.rbo_chem_step.restype = C.c_int
        L.rbo_conserv_floor.argtypes = [C.POINTER(_Grid), dp, dp, C.c_int]
        L.rbo_sediment.argtypes = [C.POINTER(_Grid), C.c_double, C.c_int, dp, dp, dp, dp, dp, dp]
        L.rbo_sediment.restype = None
        L.rbo_zenith.argtypes = [C.c_int64, C.c_double, dp, dp, dp]
        L.rbo_zenith.restype = None
        L.rbo_altitude.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double] + [dp] * 10
        L.rbo_altitude.restype = None
        L.rbo_chem_converge.argtypes = [C.POINTER(_Grid), C.c_double, C.c_double] + [dp] * 12 + [
            np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"), C.c_void_p, C.c_void_p,
            C.c_int]
        L.rbo_chem_converge.restype = C.c_int
        L.rbo_chem_converge_aa.argtypes = [C.POINTER(_Grid), C.c_double, C.c_double, C.c_int,
                                           C.c_int, C.c_int] + \
            [dp] * 12 + [np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"), C.c_void_p,
                         C.c_void_p, C.c_int]
        L.rbo_chem_converge_aa.restype = C.c_int
        L.rbo_chem_surrogate.argtypes = [C.POINTER(_Grid), C.c_double, C.c_double, C.c_double,
                                         C.c_int, C.c_int, C.c_int] + \
            [dp] * 12 + [np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"), dp, dp,
                         C.c_int]
        L.rbo_chem_surrogate.restype = C.c_int
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.rbo_transport_weights.argtypes = [C.c_int, C.c_int, C.c_int] + [dp] * 6 + [ip, dp]
        L.rbo_transport_weights.restype = None
        L.rbo_internew.argtypes = [C.c_int, C.c_int, C.c_int, ip, dp, dp, dp]
        L.rbo_internew.restype = C.c_int64
        L.rbo_backtra.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, C.c_void_p, C.c_void_p,
                                  C.c_double, C.c_double, C.c_int, C.c_int, C.c_int] + \
            [C.c_void_p] * 10 + [C.c_int]
        L.rbo_backtra.restype = C.c_int
        L.rbo_troe_constants.argtypes = [dp]
        L.rbo_troe.argtypes = [C.c_double] * 6
        L.rbo_troe.restype = C.c_double
        L.rbo_rates.argtypes = [C.c_double, C.c_double, C.c_double, dp]
        L.rbo_rates.restype = C.c_int
        L.rbo_rate_name.argtypes = [C.c_int]
        L.rbo_rate_name.restype = C.c_char_p
        L.rbo_phot_box.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, dp]
        L.rbo_column.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, dp, dp, dp]
        L.rbo_jno_box.argtypes = [C.c_double] * 5
        L.rbo_jno_box.restype = C.c_double
        L.rbo_density.argtypes = [C.c_double] * 3
        L.rbo_density.restype = C.c_double
        L.rbo_aerosol_box.argtypes = [C.c_double, C.c_double, C.c_double, dp, dp, dp, dp, dp,
                                      np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")]
        _libs[path] = L
        if path == LIB:
            _lib = L
    return L


def chem_step(inp, ajl, o3up, niter: int = 5, nthreads: int = 0, conserv: bool = True,
              cr_psc: bool = False) -> dict:
    """Run the restatement on a synth.StepInputs; returns outputs like refrun."""
    g = inp.grid
    G = _Grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau, niter)
    q = np.ascontiguousarray(inp.qj1, dtype=np.float64).copy()
    h = np.ascontiguousarray(inp.hc, dtype=np.float64).copy()
    surf = np.zeros(g.shape3)
    shb = (g.nivbas, g.nlat, g.nlon)
    vsed, p2o, pno3 = np.zeros(shb), np.zeros(shb), np.zeros(shb)
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    ajl_f = np.ascontiguousarray(np.asarray(ajl, dtype=np.float64).transpose(3, 2, 1, 0))
    rc = lib(cr_psc).rbo_chem_step(C.byref(G), float(inp.dt), c(inp.T), c(inp.p), c(inp.sza), c(inp.alt),
                             ajl_f, c(o3up), q, h, surf, vsed, p2o, pno3, nthreads)
    if rc != 0:
        raise RuntimeError(f"rbo_chem_step failed: {rc}")
    out = {"qj1_chem": q.copy(), "hc_chem": h.copy(), "surfarea": surf, "vsed3d": vsed,
           "parth2osed": p2o, "parthno3sed": pno3}
    if conserv:
        lib(cr_psc).rbo_conserv_floor(C.byref(G), q, h, nthreads)
        out["qj1"], out["hc"] = q, h
    return out


def fortran_ajl(ajl):
    """ajl (40,101,27,7) in Fortran index order -> contiguous Fortran bytes."""
    return np.ascontiguousarray(np.asarray(ajl, dtype=np.float64).transpose(3, 2, 1, 0))


def chem_converge(inp, ajl, o3up, tol: float = 1e-10, maxiter: int = 2000, nthreads: int = 0,
                  conserv: bool = True, cr_psc: bool = False, guess=None, accel: int = 0,
                  aa_start: int = 0, aa_budget: int = 240) -> dict:
    """Converged mode of the restatement: sweeps to the fixed point (per-box counts in 'iters').
    guess = (qj1, hc) mixing ratios in state layout replaces the initial iterate; with accel
    the sweep cap is min(maxiter, aa_start + aa_budget), as the library's."""
    g = inp.grid
    G = _Grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau, maxiter)
    q = np.ascontiguousarray(inp.qj1, dtype=np.float64).copy()
    h = np.ascontiguousarray(inp.hc, dtype=np.float64).copy()
    surf = np.zeros(g.shape3)
    shb = (g.nivbas, g.nlat, g.nlon)
    vsed, p2o, pno3 = np.zeros(shb), np.zeros(shb), np.zeros(shb)
    iters = np.zeros(shb, dtype=np.int32)
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    L = lib(cr_psc)
    gq = gh = None
    if guess is not None:
        gq, gh = c(guess[0]), c(guess[1])
        assert gq.shape == q.shape and gh.shape == h.shape
    rc = L.rbo_chem_converge_aa(C.byref(G), float(inp.dt), float(tol), int(accel), int(aa_start),
                                int(aa_budget), c(inp.T),
                                c(inp.p), c(inp.sza), c(inp.alt), fortran_ajl(ajl), c(o3up), q, h,
                                surf, vsed, p2o, pno3, iters,
                                None if gq is None else gq.ctypes.data,
                                None if gh is None else gh.ctypes.data, nthreads)
    if rc != 0:
        raise RuntimeError(f"rbo_chem_converge failed: {rc}")
    out = {"qj1_chem": q.copy(), "hc_chem": h.copy(), "surfarea": surf, "vsed3d": vsed,
           "parth2osed": p2o, "parthno3sed": pno3, "iters": iters}
    if conserv:
        L.rbo_conserv_floor(C.byref(G), q, h, nthreads)
        out["qj1"], out["hc"] = q, h
    return out


def chem_surrogate(inp, ajl, o3up, sur, accept_tol: float = 1e-8, tol: float = 1e-11,
                   maxiter: int = 2000, nthreads: int = 0, conserv: bool = True,
                   cr_psc: bool = False, accel: int = 0, aa_start: int = 0,
                   aa_budget: int = 240) -> dict:
    """Surrogate step of the restatement: sur = (qj1, hc) mixing ratios in state layout, a
    surrogate's answer for the step.  Boxes whose first sweep from it changes every species
    by less than accept_tol keep it (iters 0); the rest sweep on to tol ('iters')."""
    g = inp.grid
    G = _Grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau, maxiter)
    q = np.ascontiguousarray(inp.qj1, dtype=np.float64).copy()
    h = np.ascontiguousarray(inp.hc, dtype=np.float64).copy()
    surf = np.zeros(g.shape3)
    shb = (g.nivbas, g.nlat, g.nlon)
    vsed, p2o, pno3 = np.zeros(shb), np.zeros(shb), np.zeros(shb)
    iters = np.zeros(shb, dtype=np.int32)
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    L = lib(cr_psc)
    sq, sh = c(sur[0]), c(sur[1])
    assert sq.shape == q.shape and sh.shape == h.shape
    rc = L.rbo_chem_surrogate(C.byref(G), float(inp.dt), float(accept_tol), float(tol),
                              int(accel), int(aa_start), int(aa_budget), c(inp.T), c(inp.p),
                              c(inp.sza),
                              c(inp.alt), fortran_ajl(ajl), c(o3up), q, h, surf, vsed, p2o, pno3,
                              iters, sq, sh, nthreads)
    if rc != 0:
        raise RuntimeError(f"rbo_chem_surrogate failed: {rc}")
    out = {"qj1_chem": q.copy(), "hc_chem": h.copy(), "surfarea": surf, "vsed3d": vsed,
           "parth2osed": p2o, "parthno3sed": pno3, "iters": iters}
    if conserv:
        L.rbo_conserv_floor(C.byref(G), q, h, nthreads)
        out["qj1"], out["hc"] = q, h
    return out


def sediment(inp, qj1, vsed3d, parth2osed, parthno3sed, dt: float, nivh2o: int,
             cr_psc: bool = False) -> np.ndarray:
    """Sedimentation of the restatement on a copy of qj1 (state layout)."""
    g = inp.grid
    G = _Grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau, 5)
    q = np.ascontiguousarray(qj1, dtype=np.float64).copy()
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    lib(cr_psc).rbo_sediment(C.byref(G), float(dt), int(nivh2o), c(inp.p), c(inp.T), q, c(vsed3d),
                             c(parth2osed), c(parthno3sed))
    return q


def zenith(day: float, lon, lat) -> np.ndarray:
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    lon, lat = c(lon), c(lat)
    out = np.empty_like(lon)
    lib().rbo_zenith(lon.size, float(day), lon, lat, out)
    return out


def altitude(relief, psol, aaa, bbb, t, h2o, tlon, tlat, tpre, dlon=2.0, dlat=2.0) -> np.ndarray:
    """Regular-grid arrays in numpy order: relief/psol (mlat, mlon), t/h2o/traj (mniv, mlat, mlon)."""
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    mniv, mlat, mlon = np.shape(t)
    out = np.empty((mniv, mlat, mlon))
    lib().rbo_altitude(mlon, mlat, mniv, dlon, dlat, c(relief), c(psol), c(aaa), c(bbb), c(t),
                       c(h2o), c(tlon), c(tlat), c(tpre), out)
    return out


def transport_weights(tr) -> tuple:
    """Cell indices (3, n) int32 and Lagrange weights (12, n) of the departure points of
    `tr` (synth.TransportInputs): posx/posy/posz and poidx/poidy/poidz (F:7371-7647)."""
    n = tr.nlon * tr.nlat * tr.niv
    idx = np.empty((3, n), dtype=np.int32)
    w = np.empty((12, n))
    lib().rbo_transport_weights(tr.nlon, tr.nlat, tr.niv, tr.aaa, tr.bbb, tr.x0.reshape(-1),
                                tr.y0.reshape(-1), tr.z0.reshape(-1), tr.p0.reshape(-1), idx, w)
    return idx, w


class _Forcing(C.Structure):
    _fields_ = [("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p), ("t", C.c_void_p),
                ("lnps", C.c_void_p), ("day", C.c_double)]


def _forcing(rec, keep):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rec.u, rec.v, rec.w, rec.t, rec.lnps)]
    keep.extend(arrs)
    return _Forcing(*[a.ctypes.data for a in arrs], float(rec.day))


def backtra(ti, nrapp: int, nthreads: int = 0) -> dict:
    """rbo_backtra on synth.TrajectoryInputs: departure points x0 y0 z0 p0, cells idx (3, n),
    weights w (12, n) and the trajectory traj{lon,lat,pre,tem} (nrapp, niv, nlat, nlon)."""
    L = lib()
    n = ti.nlon * ti.nlat * ti.niv
    keep = []
    m1, p1 = _forcing(ti.m1, keep), _forcing(ti.p1, keep)
    out = {k: np.zeros(n) for k in ("x0", "y0", "z0", "p0")}
    idx = np.zeros((3, n), dtype=np.int32)
    w = np.zeros((12, n))
    tr = {k: np.zeros((nrapp, ti.niv, ti.nlat, ti.nlon)) for k in ("trajlon", "trajlat", "trajpre", "trajtem")}
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    rc = L.rbo_backtra(ti.nlon, ti.nlat, ti.niv, c(ti.aaa), c(ti.bbb), C.byref(m1), C.byref(p1),
                       float(ti.daynum), float(ti.dtdyn), int(nrapp), int(ti.ntime),
                       int(ti.nintecmwf), *[out[k].ctypes.data for k in ("x0", "y0", "z0", "p0")],
                       idx.ctypes.data, w.ctypes.data,
                       *[tr[k].ctypes.data for k in ("trajlon", "trajlat", "trajpre", "trajtem")],
                       int(nthreads))
    if rc < 0:
        raise RuntimeError(f"rbo_backtra failed: {rc}")
    out.update(tr)
    out["outside"] = rc
    out["idx"], out["w"] = idx, w
    return out


def internew(tr, idx, w, q) -> tuple:
    """One tracer field q (niv, nlat, nlon) interpolated at the departure points (internew,
    F:8339-8618); returns (field, number of points outside the grid)."""
    out = np.zeros_like(np.ascontiguousarray(q, dtype=np.float64))
    bad = lib().rbo_internew(tr.nlon, tr.nlat, tr.niv, np.ascontiguousarray(idx, dtype=np.int32),
                             np.ascontiguousarray(w), np.ascontiguousarray(q, dtype=np.float64), out)
    return out, int(bad)
