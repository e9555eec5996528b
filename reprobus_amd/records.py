"""The model's on-disk records through librbchem.so (C ABI: include/rbio.h): the
big-endian history/restart record and the ECMWF forcing file."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import RbError

NLON, NLAT, NIV = 180, 91, 137      # the model's grid


class HistoryHeader(C.Structure):
    _fields_ = [("namexp", C.c_char * 7), ("ian", C.c_int32), ("imois", C.c_int32),
                ("ijour", C.c_int32), ("iheure", C.c_int32), ("imin", C.c_int32)]


@dataclass
class History:
    namexp: str
    date: tuple              # (ian, imois, ijour, iheure, imin)
    pj1: np.ndarray          # (nlat, nlon)
    uj1: np.ndarray          # (niv, nlat, nlon)
    vj1: np.ndarray
    alt: np.ndarray
    tj1: np.ndarray
    qj1: np.ndarray          # (43, niv, nlat, nlon)
    surfarea: np.ndarray
    hc: np.ndarray           # (15, niv, nlat, nlon)

    FIELDS = ("pj1", "uj1", "vj1", "alt", "tj1", "qj1", "surfarea", "hc")


def _check(L, rc):
    if rc != 0:
        raise RbError(rc, L.rb_io_last_error().decode())


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data


def write_history(path: str, hist: History) -> None:
    L = _lib.load()
    niv, nlat, nlon = hist.uj1.shape
    h = HistoryHeader(hist.namexp.encode()[:6], *hist.date)
    keep = [_c(getattr(hist, f)) for f in History.FIELDS]
    _check(L, L.rb_history_write(path.encode(), nlon, nlat, niv, C.byref(h), *[k[1] for k in keep]))


def read_history(path: str, nlon: int = NLON, nlat: int = NLAT, niv: int = NIV) -> History:
    L = _lib.load()
    shp = (niv, nlat, nlon)
    arrs = {"pj1": np.empty((nlat, nlon)), "uj1": np.empty(shp), "vj1": np.empty(shp),
            "alt": np.empty(shp), "tj1": np.empty(shp), "qj1": np.empty((43,) + shp),
            "surfarea": np.empty(shp), "hc": np.empty((15,) + shp)}
    h = HistoryHeader()
    _check(L, L.rb_history_read(path.encode(), nlon, nlat, niv, C.byref(h),
                                *[arrs[f].ctypes.data for f in History.FIELDS]))
    return History(h.namexp.decode(), (h.ian, h.imois, h.ijour, h.iheure, h.imin), **arrs)


def write_ecmwf_time(path: str, append: bool, idate_ps, ps, idate_lev, t, w, h, u, v) -> None:
    """Append one analysis time: ps (nlat, nlon); t..v (niv, nlat, nlon); idate_ps (18,);
    idate_lev (niv, 5, 18) in the order T, W, H, U, V."""
    L = _lib.load()
    niv, nlat, nlon = np.shape(t)
    ip = np.ascontiguousarray(idate_ps, dtype=np.int32)
    il = np.ascontiguousarray(idate_lev, dtype=np.int32)
    assert ip.shape == (18,) and il.shape == (niv, 5, 18)
    keep = [_c(a) for a in (ps, t, w, h, u, v)]
    _check(L, L.rb_ecmwf_write_time(path.encode(), 1 if append else 0, nlon, nlat, niv,
                                    ip.ctypes.data, keep[0][1], il.ctypes.data,
                                    *[k[1] for k in keep[1:]]))


def read_ecmwf_time(path: str, itime: int, nlon: int = NLON, nlat: int = NLAT, niv: int = NIV) -> dict:
    L = _lib.load()
    idate = np.empty(18, dtype=np.int32)
    out = {"ps": np.empty((nlat, nlon))}
    for k in ("t", "w", "h", "u", "v"):
        out[k] = np.empty((niv, nlat, nlon))
    _check(L, L.rb_ecmwf_read_time(path.encode(), itime, nlon, nlat, niv, idate.ctypes.data,
                                   *[out[k].ctypes.data for k in ("ps", "t", "w", "h", "u", "v")]))
    out["idate"] = idate
    return out
