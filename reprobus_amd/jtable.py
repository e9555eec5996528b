"""Synthetic photolysis table in the layout REPROBUS reads from ``jstrato.txt``.

The model's own table (``src/data.zip``) is not distributed with the source, so
this module builds a physically shaped stand-in with the same shape, node grid
and text format, and both the reference build and this package read the same
file.

Layout (reference ``src/reprobus_l137_20230723.f``):
  * ``ajl(40, 101, 27, 7)`` -- 40 J-values x altitude node (1 km, z = 0..100 km)
    x SZA node (``dsza`` F:5035-5038) x O3-column ratio node (``xv3`` F:5039);
  * ``o3up(0:79)`` -- standard O3 column above z = 0..79 km (cm-2);
  * read order F:5953-5973: for l in 1..7, k in 1..27, j in 101..1: one
    list-directed line (the altitude index), then the 40 values in ``7e11.4``;
    then one skipped line and 80 lines ``z o3up``.

Table model: J_n(z, chi, r) = J0_n * exp(-min(30, tau_n(z) * m(chi) * r**w_n)),
tau_n(z) = tau20_n * exp(-(z - 20)/7): an absorber column decaying with a 7 km
scale height, ``m`` an air-mass factor that keeps growing past 90 deg (light
reaching the upper levels around the terminator), and ``w_n`` the weight of
O3 absorption in the species' action spectrum.  The cap keeps every entry
strictly positive, which the reference needs: it divides by J(BrCl) in daytime
(F:1525-1526, F:1570-1571).  Values are rounded to the 4 significant digits
the ``7e11.4`` format carries, so the text file is the exact table.
"""
from __future__ import annotations

import math
import os

import numpy as np

NPHOT_TAB = 40          # table J-values (J(NO) is computed by jno.f90)
NZ, NSZA, NOZO = 101, 27, 7

# SZA nodes (deg), F:5035-5038, and O3-ratio nodes, F:5039.
DSZA = np.array([0., 5., 10., 15., 20., 25., 30., 35., 40., 45., 50., 55., 60., 65., 70., 75.,
                 80., 82., 84., 86., 88., 90., 91., 92., 93., 94., 95.])
XV3 = np.array([0.4, 0.6, 0.8, 1.0, 1.2, 1.4, 1.6])

# (J0 overhead at the top [s-1], tau at 20 km overhead, O3 weight) per table slot,
# in the reference's numbering F:754-793.
_SPECIES = [
    ("o2",          1.0e-9,  40.0, 0.0),
    ("no2",         9.0e-3,  0.15, 0.2),
    ("o3_o1d",      5.0e-3,  6.0,  1.0),
    ("o3_o3p",      5.0e-4,  0.10, 0.3),
    ("no3_no",      2.0e-2,  0.08, 0.1),
    ("no3_no2",     1.7e-1,  0.08, 0.1),
    ("n2o5",        5.0e-5,  3.0,  0.9),
    ("hno3",        5.0e-5,  5.0,  0.9),
    ("hno4",        1.0e-4,  3.0,  0.9),
    ("h2o2",        2.0e-5,  3.0,  0.9),
    ("h2o",         1.0e-9,  35.0, 0.0),
    ("n2o",         1.0e-6,  12.0, 0.5),
    ("hocl",        3.0e-4,  0.6,  0.6),
    ("clono2_cl",   1.0e-4,  1.5,  0.8),
    ("clono2_clo",  3.0e-5,  1.5,  0.8),
    ("cl2o2",       1.0e-3,  0.8,  0.7),
    ("hcl",         1.0e-7,  25.0, 0.1),
    ("cl2",         2.5e-3,  0.12, 0.2),
    ("clno2",       3.0e-4,  0.6,  0.6),
    ("bro",         4.0e-2,  0.20, 0.4),
    ("brono2",      1.0e-3,  0.5,  0.5),
    ("hobr",        2.0e-3,  0.3,  0.4),
    ("oclo",        7.0e-2,  0.20, 0.4),
    ("brcl",        1.0e-2,  0.12, 0.2),
    ("ch3o2h",      2.0e-5,  3.0,  0.9),
    ("ch2o_hco",    5.0e-5,  1.0,  0.9),
    ("ch2o_co",     7.0e-5,  1.0,  0.9),
    ("cfc11",       1.0e-5,  10.0, 0.6),
    ("cfc12",       1.0e-6,  14.0, 0.5),
    ("cfc113",      2.0e-6,  12.0, 0.5),
    ("ccl4",        3.0e-5,  8.0,  0.7),
    ("ch3ccl3",     1.0e-5,  9.0,  0.7),
    ("ch3cl",       1.0e-6,  15.0, 0.4),
    ("hcfc22",      1.0e-8,  20.0, 0.2),
    ("ch3br",       5.0e-6,  10.0, 0.6),
    ("h1211",       2.0e-4,  4.0,  0.9),
    ("h1301",       5.0e-6,  9.0,  0.7),
    ("co2",         1.0e-10, 45.0, 0.0),
    ("ch2br2",      1.0e-5,  6.0,  0.8),
    ("hbr",         2.0e-6,  8.0,  0.5),
]
assert len(_SPECIES) == NPHOT_TAB
J_NAMES = [s[0] for s in _SPECIES] + ["no"]


def _airmass(chi_deg: np.ndarray) -> np.ndarray:
    """Kasten-Young air mass up to 90 deg, growing e-fold per 2 deg beyond."""
    c = np.minimum(chi_deg, 90.0)
    m = 1.0 / (np.cos(np.radians(c)) + 0.50572 * (96.07995 - c) ** -1.6364)
    return m * np.exp(np.maximum(chi_deg - 90.0, 0.0) / 2.0)


def _round4(x: np.ndarray) -> np.ndarray:
    """Round to the 4 significant digits of a Fortran ``e11.4`` field."""
    out = np.empty_like(x)
    flat_in, flat_out = x.ravel(), out.ravel()
    for i, v in enumerate(flat_in):
        flat_out[i] = float(_e11_4(v))
    return out


def _e11_4(v: float) -> str:
    """Format one value like Fortran ``E11.4``: ' 0.dddd E+ee' (11 chars)."""
    if v == 0.0:
        return "  0.0000E+00"[-11:]
    m, e = f"{v:.3e}".split("e")          # d.ddd, exponent
    digits = m.replace(".", "").replace("-", "")
    e = int(e) + 1                        # 0.dddd convention
    sign = "-" if v < 0 else " "
    s = f"{sign}0.{digits}E{e:+03d}"
    return s.rjust(11)


def build_ajl() -> np.ndarray:
    """Return ajl with shape (40, 101, 27, 7) (Fortran index order), rounded."""
    z = np.arange(NZ, dtype=np.float64)                    # node j <-> z = j-1 km
    m = _airmass(DSZA)
    ajl = np.empty((NPHOT_TAB, NZ, NSZA, NOZO))
    for n, (_, j0, tau20, w) in enumerate(_SPECIES):
        tau = tau20 * np.exp(-(z - 20.0) / 7.0)             # (NZ,)
        od = tau[:, None, None] * m[None, :, None] * (XV3[None, None, :] ** w)
        ajl[n] = j0 * np.exp(-np.minimum(od, 30.0))
    return _round4(ajl)


def build_o3up() -> np.ndarray:
    """Standard O3 column above z (cm-2), z = 0..79 km, from a smooth vmr profile."""
    zf = np.linspace(0.0, 120.0, 12001)
    p_pa = 101325.0 * np.exp(-zf / 7.0)
    q = 8.0e-6 * np.exp(-((zf - 32.0) / 9.0) ** 2) + 2.0e-8
    zcmt = 2.12e20                                           # molecules per (Pa * vmr)
    # column above zf[i] = integral over pressure from 0 to p(z)
    seg = 0.5 * (q[1:] + q[:-1]) * (p_pa[:-1] - p_pa[1:]) * zcmt
    above = np.concatenate([np.cumsum(seg[::-1])[::-1], [0.0]])
    col = np.interp(np.arange(80.0), zf, above)
    return np.array([float(f"{v:.6e}") for v in col])


def write_jstrato(path: str, ajl: np.ndarray, o3up: np.ndarray) -> None:
    """Write the table in the record order of F:5953-5973 (format F:5987)."""
    lines = []
    for l in range(NOZO):
        for k in range(NSZA):
            for j in range(NZ - 1, -1, -1):
                lines.append(f"{j + 1:4d}")
                vals = [_e11_4(v) for v in ajl[:, j, k, l]]
                for c in range(0, NPHOT_TAB, 7):
                    lines.append("".join(vals[c:c + 7]))
    lines.append("   z(km)  o3 column above z (cm-2)")
    for i in range(80):
        lines.append(f"{float(i):6.1f} {o3up[i]:.6e}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_jstrato(path: str) -> tuple[np.ndarray, np.ndarray]:
    """Parse a ``jstrato.txt`` written in the reference layout."""
    ajl = np.empty((NPHOT_TAB, NZ, NSZA, NOZO))
    with open(path) as f:
        rows = f.read().split("\n")
    pos = 0
    for l in range(NOZO):
        for k in range(NSZA):
            for j in range(NZ - 1, -1, -1):
                pos += 1                                    # altitude index line
                vals = []
                while len(vals) < NPHOT_TAB:
                    row = rows[pos]
                    pos += 1
                    vals += [float(row[c:c + 11]) for c in range(0, len(row.rstrip()), 11)]
                ajl[:, j, k, l] = vals
    pos += 1
    o3up = np.array([float(rows[pos + i].split()[1]) for i in range(80)])
    return ajl, o3up


_CACHE: dict = {}


def default_table() -> tuple[np.ndarray, np.ndarray]:
    """The project's synthetic table (ajl (40,101,27,7), o3up (80,)), cached."""
    if "t" not in _CACHE:
        _CACHE["t"] = (build_ajl(), build_o3up())
    return _CACHE["t"]


def ensure_file(path: str) -> str:
    if not os.path.exists(path):
        ajl, o3up = default_table()
        write_jstrato(path, ajl, o3up)
    return path


if __name__ == "__main__":  # pragma: no cover
    a, o = default_table()
    print(a.shape, a.min(), a.max(), o[:3], o[-3:])
    print(math.isfinite(float(a.sum())))
