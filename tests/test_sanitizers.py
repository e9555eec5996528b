"""Host-side AddressSanitizer and UndefinedBehaviorSanitizer runs (oracle/Makefile target
`sanitize`): the record reader/writer (reprobus_amd/csrc/records.cpp) through round trips
and error paths, and the restatement through the chemistry step, conserv, the converged
mode (plain and Anderson), the back-trajectories and internew -- each run must finish
clean, and the sanitized restatement must return the ordinary build's bits."""
import os
import subprocess

import numpy as np
import pytest

import cport
from conftest import ROOT
from reprobus_amd import synth

SAN = os.path.join(ROOT, "oracle", "_ref")


def _binary(name):
    path = os.path.join(SAN, name)
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    return path


def test_records_under_sanitizers(tmp_path):
    r = subprocess.run([_binary("sanitize_records"), str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed checks" in r.stdout


def test_restatement_under_sanitizers(tmp_path, table):
    inp = synth.make_c2(3, nlon=6, nlat=5)
    g = inp.grid
    ti = synth.make_trajectory_inputs(1, nlon=12, nlat=7, niv=6)
    nrapp = 2
    tq = synth.make_transport(2, nlon=12, nlat=7, niv=6).q[0]
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    with open(fin, "wb") as f:
        np.array([g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau, 5], dtype=np.int32).tofile(f)
        np.array([inp.dt]).tofile(f)
        for a in (inp.T, inp.p, inp.sza, inp.alt, inp.qj1, inp.hc, cport.fortran_ajl(table[0]),
                  table[1]):
            c(a).tofile(f)
        np.array([ti.nlon, ti.nlat, ti.niv, nrapp, ti.ntime, ti.nintecmwf],
                 dtype=np.int32).tofile(f)
        np.array([ti.daynum, ti.dtdyn]).tofile(f)
        c(ti.aaa).tofile(f)
        c(ti.bbb).tofile(f)
        for r in (ti.m1, ti.p1):
            for a in (r.u, r.v, r.w, r.t, r.lnps):
                c(a).tofile(f)
            np.array([r.day]).tofile(f)
        c(tq).tofile(f)
    r = subprocess.run([_binary("sanitize_restatement"), fin, fout], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    field, fb = g.nlon * g.nlat * g.niv, g.nlon * g.nlat * g.nivbas
    tn = ti.nlon * ti.nlat * ti.niv
    with open(fout, "rb") as f:
        rd = lambda n, dt=np.float64: np.fromfile(f, dtype=dt, count=n)
        step = {"qj1_chem": rd(43 * field), "hc_chem": rd(15 * field), "qj1": rd(43 * field),
                "hc": rd(15 * field)}
        conv = [{"q": rd(43 * field), "h": rd(15 * field), "it": rd(fb, np.int32)} for _ in range(2)]
        bt = {k: rd(tn) for k in ("x0", "y0", "z0", "p0")}
        bt["idx"] = rd(3 * tn, np.int32)
        bt["w"] = rd(12 * tn)
        interp = rd(tn)
    ref = cport.chem_step(inp, *table)
    for k in step:
        assert np.array_equal(step[k], ref[k].ravel()), k
    for j, accel in enumerate((0, 2)):
        cr = cport.chem_converge(inp, *table, tol=1e-11, maxiter=2000, conserv=False,
                                 accel=accel, aa_start=10)
        assert np.array_equal(conv[j]["q"], cr["qj1_chem"].ravel()), accel
        assert np.array_equal(conv[j]["h"], cr["hc_chem"].ravel()), accel
        assert np.array_equal(conv[j]["it"], cr["iters"].ravel()), accel
    br = cport.backtra(ti, nrapp)
    for k in ("x0", "y0", "z0", "p0"):
        assert np.array_equal(bt[k], br[k]), k
    assert np.array_equal(bt["idx"], br["idx"].ravel())
    assert np.array_equal(bt["w"], br["w"].ravel())
    tr = synth.make_transport(2, nlon=12, nlat=7, niv=6)
    out, _ = cport.internew(tr, br["idx"], br["w"], tq)
    assert np.array_equal(interp, out.ravel())
