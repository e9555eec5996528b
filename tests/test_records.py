"""The model's on-disk records (reprobus_amd.records over include/rbio.h): the ECMWF
forcing reader against the reference's own reader (golden fixture from the flang build
of F:4713-4935), and the history record against the Fortran runtime's own bytes for the
model's write statement (F:587-588) and the layout MODEL_history_convert_to_netcdf.py
reads (:113-129)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

import forcing_synth as fs


def test_ecmwf_reader_matches_reference(tmp_path):
    from reprobus_amd import records

    z = np.load(os.path.join(GOLDEN, "ecmwf_seed13.npz"), allow_pickle=False)
    y, m, d = fs.DATE
    path = str(tmp_path / f"ecmwf_{y:04d}{m:02d}{d:02d}")
    fs.write_file(path, int(z["seed"]), 3)
    idx = z["idx"].astype(np.int64) - 1
    # the model at 03 UT loads the 3 h and 6 h analyses (times 1 and 2)
    for key, itime in (("m1", 1), ("p1", 2)):
        r = records.read_ecmwf_time(path, itime)
        assert tuple(r["idate"][9:13]) == (y, m, d, 3 * itime)
        got = np.stack([r["ps"].ravel()[idx % (180 * 91)]] +
                       [r[f].ravel()[idx] for f in ("t", "w", "h", "u", "v")])
        assert np.array_equal(got, z[key]), key
        sums = np.array([r[f].sum() for f in ("ps", "t", "w", "h", "u", "v")])
        ref = z["sums"][0 if key == "m1" else 1]
        assert np.all(np.abs(sums - ref) <= 1e-12 * np.abs(ref).max()), key
        # the floor is exercised (the synthetic humidity has zeros and negatives)
        assert r["h"].min() == 1e-30 * 18.0 / 28.97


def test_ecmwf_round_trip_and_errors(tmp_path):
    from reprobus_amd import records
    from reprobus_amd._lib import RB_EIO, RbError

    path = str(tmp_path / "ecmwf_small")
    fs.write_file(path, 5, 2, nlon=7, nlat=4, niv=3)
    for it in (0, 1):
        r = records.read_ecmwf_time(path, it, 7, 4, 3)
        ref = fs.fields(5, it, 7, 4, 3)
        for k in ("ps", "t", "w", "u", "v"):
            assert np.array_equal(r[k], ref[k]), k
        assert np.array_equal(r["h"], np.maximum(ref["h"], 1e-30 * 18.0 / 28.97))
        assert np.array_equal(r["idate"], ref["idate_ps"])
    with pytest.raises(RbError) as e:
        records.read_ecmwf_time(path, 2, 7, 4, 3)              # past the last time
    assert e.value.code == RB_EIO
    with pytest.raises(RbError):
        records.read_ecmwf_time(path, 0, 8, 4, 3)              # wrong grid: record size


def _small_history(nlon, nlat, niv):
    """The integer-valued state oracle/hist_writer.f90 writes."""
    from reprobus_amd.records import History

    i = np.arange(1, nlon + 1)[None, None, :]
    j = np.arange(1, nlat + 1)[None, :, None]
    k = np.arange(1, niv + 1)[:, None, None]
    u = (1 * i + 7 * j + 31 * k).astype(np.float64)
    n43 = np.arange(1, 44)[:, None, None, None]
    n15 = np.arange(1, 16)[:, None, None, None]
    return History("001442", (2023, 1, 15, 12, 0), (1000 + i + 7 * j)[0].astype(np.float64),
                   u, -u, u + 0.5, u + 0.25, u[None] + 1000 * n43, u * 2, u[None] - 1000 * n15)


def test_history_matches_fortran_runtime_bytes(tmp_path):
    from reprobus_amd import records

    exe = os.path.join(ROOT, "oracle", "hist_writer")
    if not os.path.exists(exe):
        pytest.fail("build first: make -C oracle restatement")
    nlon, nlat, niv = 6, 4, 5
    fort = str(tmp_path / "fortran_history")
    subprocess.run([exe, str(nlon), str(nlat), str(niv), fort], check=True)
    ours = str(tmp_path / "our_history")
    h = _small_history(nlon, nlat, niv)
    records.write_history(ours, h)
    assert open(ours, "rb").read() == open(fort, "rb").read()
    back = records.read_history(fort, nlon, nlat, niv)
    assert back.namexp == "001442" and back.date == (2023, 1, 15, 12, 0)
    for f in records.History.FIELDS:
        assert np.array_equal(getattr(back, f), getattr(h, f)), f


def test_history_layout_is_what_the_converter_reads(tmp_path):
    """MODEL_history_convert_to_netcdf.py: 10 bytes (record length + 6-character name),
    five big-endian int32, then big-endian float64 arrays in the write order."""
    from reprobus_amd import records

    nlon, nlat, niv = 5, 3, 4
    h = _small_history(nlon, nlat, niv)
    path = str(tmp_path / "MODEL_history")
    records.write_history(path, h)
    raw = open(path, "rb").read()
    plane, field = nlon * nlat, nlon * nlat * niv
    nbytes = 6 + 20 + 8 * (plane + field * (4 + 43 + 1 + 15))
    assert len(raw) == nbytes + 8
    assert int.from_bytes(raw[:4], "big") == nbytes == int.from_bytes(raw[-4:], "big")
    assert raw[4:10] == b"001442"
    assert tuple(np.frombuffer(raw[10:30], dtype=">i4")) == (2023, 1, 15, 12, 0)
    data = np.frombuffer(raw[30:-4], dtype=">f8")
    o = 0
    for f, n in (("pj1", plane), ("uj1", field), ("vj1", field), ("alt", field), ("tj1", field),
                 ("qj1", 43 * field), ("surfarea", field), ("hc", 15 * field)):
        assert np.array_equal(data[o:o + n], getattr(h, f).ravel()), f
        o += n
    assert o == data.size


def test_history_errors(tmp_path):
    from reprobus_amd import records
    from reprobus_amd._lib import RB_EIO, RbError

    path = str(tmp_path / "h")
    records.write_history(path, _small_history(3, 2, 2))
    with pytest.raises(RbError) as e:
        records.read_history(path, 3, 2, 3)                     # wrong grid
    assert e.value.code == RB_EIO
    with open(path, "r+b") as f:
        f.truncate(100)
    with pytest.raises(RbError):
        records.read_history(path, 3, 2, 2)                     # short file
    with pytest.raises(RbError):
        records.read_history(str(tmp_path / "missing"), 3, 2, 2)
