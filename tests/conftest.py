import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
FLOOR = 1e-30          # mole-fraction floor of the parity metric
PARITY_TOL = 1e-9      # relative, per species, above FLOOR


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


def load_fixture(name):
    """(StepInputs, reference outputs) of a committed golden fixture."""
    from reprobus_amd import synth

    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    nlon, nlat, niv, nivbas, nivhau = (int(v) for v in z["grid"])
    g = synth.Grid(nlon, nlat, niv, nivbas, nivhau)
    inp = synth.StepInputs(g, float(z["dt"]), z["T"], z["p"], z["sza"], z["alt"], z["qj1_in"],
                           z["hc_in"], None, {"name": name})
    ref = {k: z[k] for k in ("qj1_chem", "hc_chem", "surfarea", "vsed3d", "parth2osed",
                             "parthno3sed", "qj1", "hc")}
    if "qj1_sedi" in z.files:      # the reference ran sedimentation before conserv
        ref["qj1_sedi"] = z["qj1_sedi"]
        ref["dtsedi"], ref["nivh2o"] = float(z["dtsedi"]), int(z["nivh2o"])
    return inp, ref


def rel_err(a, b, floor=FLOOR):
    """|a-b|/|b| where max(|a|,|b|) > floor, else 0 (per element)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    mask = np.maximum(np.abs(a), np.abs(b)) > floor
    d = np.abs(a - b) / np.maximum(np.abs(b), floor)
    d = np.where(mask, d, 0.0)
    d = np.where(np.isnan(a) != np.isnan(b), np.inf, d)
    return np.nan_to_num(d, nan=0.0)


def max_rel_by_species(a, b, floor=FLOOR):
    d = rel_err(a, b, floor)
    return d.reshape(d.shape[0], -1).max(axis=1)


@pytest.fixture(scope="session")
def table():
    from reprobus_amd import jtable

    return jtable.default_table()


def box_rel(a_q, a_h, b_q, b_h, floor=FLOOR):
    """Per box: the largest relative difference of any long- or short-lived species."""
    return np.maximum(rel_err(a_q[:42], b_q[:42], floor).max(axis=0),
                      rel_err(a_h, b_h, floor).max(axis=0))
