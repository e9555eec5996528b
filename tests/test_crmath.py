"""The product's correctly rounded elementary functions (reprobus_amd/csrc/crmath.h,
host build) against binary128 evaluation rounded to double (oracle/quadref.c).

The host and gfx950 builds run the same IEEE operation sequence (add, mul, div,
fma only), so agreement here is agreement of the device code.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT

HOST = os.path.join(ROOT, "reprobus_amd", "libcrmath_host.so")
QUAD = os.path.join(ROOT, "oracle", "libquadref.so")


@pytest.fixture(scope="module")
def libs():
    if not (os.path.exists(HOST) and os.path.exists(QUAD)):
        pytest.fail("build first: __graft_entry__.build()")
    H, Q = C.CDLL(HOST), C.CDLL(QUAD)
    for L, pre in ((H, "rb_cr_"), (Q, "qr_")):
        for n in ("exp", "log", "log10", "atan", "cos", "tanh", "sin", "asin", "acos"):
            f = getattr(L, pre + n)
            f.argtypes, f.restype = [C.c_double], C.c_double
        f = getattr(L, pre + "pow")
        f.argtypes, f.restype = [C.c_double, C.c_double], C.c_double
    H.rb_cr_pow10.argtypes, H.rb_cr_pow10.restype = [C.c_double], C.c_double
    return H, Q


CASES = [
    ("exp", lambda r, n: r.uniform(-700, 700, n)),
    ("exp", lambda r, n: r.uniform(-40, 40, n)),
    ("log", lambda r, n: np.exp(r.uniform(-700, 700, n))),
    ("log", lambda r, n: r.uniform(1e-9, 3e-6, n)),          # water / HNO3 partial pressures
    ("log10", lambda r, n: np.exp(r.uniform(-40, 40, n))),
    ("atan", lambda r, n: r.standard_normal(n) * np.exp(r.uniform(-12, 12, n))),
    ("cos", lambda r, n: r.uniform(-7, 7, n)),
    ("tanh", lambda r, n: np.exp(r.uniform(-30, 4, n))),
    ("sin", lambda r, n: r.uniform(-7, 7, n)),
    ("sin", lambda r, n: r.standard_normal(n) * np.exp(r.uniform(-30, 0, n))),
    ("asin", lambda r, n: r.uniform(-1, 1, n)),
    ("asin", lambda r, n: np.sign(r.standard_normal(n)) * (1 - np.exp(r.uniform(-36, -1, n)))),
    ("acos", lambda r, n: r.uniform(-1, 1, n)),
    ("acos", lambda r, n: np.sign(r.standard_normal(n)) * (1 - np.exp(r.uniform(-36, -1, n)))),
]


@pytest.mark.parametrize("name,gen", CASES)
def test_correctly_rounded_unary(libs, name, gen):
    H, Q = libs
    fh, fq = getattr(H, "rb_cr_" + name), getattr(Q, "qr_" + name)
    x = gen(np.random.default_rng(hash(name) % 2**32), 20000)
    bad = [v for v in x if fh(float(v)) != fq(float(v))]
    assert not bad, (name, bad[:5])


def test_sincos_is_sin_and_cos(libs):
    """sincos_cr, which the back-trajectories use, gives sin_cr's and cos_cr's bits."""
    H, _ = libs
    H.rb_cr_sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    H.rb_cr_sincos.restype = None
    r = np.random.default_rng(5)
    xs = np.concatenate([r.uniform(-7, 7, 5000), [0.0, -0.0, 1e-300, 150.0, np.pi]])
    s, c = C.c_double(), C.c_double()
    for x in xs:
        H.rb_cr_sincos(float(x), C.byref(s), C.byref(c))
        assert s.value == H.rb_cr_sin(float(x)) and c.value == H.rb_cr_cos(float(x)), x
        assert np.signbit(s.value) == np.signbit(H.rb_cr_sin(float(x)))


def test_correctly_rounded_pow(libs):
    H, Q = libs
    r = np.random.default_rng(11)
    xs = np.concatenate([np.exp(r.uniform(-20, 20, 10000)), r.uniform(0.01, 1.0, 10000)])
    ys = np.concatenate([r.uniform(-3, 3, 10000), r.uniform(-0.5, 0.8, 10000)])
    bad = [(x, y) for x, y in zip(xs, ys) if H.rb_cr_pow(float(x), float(y)) != Q.qr_pow(float(x), float(y))]
    assert not bad, bad[:5]
    # the exponents the aerosol code uses
    for x, y in ((0.5e-4, 3.0), (6.5e-4, 3.0), (10e-4, 3.0), (10.0, -7.3), (2.2, 0.5), (5.0, 0.1)):
        assert H.rb_cr_pow(x, y) == Q.qr_pow(x, y)


def test_special_values(libs):
    H, _ = libs
    assert H.rb_cr_exp(0.0) == 1.0 and H.rb_cr_log(1.0) == 0.0
    assert np.isnan(H.rb_cr_log(-1.0)) and H.rb_cr_exp(-1e4) == 0.0 and np.isinf(H.rb_cr_exp(1e4))
    assert H.rb_cr_atan(float("inf")) == np.pi / 2 and H.rb_cr_tanh(-30.0) == -1.0
    assert H.rb_cr_pow(7.0, 0.0) == 1.0 and H.rb_cr_pow(1.0, 123.4) == 1.0


def test_pow10(libs):
    H, Q = libs
    ys = np.random.default_rng(5).uniform(-30, 30, 20000)
    bad = [y for y in ys if H.rb_cr_pow10(float(y)) != Q.qr_pow(10.0, float(y))]
    assert not bad, bad[:5]


def test_kernel_constants_are_correctly_rounded(libs):
    """The kernel hoists the reference's constant libm calls as literals."""
    import re

    _, Q = libs
    src = open(os.path.join(ROOT, "reprobus_amd", "csrc", "chem_kernels.hip")).read()
    vals = dict(re.findall(r"(LOG760|LOG10|RNATS3|RNATL3|RICE3) = ([0-9.e+-]+)", src))
    assert float(vals["LOG760"]) == Q.qr_log(760.0)
    assert float(vals["LOG10"]) == Q.qr_log(10.0)
    assert float(vals["RNATS3"]) == Q.qr_pow(0.5e-4, 3.0)
    assert float(vals["RNATL3"]) == Q.qr_pow(6.5e-4, 3.0)
    assert float(vals["RICE3"]) == Q.qr_pow(10.0e-4, 3.0)
