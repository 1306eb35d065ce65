"""The polynomial forms of the electron heat-capacity functions compiled into
the dTe kernels (spilady_amd/csrc/heatcap_poly.h), checked on the CPU against
the defining formulas of heatcapacity_CPU.cpp:34-57 evaluated in long double:
C(T) = 3 tanh(bT), E(T) = 3/b log cosh(bT), T(E) = acosh(exp(bE/3))/b."""
import os
import re

import numpy as np
import pytest

HDR = os.path.join(os.path.dirname(__file__), "..", "spilady_amd", "csrc", "heatcap_poly.h")


def _consts():
    s = open(HDR).read()
    out = {}
    for name, body in re.findall(r"constexpr double (\w+)\[\d+\] = \{(.*?)\};", s, re.S):
        out[name] = [float(v) for v in body.replace("\n", " ").split(",")]
    for name, val in re.findall(r"constexpr double (\w+) = ([0-9.eE+-]+);", s):
        out[name] = float(val)
    return out


def _horner(c, t):
    acc = np.full_like(t, c[-1])
    for k in range(len(c) - 2, -1, -1):
        acc = acc * t + c[k]
    return acc


@pytest.fixture(scope="module")
def K():
    return _consts()


def test_header_has_the_three_fits(K):
    assert K["kPolyXMax"] == 1.25 and K["kPolyYMax"] == 0.75
    # the leading terms are fixed by the series: log cosh x = x^2/2 - x^4/12 ...,
    # tanh x = x - x^3/3 ..., acosh(e^y) = sqrt(2y) (1 + y/6 ...)
    assert K["kLogCoshQ"][0] == pytest.approx(0.5, rel=1e-14)
    assert K["kLogCoshQ"][1] == pytest.approx(-1 / 12, rel=1e-10)
    assert K["kTanhR"][0] == pytest.approx(1.0, rel=1e-14)
    assert K["kTanhR"][1] == pytest.approx(-1 / 3, rel=1e-10)
    assert K["kAcoshExpG"][0] == pytest.approx(1.0, rel=1e-14)
    assert K["kAcoshExpG"][1] == pytest.approx(1 / 6, rel=1e-10)


def test_fits_match_the_formulas_over_their_range(K):
    ld = np.longdouble
    x = np.linspace(1e-6, K["kPolyXMax"], 20001)
    xl = x.astype(ld)
    logcosh = np.log1p(2 * np.sinh(xl / 2) ** 2)   # cosh x - 1 = 2 sinh^2(x/2), no cancellation
    got = x * x * _horner(K["kLogCoshQ"], x * x)
    assert np.max(np.abs((got - logcosh) / logcosh)) < 1e-15
    got = x * _horner(K["kTanhR"], x * x)
    ref = np.tanh(xl)
    assert np.max(np.abs((got - ref) / ref)) < 1e-15
    y = np.linspace(1e-9, K["kPolyYMax"], 20001)
    yl = y.astype(ld)
    ref = yl + np.log1p(np.sqrt(-np.expm1(-2 * yl)))
    got = np.sqrt(2 * y) * _horner(K["kAcoshExpG"], y)
    assert np.max(np.abs((got - ref) / ref)) < 1e-15


def test_inverse_pair_round_trips(K):
    # T(E(T)) = T through the two fits, where both apply (T up to 6,250 K)
    b = 2e-4 / 8.617343e-5
    t = np.linspace(1e-3, 6000.0, 5001) * 8.617343e-5
    x = b * t
    e = 3.0 / b * x * x * _horner(K["kLogCoshQ"], x * x)
    y = b / 3.0 * e
    back = np.sqrt(2 * y) * _horner(K["kAcoshExpG"], y) * (1.0 / b)
    assert np.max(np.abs(back / t - 1)) < 5e-15
