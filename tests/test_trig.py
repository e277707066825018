"""The bounded-argument sine and cosine of the cosine integrands' factored
sums (pagani::math::sincos_bounded in include/pagani/trig.hpp).

CPU: its constants are recomputed from their definition (pi from Machin's
formula in 60-digit decimal arithmetic, pi/2 split into two doubles; the
polynomials interpolate the sine and cosine series at the 6 Chebyshev nodes
of [0, (pi/4)^2]) and the polynomials' error bound is checked in extended
precision.
GPU: device results against extended-precision sine and cosine over the
accepted range, with the device library's sincos alongside.
"""
import ctypes
import math
import os
import re
from decimal import Decimal, getcontext

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HEADER = os.path.join(ROOT, "include", "pagani", "trig.hpp")
REFTESTS = os.path.join(HERE, "cxx", "libreftests.so")

LD = np.longdouble
EPS = 2.0 ** -53


def header_constants():
    text = open(HEADER).read()
    body = text[text.index("void sincos_bounded(double x"):]
    body = body[:body.index("__builtin_rint")]
    named = {k: float.fromhex(v) for k, v in re.findall(r"k(\w+) = (-?0x[0-9a-fp.+-]+);", body)}
    arrays = {}
    for name in ("ks", "kc"):
        block = body[body.index(name + "[6]"):]
        block = block[:block.index("};")]
        arrays[name] = [float.fromhex(v) for v in re.findall(r"(-?0x[0-9a-f.]+p[+-]\d+)", block)]
    return named, arrays["ks"], arrays["kc"]


def machin_pi():
    getcontext().prec = 60

    def atan_inv(x):
        x = Decimal(x)
        s, p, k = Decimal(0), 1 / x, 0
        while True:
            t = p / (2 * k + 1)
            if t < Decimal(10) ** -70:
                return s
            s += t if k % 2 == 0 else -t
            p /= x * x
            k += 1

    return 16 * atan_inv(5) - 4 * atan_inv(239)


def series_s(t):
    # (sin(sqrt t) - sqrt t) / t^(3/2)
    return sum((-1) ** k * t ** (k - 1) / Decimal(math.factorial(2 * k + 1)) for k in range(1, 30))


def series_c(t):
    # (cos(sqrt t) - 1 + t/2) / t^2
    return sum((-1) ** k * t ** (k - 2) / Decimal(math.factorial(2 * k)) for k in range(2, 30))


def interpolant(fn, T, N=6):
    nodes = [T / 2 * (1 + Decimal(math.cos(math.pi * (i + 0.5) / N))) for i in range(N)]
    A = [[x ** k for k in range(N)] + [fn(x)] for x in nodes]
    for c in range(N):
        p = max(range(c, N), key=lambda r: abs(A[r][c]))
        A[c], A[p] = A[p], A[c]
        for r in range(c + 1, N):
            f = A[r][c] / A[c][c]
            A[r] = [a - f * b for a, b in zip(A[r], A[c])]
    x = [Decimal(0)] * N
    for r in range(N - 1, -1, -1):
        x[r] = (A[r][N] - sum(A[r][k] * x[k] for k in range(r + 1, N))) / A[r][r]
    return [float(v) for v in x]


def test_reduction_constants():
    pi = machin_pi()
    # Machin's formula against the digits of pi
    assert str(pi).startswith("3.14159265358979323846264338327950288419716939937510")
    named, _, _ = header_constants()
    assert named["TwoOverPi"] == float(2 / pi)
    assert named["Pio2Hi"] == float(pi / 2)
    assert named["Pio2Lo"] == float(pi / 2 - Decimal(named["Pio2Hi"]))
    # x - n Pio2Hi is exact for |n| < 2^20: Pio2Hi is a multiple of 2^-52,
    # so the difference is a multiple of min(ulp(x), 2^-52) below pi/4 + 1
    assert (named["Pio2Hi"] * 2.0 ** 52).is_integer()


def test_coefficients_are_the_chebyshev_interpolants():
    pi = machin_pi()
    T = (pi / 4) ** 2
    _, ks, kc = header_constants()
    want_s = interpolant(series_s, T)
    want_c = interpolant(series_c, T)
    assert len(ks) == 6 and len(kc) == 6
    for k in range(6):
        assert ks[k] == pytest.approx(want_s[k], rel=4e-16, abs=0), k
        assert kc[k] == pytest.approx(want_c[k], rel=4e-16, abs=0), k
    # the polynomials with the header's (double) coefficients, in extended
    # precision, against sin and cos on [0, pi/4]
    r = np.linspace(0, float(pi / 4), 100001).astype(LD)
    t = r * r
    ps = np.zeros_like(t)
    pc = np.zeros_like(t)
    for c in reversed(ks):
        ps = ps * t + LD(c)
    for c in reversed(kc):
        pc = pc * t + LD(c)
    sr = r + r * t * ps
    cr = 1 - t / 2 + t * t * pc
    assert np.max(np.abs(sr[1:] / np.sin(r[1:]) - 1)) < 0.25 * EPS
    assert np.max(np.abs(cr - np.cos(r))) < 0.05 * EPS
    # a plausible slip (a Taylor coefficient instead of the fitted one) fails
    bad = list(ks)
    bad[5] = 1 / math.factorial(13)
    q = np.zeros_like(t)
    for c in reversed(bad):
        q = q * t + LD(c)
    assert np.max(np.abs((r + r * t * q)[1:] / np.sin(r[1:]) - 1)) > 0.25 * EPS


@pytest.mark.gpu
def test_device_sincos_accuracy(pg, gpu):
    lib = ctypes.CDLL(REFTESTS)
    rng = np.random.default_rng(9)
    k = np.arange(-2000, 2001)
    x = np.concatenate([
        rng.uniform(-10, 10, 200000),
        rng.uniform(-1e3, 1e3, 100000),
        rng.uniform(-2.0 ** 20, 2.0 ** 20, 50000),
        np.exp(rng.uniform(-700, 0, 20000)) * rng.choice([-1, 1], 20000),
        k * (math.pi / 4),                      # octant boundaries
        k * (math.pi / 2) + 1e-9,               # near the zeros of sin or cos
        np.array([0.0, -0.0, 5e-324, 1.0, -1.0, 2.0 ** 20, -2.0 ** 20]),
    ]).astype(np.float64)
    n = len(x)
    s, c, ls, lc = (np.zeros(n) for _ in range(4))
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.reftest_sincos(x.ctypes.data_as(dp), ctypes.c_size_t(n), s.ctypes.data_as(dp),
                            c.ctypes.data_as(dp), ls.ctypes.data_as(dp), lc.ctypes.data_as(dp))
    assert rc == 0
    xl = x.astype(LD)
    rs, rcos = np.sin(xl), np.cos(xl)
    # absolute error in units of 2^-53 (the values multiply into sums whose
    # scale is at least the largest point value)
    es = np.abs(s.astype(LD) - rs) / LD(EPS)
    ec = np.abs(c.astype(LD) - rcos) / LD(EPS)
    assert es.max() <= 2.0, (x[np.argmax(es)], es.max())
    assert ec.max() <= 2.0, (x[np.argmax(ec)], ec.max())
    # no worse on average than the device library
    els = np.abs(ls.astype(LD) - rs) / LD(EPS)
    elc = np.abs(lc.astype(LD) - rcos) / LD(EPS)
    assert es.mean() <= max(1.5 * els.mean(), 0.3) and ec.mean() <= max(1.5 * elc.mean(), 0.3)
    # small arguments keep their relative accuracy
    tiny = (np.abs(x) < 1e-3) & (x != 0)
    rel = np.abs(s[tiny].astype(LD) / rs[tiny] - 1)
    assert rel.max() <= 2 * EPS
    assert np.all(s[x == 0.0] == 0.0) and np.all(c[x == 0.0] == 1.0)
