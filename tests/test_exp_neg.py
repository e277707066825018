"""The exponential used by the Gaussian, C0 and discontinuous integrands
(pagani::math::exp_bounded in include/pagani/genz.hpp, valid for arguments
up to 709; exp_neg is the same function for non-positive arguments).

CPU: its constants are recomputed from their definition (ln2 split into two
doubles; e^r interpolated at the 12 Chebyshev nodes of [-ln2/2, ln2/2] in
extended precision) and the polynomial's error bound is checked.
GPU: device results against the correctly rounded exponential (long double
on the host), with the device library's exp alongside.
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
HEADER = os.path.join(ROOT, "include", "pagani", "genz.hpp")
REFTESTS = os.path.join(HERE, "cxx", "libreftests.so")

LD = np.longdouble


def header_constants():
    text = open(HEADER).read()
    body = text[text.index("double exp_bounded(double x)"):]
    body = body[:body.index("__builtin_ldexp")]
    named = dict(re.findall(r"k(\w+) = (0x[0-9a-fp.+-]+);", body))
    coeffs = re.findall(r"(0x[0-9a-f.]+p[+-]\d+)", body[body.index("c[12]"):])
    return {k: float.fromhex(v) for k, v in named.items()}, [float.fromhex(c) for c in coeffs]


def chebyshev_coefficients():
    getcontext().prec = 60
    h = LD(str(Decimal(2).ln() / 2))
    pi = LD("3.14159265358979323846264338327950288")
    N = 12
    nodes = h * np.cos(pi * (np.arange(N, dtype=LD) + LD(0.5)) / LD(N))
    M = np.concatenate([np.array([[x ** k for k in range(N)] for x in nodes], dtype=LD),
                        np.exp(nodes)[:, None]], axis=1)
    for c in range(N):
        p = c + int(np.argmax(np.abs(M[c:, c])))
        M[[c, p]] = M[[p, c]]
        for r in range(c + 1, N):
            M[r, c:] -= (M[r, c] / M[c, c]) * M[c, c:]
    x = np.zeros(N, dtype=LD)
    for r in range(N - 1, -1, -1):
        x[r] = (M[r, N] - np.dot(M[r, r + 1:N], x[r + 1:N])) / M[r, r]
    return [float(v) for v in x], h


def test_split_ln2_constants():
    getcontext().prec = 60
    ln2 = Decimal(2).ln()
    named, _ = header_constants()
    assert named["Ln2Hi"] == float(ln2)
    assert named["Ln2Lo"] == float(ln2 - Decimal(float(ln2)))
    assert named["Log2e"] == float(1 / ln2)


def test_coefficients_are_the_chebyshev_interpolant():
    _, coeffs = header_constants()
    want, h = chebyshev_coefficients()
    assert len(coeffs) == 12
    for k, (got, w) in enumerate(zip(coeffs, want)):
        assert got == pytest.approx(w, rel=4e-16, abs=0), k
    # the polynomial with the header's (double) coefficients, in extended
    # precision, against e^r over the reduced range: well under one ulp
    r = np.linspace(-float(h), float(h), 100001).astype(LD)
    p = np.zeros_like(r)
    for c in reversed(coeffs):
        p = p * r + LD(c)
    rel = np.max(np.abs(p / np.exp(r) - 1))
    assert rel < 0.2 * 2.0 ** -53
    # a plausible slip (a Taylor coefficient instead of the fitted one) is
    # caught by the same bound
    bad = list(coeffs)
    bad[11] = 1 / math.factorial(11)
    q = np.zeros_like(r)
    for c in reversed(bad):
        q = q * r + LD(c)
    assert np.max(np.abs(q / np.exp(r) - 1)) > 0.2 * 2.0 ** -53


def _ulp_err(got, x):
    ref = np.exp(x.astype(LD))
    ref_d = ref.astype(np.float64)
    out = np.zeros(len(x))
    for i, (g, rd, rl) in enumerate(zip(got, ref_d, ref)):
        if rd == 0.0 or abs(rd) < 2.2250738585072014e-308:
            out[i] = abs(LD(g) - rl) / LD(5e-324)
        else:
            out[i] = abs(LD(g) - rl) / LD(math.ulp(rd))
    return out


@pytest.mark.gpu
def test_device_exp_neg_accuracy(pg, gpu):
    lib = ctypes.CDLL(REFTESTS)
    rng = np.random.default_rng(5)
    x = np.concatenate([
        -rng.uniform(0, 745.2, 200000),
        -np.exp(rng.uniform(-700, 6.6, 50000)),           # tiny to moderate magnitudes
        -(np.arange(1, 1100) * math.log(2) / 2),            # reduction boundaries
        np.array([0.0, -0.0, -5e-324, -1e-300, -708.39, -708.4, -745.13, -745.14, -800.0,
                  -1e4, -1e10]),
        # the positive side the discontinuous family uses (exponents up to sum a_i)
        rng.uniform(0, 709.7, 50000),
        np.arange(1, 1024) * math.log(2) / 2,
        np.array([1e-300, 5e-324, 1.0, 100.0, 709.0, 709.78]),
    ]).astype(np.float64)
    n = len(x)
    mine = np.zeros(n)
    ref = np.zeros(n)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.reftest_exp_neg(x.ctypes.data_as(dp), ctypes.c_size_t(n), mine.ctypes.data_as(dp),
                             ref.ctypes.data_as(dp))
    assert rc == 0
    e_mine = _ulp_err(mine, x)
    e_lib = _ulp_err(ref, x)
    assert e_mine.max() <= 1.0, (x[np.argmax(e_mine)], e_mine.max())
    # no worse on average than the device library's exp
    assert e_mine.mean() <= max(e_lib.mean() * 1.5, 0.3)
    assert mine[np.where(x == 0.0)[0]].tolist() == [1.0, 1.0]
    assert np.all(mine[x < -745.14] == 0.0)
