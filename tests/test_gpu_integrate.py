"""Full PAGANI integrations on the GPU.

Known answers from the reference's tests (Easy_Integrals.cu:54-148,
BoxIntegrals.cu:16-66, accuracy_improves_with_epsrel.cu:38-67,
NanIntegral.cu:16-37, test_integration_from_c.c:16-21), run through both
boundaries: the precompiled Genz descriptors of the C ABI and the C++ drop-in
header (tests/cxx/reftests.hip).  The oracle's loop (same memory model, 16e9
bytes as in the reference) must agree: estimates within the combined error
estimates, and the same per-iteration region counts.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REFTESTS = os.path.join(HERE, "cxx", "libreftests.so")
C_EXE = os.path.join(HERE, "cxx", "test_integration_from_c")

# name, family, d, a, u, true value (reference constants)
GENZ_KATS = [
    ("F_2_6D", "product_peak", 6, [50.0] * 6, [0.5] * 6, 1.286889807581113e13),
    ("F_3_8D", "corner_peak", 8, [3, 2, 1, 4, 5, 6, 7, 8], None, 2.2751965817917756076e-10),
    ("F_4_5D", "gaussian", 5, [25.0] * 5, [0.5] * 5, 1.79132603674879e-06),
    ("F_5_8D", "c0", 8, [10.0] * 8, [0.5] * 8, 2.425217625641885e-06),
    ("F_6_6D", "discontinuous", 6, [10, 9, 8, 7, 6, 5], [.9, .8, .7, .6, .5, .4],
     1.5477367885091207413e8),
]


class _Res(ctypes.Structure):
    _fields_ = [("estimate", ctypes.c_double), ("errorest", ctypes.c_double),
                ("neval", ctypes.c_size_t), ("nregions", ctypes.c_size_t),
                ("nFinishedRegions", ctypes.c_size_t), ("status", ctypes.c_int),
                ("lastPhase", ctypes.c_int), ("chi_sq", ctypes.c_double),
                ("iters", ctypes.c_size_t)]


@pytest.fixture(scope="module")
def reflib():
    lib = ctypes.CDLL(REFTESTS)
    lib.reftest_integrate.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(_Res)]
    lib.reftest_integrate_split.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_size_t,
                                            ctypes.POINTER(_Res),
                                            ctypes.POINTER(ctypes.c_size_t)]
    return lib


def ref_integrate(lib, name, epsrel, epsabs=1e-12):
    r = _Res()
    rc = lib.reftest_integrate(name.encode(), epsrel, epsabs, ctypes.byref(r))
    assert rc == 0, f"{name}: rc {rc}"
    return r


@pytest.mark.parametrize("kat", GENZ_KATS, ids=[k[0] for k in GENZ_KATS])
def test_genz_kat_c_abi(pg, gpu, kat):
    name, fam, d, a, u, true = kat
    with pg.Workspace(d, 0) as ws:
        ws.set_device_mem(int(16e9))
        r = ws.integrate_genz(fam, a, u, epsrel=1e-3, epsabs=1e-12)
        tr = ws.trace()
    assert r.status == 0
    assert abs(r.estimate - true) <= 1e-3 * abs(true)
    # the oracle's loop with the same memory model
    o, otr = O.integrate(d, fam, O.params(a, u), 1e-3, 1e-12, trace_cap=800)
    assert o.status == r.status
    assert abs(r.estimate - o.estimate) <= r.errorest + o.errorest
    assert list(tr["n"]) == list(otr["n"]), "per-iteration region counts differ"
    assert r.nregions == o.nregions and r.iters == o.iters


@pytest.mark.parametrize("kat", GENZ_KATS, ids=[k[0] for k in GENZ_KATS])
def test_genz_kat_cxx_header(gpu, reflib, kat):
    name, fam, d, a, u, true = kat
    r = ref_integrate(reflib, name, 1e-3)
    assert r.status == 0
    assert abs(r.estimate - true) <= 1e-3 * abs(true)


def test_constant_kats(gpu, reflib):
    for name, val in [("PTest", 15.37), ("NTest", -15.37)]:
        r = ref_integrate(reflib, name, 1e-3)
        assert r.status == 0 and abs(r.estimate - val) <= 1e-6 * abs(val)
    r = ref_integrate(reflib, "ZTest", 1e-3)
    assert r.estimate == 0.0


@pytest.mark.parametrize("name,true", [("BoxIntegral8_15", 8879.851175413485),
                                       ("BoxIntegral8_25", 14996089.096112404019),
                                       ("BoxIntegral8_22", 1495369.283757217694)])
def test_box_integrals(gpu, reflib, name, true):
    r = ref_integrate(reflib, name, 1e-3)
    assert abs(true - r.estimate) / true <= 1e-3


def test_nan_integral(gpu, reflib):
    r = ref_integrate(reflib, "NaN_Integral", 1e-3)
    assert np.isnan(r.estimate)


def test_accuracy_improves_with_epsrel(gpu, reflib):
    epsrel, prev = 1.0e-3, 1.0
    while epsrel > 1.0e-6:
        r = ref_integrate(reflib, "Fun6", epsrel, 1.0e-40)
        assert r.status == 0
        assert r.errorest / r.estimate <= epsrel
        assert r.errorest <= prev
        assert abs(r.estimate - 1.0) <= max(3 * r.errorest, 1e-12)
        prev = r.errorest
        epsrel /= 2.0


def test_sub_regions_overload(gpu, reflib):
    r = _Res()
    fin = ctypes.c_size_t(0)
    rc = reflib.reftest_integrate_split(b"F_4_5D", 3, 1e-3, 1e-12, int(16e9), ctypes.byref(r),
                                        ctypes.byref(fin))
    assert rc == 0 and r.status == 0
    assert abs(r.estimate - 1.79132603674879e-06) <= 1e-3 * 1.79132603674879e-06
    assert fin.value > 0
    o = O.integrate(5, "gaussian", O.params([25.0] * 5, [0.5] * 5), 1e-3, 1e-12, init_k=3,
                    user_regions=True)
    assert r.nregions == o.nregions and r.iters == o.iters
    assert abs(r.estimate - o.estimate) <= r.errorest + o.errorest


def test_c_interop_shim(pg, gpu):
    rc, val = pg.do_integration_from_c()
    assert rc == 0
    assert abs(val - 0.25) / 0.25 <= 1e-6


def test_c_interop_program(gpu):
    p = subprocess.run([C_EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout


def test_fp32_integrand_converges(pg, gpu):
    """Mixed precision on the 6D corner peak (one of the mixed-precision workloads)."""
    d = 6
    rng = np.random.default_rng(3)
    a = 5 + 10 * rng.random(d)
    with pg.Workspace(d, 0) as ws:
        r32 = ws.integrate_genz("corner_peak", a, None, epsrel=1e-4, precision=32)
        r64 = ws.integrate_genz("corner_peak", a, None, epsrel=1e-4, precision=64)
    exact = O.exact("corner_peak", d, O.params(a))
    assert r64.status == 0 and r32.status == 0
    assert abs(r64.estimate - exact) <= 1e-4 * abs(exact)
    assert abs(r32.estimate - exact) <= 3e-4 * abs(exact)
