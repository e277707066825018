"""PAGANI above 8 dimensions: the reference instantiates Workspace<double, 10>
(cuda/pagani/demos/oscillatory_integrands.cu:23, 91-93) and has 9D demos;
its default initial split is 2 per axis up to 10D and 1 above
(Workspace.cuh:379-387, new_time_and_call.cuh:116-126).

- per-region parity at d = 9, 10 for the six Genz families (precompiled
  descriptors) and for plain functors through the header, against the
  oracle at the kernel's own points (the bound of test_gpu_eval.py);
- the reference's 10D oscillatory demo, cos(sum x_i) over [pi, 1.1 pi]^10 at
  epsrel 1e-3 through Workspace<double, 10> with a Sub_regions list of 2 per
  axis: the oracle's per-iteration region counts and the closed form
  Re prod_j (e^{i b} - e^{i a}) / i;
- Genz descriptors at 9D and 10D through the C ABI against the oracle's
  loop and the closed forms; an 11D descriptor is refused.
"""
import ctypes
import math
import os

import numpy as np
import pytest

import oracle_api as O
from test_gpu_eval import (FAMILIES, check_region_parity, genz_params, gpu_eval,
                           random_regions)

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("d", [9, 10])
@pytest.mark.parametrize("slices", [0, 1])
def test_eval_genz_high_dims(pg, gpu, family, d, slices):
    n = 500
    a, u = genz_params(family, d, seed=900 + d)
    L, H = random_regions(d, n, seed=31 * d)
    with pg.Workspace(d, 0) as ws:
        if slices:
            ws.set_slices(slices)
        ge, gerr, gdim = gpu_eval(ws, gpu, family, a, u, L, H, n)
    oe, oerr, odim, fd, rs = O.eval_regions_kform(d, family, O.params(a, u), L, H, n)
    check_region_parity(d, family, ge, gerr, gdim, oe, oerr, odim, fd, rs, floor=True)


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("d", [9, 10])
def test_eval_plain_functor_high_dims(pg, gpu, family, d):
    lib = ctypes.CDLL(os.path.join(HERE, "cxx", "libreftests.so"))
    n = 400
    a, u = genz_params(family, d, seed=950 + d)
    a = np.ascontiguousarray(a, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    L, H = random_regions(d, n, seed=37 * d)
    est, err = np.zeros(n), np.zeros(n)
    dim = np.zeros(n, dtype=np.int32)
    lib.reftest_sample_plain_genz.argtypes = [
        ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_void_p]
    rc = lib.reftest_sample_plain_genz(pg.FAMILY[family], d, a.ctypes.data, u.ctypes.data, n,
                                       L.ctypes.data, H.ctypes.data, 0, est.ctypes.data,
                                       err.ctypes.data, dim.ctypes.data)
    assert rc == 0
    oe, oerr, odim, fd, rs = O.eval_regions_kform(d, family, O.params(a, u), L, H, n,
                                                  plain=True)
    check_region_parity(d, family, est, err, dim, oe, oerr, odim, fd, rs, floor=True)


class _Res(ctypes.Structure):
    _fields_ = [("estimate", ctypes.c_double), ("errorest", ctypes.c_double),
                ("neval", ctypes.c_size_t), ("nregions", ctypes.c_size_t),
                ("nFinishedRegions", ctypes.c_size_t), ("status", ctypes.c_int),
                ("lastPhase", ctypes.c_int), ("chi_sq", ctypes.c_double),
                ("iters", ctypes.c_size_t)]


def cos_sum_box(d, lo, hi):
    """Integral of cos(sum x_i) over [lo, hi]^d: Re prod_j (e^{i hi} - e^{i lo}) / i."""
    z = 1.0 + 0.0j
    for _ in range(d):
        z *= (complex(math.cos(hi), math.sin(hi)) - complex(math.cos(lo), math.sin(lo))) / 1j
    return z.real


def test_oscillatory_10d_demo_through_header(gpu):
    d, lo, hi, eps = 10, math.pi, 1.1 * math.pi, 1e-3
    lib = ctypes.CDLL(os.path.join(HERE, "cxx", "libreftests.so"))
    lib.reftest_oscillatory10.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.POINTER(_Res),
                                          ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t)]
    cap = 800
    tn = (ctypes.c_uint64 * cap)()
    tm = (ctypes.c_uint64 * cap)()
    nt = ctypes.c_size_t(0)
    r = _Res()
    assert lib.reftest_oscillatory10(lo, hi, eps, ctypes.byref(r), tn, tm, cap,
                                     ctypes.byref(nt)) == 0
    exact = cos_sum_box(d, lo, hi)
    # the integrand changes sign across the box (the sum runs from 10 pi to
    # 11 pi), so the integral is small against |f| and the run may stop on
    # the memory model (status 1) before 1e-3; its estimate stays within
    # its error estimate
    print(f"10D demo: status {r.status}, estimate {r.estimate!r}, exact {exact!r}, "
          f"errorest {r.errorest:.3g}, nregions {r.nregions}")
    assert abs(r.estimate - exact) <= max(r.errorest, eps * abs(exact))
    # the oracle's first iterations (its whole 10D run takes the host
    # minutes): cos(2 pi u_0 + sum a_j x_j) with a = 1, u_0 = 0, same box,
    # the Sub_regions overload's bookkeeping from 2 per axis
    prm = O.params([1.0] * d, [0.0] * d)
    k_or = 7
    o, otr = O.integrate(d, "oscillatory", prm, eps, 1e-40, lows=np.full(d, lo),
                         highs=np.full(d, hi), init_k=2, user_regions=True, trace_cap=800,
                         max_iters=k_or)
    k = nt.value
    assert k > k_or
    assert [int(tn[i]) for i in range(k_or)] == [int(v) for v in otr["n"]]
    assert [int(tm[i]) for i in range(k_or)] == [int(v) for v in otr["m"]]


@pytest.mark.parametrize("family,d,eps", [("product_peak", 10, 1e-1), ("corner_peak", 9, 1e-2),
                                          ("oscillatory", 9, 1e-5), ("oscillatory", 10, 1e-4)])
def test_genz_descriptors_high_dims(pg, gpu, family, d, eps):
    a, u = genz_params(family, d, seed=1000 + d)
    prm = O.params(a, u)
    exact = O.exact(family, d, prm)
    with pg.Workspace(d, 0) as ws:
        ws.set_device_mem(int(16e9))
        r = ws.integrate_genz(family, a, u, epsrel=eps, epsabs=1e-20)
        tr = ws.trace()
    o, otr = O.integrate(d, family, prm, eps, 1e-20, trace_cap=800)
    assert list(tr["n"]) == [int(v) for v in otr["n"]]
    assert (r.nregions, r.iters, r.status) == (o.nregions, o.iters, o.status)
    assert abs(r.estimate - o.estimate) <= r.errorest + o.errorest
    if r.status == 0:
        assert abs(r.estimate - exact) <= max(r.errorest, eps * abs(exact))


def test_genz_descriptor_above_10d_refused(pg, gpu):
    with pg.Workspace(11, 0) as ws:
        with pytest.raises(pg.PaganiError) as e:
            ws.integrate_genz("gaussian", [10.0] * 11, [0.5] * 11, epsrel=1e-3)
        assert e.value.code == pg.EUNSUPPORTED
