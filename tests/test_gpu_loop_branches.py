"""Branches of the iteration loop and boundary semantics that the known-answer
integrals never reach, against the oracle's loop (same 16e9 memory model as
the reference):

- fix_error_budget_overflow (Workspace.cuh:121-146): cos(sum (i+1) x_i) at
  5D and 6D, whose integral is small against its integrand, finishes more
  error than the budget allows; the branch reactivates every region.  The
  oracle counts the iterations that take it (pinned on the CPU in
  test_oracle.py), the library counts its own, both loops (host-driven and
  device-resident) are run.
- predict_split (Workspace.cuh:206-210, 415-420): relerr classification off
  before iteration 15 while nothing finished, on after; runs that go past
  iteration 15, so the switch shows in the trajectory.
- the C++ header: the reference's integrands as drop-in functors, with the
  oracle's per-iteration region counts.

Bar: per-iteration region counts n and m, nregions, iters and status
identical to the oracle's; estimates within the combined error estimates.
"""
import ctypes
import os

import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REFTESTS = os.path.join(HERE, "cxx", "libreftests.so")
MEM = int(16e9)

BUDGET_CASES = [(5, 1e-5), (5, 1e-6), (6, 1e-4), (6, 1e-6)]


class _Res(ctypes.Structure):
    _fields_ = [("estimate", ctypes.c_double), ("errorest", ctypes.c_double),
                ("neval", ctypes.c_size_t), ("nregions", ctypes.c_size_t),
                ("nFinishedRegions", ctypes.c_size_t), ("status", ctypes.c_int),
                ("lastPhase", ctypes.c_int), ("chi_sq", ctypes.c_double),
                ("iters", ctypes.c_size_t)]


def _same_run(r, tr, o, otr):
    assert list(tr["n"]) == [int(v) for v in otr["n"]], "per-iteration n differs"
    assert list(tr["m"]) == [int(v) for v in otr["m"]], "per-iteration m differs"
    assert (r.nregions, r.iters, r.status) == (o.nregions, o.iters, o.status)
    assert abs(r.estimate - o.estimate) <= r.errorest + o.errorest


@pytest.mark.parametrize("d,eps", BUDGET_CASES)
def test_budget_fix_trajectory(pg, gpu, d, eps):
    a, u = [float(i + 1) for i in range(d)], [0.0] * d
    o, otr = O.integrate(d, "oscillatory", O.params(a, u), eps, 1e-20, trace_cap=800)
    assert o.budget_fixes > 0
    exact = O.exact("oscillatory", d, O.params(a, u))
    for device_loop in (False, True):
        with pg.Workspace(d, 0) as ws:
            ws.set_device_mem(MEM)
            ws.set_device_loop(device_loop)
            r = ws.integrate_genz("oscillatory", a, u, epsrel=eps, epsabs=1e-20)
            tr, st = ws.trace(), ws.stats()
        _same_run(r, tr, o, otr)
        assert st["budget_fixes"] == o.budget_fixes
        assert abs(r.estimate - exact) <= max(r.errorest, eps * abs(exact))


PREDICT_CASES = [("c0", 3, [10.0, 10.0, 10.0], [0.3, 0.5, 0.7], 1e-6),
                 ("discontinuous", 3, [5.0, 6.0, 7.0], [0.4, 0.5, 0.6], 1e-5),
                 ("gaussian", 4, [10.0] * 4, [0.4, 0.5, 0.6, 0.45], 1e-6)]


@pytest.mark.parametrize("case", PREDICT_CASES, ids=[c[0] for c in PREDICT_CASES])
def test_predict_split_relerr_rule(pg, gpu, case):
    fam, d, a, u, eps = case
    prm = O.params(a, u)
    o, otr = O.integrate(d, fam, prm, eps, 1e-20, trace_cap=800, predict_split=True)
    plain, ptr = O.integrate(d, fam, prm, eps, 1e-20, trace_cap=800)
    # the rule changes the run (else this case could not tell them apart)
    assert list(otr["n"]) != list(ptr["n"])
    with pg.Workspace(d, 0) as ws:
        ws.set_device_mem(MEM)
        ws.set_predict_split(True)
        r = ws.integrate_genz(fam, a, u, epsrel=eps, epsabs=1e-20, relerr=True)
        tr = ws.trace()
        # the rule is a workspace setting: off again, the plain run returns
        ws.set_predict_split(False)
        r2 = ws.integrate_genz(fam, a, u, epsrel=eps, epsabs=1e-20, relerr=True)
        tr2 = ws.trace()
    _same_run(r, tr, o, otr)
    _same_run(r2, tr2, plain, ptr)


@pytest.fixture(scope="module")
def reflib():
    lib = ctypes.CDLL(REFTESTS)
    lib.reftest_integrate_traced.argtypes = [
        ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_size_t, ctypes.c_int,
        ctypes.c_int, ctypes.POINTER(_Res), ctypes.POINTER(ctypes.c_uint64),
        ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    return lib


def header_run(lib, name, eps, predict_split=False, relerr=True):
    cap = 800
    tn = (ctypes.c_uint64 * cap)()
    tm = (ctypes.c_uint64 * cap)()
    nt = ctypes.c_size_t(0)
    r = _Res()
    rc = lib.reftest_integrate_traced(name.encode(), eps, 1e-20, MEM, int(predict_split),
                                      int(relerr), ctypes.byref(r), tn, tm, cap, ctypes.byref(nt))
    assert rc == 0, f"{name}: rc {rc}"
    k = nt.value
    return r, {"n": [int(tn[i]) for i in range(k)], "m": [int(tm[i]) for i in range(k)]}


# the reference's integrands (integrands.cuh) and their oracle descriptors
HEADER_KATS = [
    ("F_2_6D", "product_peak", 6, [50.0] * 6, [0.5] * 6, 1e-3),
    ("F_3_8D", "corner_peak", 8, [3, 2, 1, 4, 5, 6, 7, 8], None, 1e-3),
    ("F_4_5D", "gaussian", 5, [25.0] * 5, [0.5] * 5, 1e-3),
    ("F_5_8D", "c0", 8, [10.0] * 8, [0.5] * 8, 1e-3),
    ("F_6_6D", "discontinuous", 6, [10, 9, 8, 7, 6, 5], [.9, .8, .7, .6, .5, .4], 1e-3),
    ("CosSum5", "oscillatory", 5, [1.0, 2.0, 3.0, 4.0, 5.0], [0.0] * 5, 1e-6),
    ("CosSum6", "oscillatory", 6, [1.0, 2.0, 3.0, 4.0, 5.0, 6.0], [0.0] * 6, 1e-4),
]


@pytest.mark.parametrize("kat", HEADER_KATS, ids=[k[0] for k in HEADER_KATS])
def test_header_trajectory(gpu, reflib, kat):
    name, fam, d, a, u, eps = kat
    o, otr = O.integrate(d, fam, O.params(a, u), eps, 1e-20, trace_cap=800)
    r, tr = header_run(reflib, name, eps)
    _same_run(r, tr, o, otr)


def test_header_predict_split_relerr(gpu, reflib):
    """integrate<F, true> with relerr_classification = true: the rule, not the
    flag, decides each iteration (the oracle with the same rule); the run
    goes past iteration 15, where the rule switches relerr on."""
    d, a, u, eps = 3, [10.0] * 3, [0.3, 0.5, 0.7], 1e-6
    o, otr = O.integrate(d, "c0", O.params(a, u), eps, 1e-20, trace_cap=800,
                         predict_split=True)
    assert len(otr["n"]) > 16
    r, tr = header_run(reflib, "C0Three", eps, predict_split=True, relerr=True)
    _same_run(r, tr, o, otr)
    plain, ptr = O.integrate(d, "c0", O.params(a, u), eps, 1e-20, trace_cap=800)
    r2, tr2 = header_run(reflib, "C0Three", eps, predict_split=False, relerr=True)
    _same_run(r2, tr2, plain, ptr)
    assert tr["n"] != tr2["n"]
