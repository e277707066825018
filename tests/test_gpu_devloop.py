"""The device-resident loop (iterations decided on the device in rounds)
against the host-driven loop: identical results, traces and region counts,
with fewer host round trips.
"""
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # family, d, epsrel: small lists that stay on the device throughout, and
    # larger ones that hand over to the host (classifier, budget fix)
    ("gaussian", 3, 1e-6),
    ("gaussian", 5, 1e-6),
    ("gaussian", 8, 1e-4),
    ("oscillatory", 5, 1e-3),
    ("oscillatory", 6, 1e-7),
    ("product_peak", 4, 1e-6),
    ("corner_peak", 6, 1e-6),
    ("c0", 5, 1e-5),
    ("discontinuous", 4, 1e-4),
    ("discontinuous", 6, 1e-5),
]


def _params(fam, d, seed):
    from gpuintegration_amd.workloads import MT19937_64
    g = MT19937_64(seed)
    if fam == "oscillatory":
        a = [1.0 + 0.5 * g.uniform() for _ in range(d)]
    elif fam == "gaussian":
        a = [10.0 * (1.0 + 0.05 * (2.0 * g.uniform() - 1.0)) for _ in range(d)]
    else:
        a = [5.0 + 10.0 * g.uniform() for _ in range(d)]
    u = [0.3 + 0.4 * g.uniform() for _ in range(d)]
    return a, u


def _run(pg, fam, a, u, eps, device, mem=None, relerr=True):
    with pg.Workspace(len(a), 0) as ws:
        ws.set_device_loop(device)
        if mem:
            ws.set_device_mem(mem)
        r = ws.integrate_genz(fam, a, u, epsrel=eps, epsabs=1e-20, relerr=relerr)
        return r, ws.trace(), ws.stats()


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1]}d-{c[2]:g}" for c in CASES])
def test_device_loop_matches_host_loop(pg, gpu, case):
    fam, d, eps = case
    a, u = _params(fam, d, seed=d * 7 + len(fam))
    for mem in (None, int(16e9)):
        rh, th, sh = _run(pg, fam, a, u, eps, False, mem)
        rd, td, sd = _run(pg, fam, a, u, eps, True, mem)
        assert (rd.estimate, rd.errorest, rd.nregions, rd.iters, rd.status) == \
               (rh.estimate, rh.errorest, rh.nregions, rh.iters, rh.status)
        for k in ("n", "m", "estimate", "errorest"):
            assert list(td[k]) == list(th[k]), k
        assert sd["regions_evaluated"] == sh["regions_evaluated"]
        assert sd["iterations"] == sh["iterations"]
        assert sh["device_iterations"] == 0
        assert sd["device_iterations"] > 0
        assert sd["device_rounds"] <= sd["device_iterations"]


def test_device_loop_without_relerr_classification(pg, gpu):
    a, u = _params("gaussian", 4, seed=3)
    rh, th, _ = _run(pg, "gaussian", a, u, 1e-6, False, relerr=False)
    rd, td, sd = _run(pg, "gaussian", a, u, 1e-6, True, relerr=False)
    assert (rd.estimate, rd.errorest, rd.nregions, rd.iters) == \
           (rh.estimate, rh.errorest, rh.nregions, rh.iters)
    assert list(td["n"]) == list(th["n"])
    assert sd["device_iterations"] > 0


def test_device_loop_round_trips(pg, gpu):
    # a run that stays small: far fewer host round trips than iterations
    a, u = _params("gaussian", 5, seed=11)
    r, tr, st = _run(pg, "gaussian", a, u, 1e-6, True)
    assert st["device_iterations"] >= 4
    assert st["device_rounds"] < st["device_iterations"]


@pytest.mark.parametrize("case", [("gaussian", 5, 1e-4), ("product_peak", 6, 1e-5),
                                  ("corner_peak", 4, 1e-6)])
def test_device_loop_fp32_integrand(pg, gpu, case):
    fam, d, eps = case
    a, u = _params(fam, d, seed=d + 101)
    out = []
    for device in (False, True):
        with pg.Workspace(d, 0) as ws:
            ws.set_device_loop(device)
            r = ws.integrate_genz(fam, a, u, epsrel=eps, epsabs=1e-20, precision=pg.FP32)
            out.append((r.estimate, r.errorest, r.nregions, r.iters, r.status,
                        ws.stats()["device_iterations"]))
    assert out[0][:5] == out[1][:5]
    assert out[1][5] > 0


def test_device_loop_user_functors(pg, gpu):
    """The reference tests' integrands through the C++ header, both loops."""
    import ctypes
    import os
    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx",
                                   "libreftests.so"))

    class _Res(ctypes.Structure):
        _fields_ = [("estimate", ctypes.c_double), ("errorest", ctypes.c_double),
                    ("neval", ctypes.c_size_t), ("nregions", ctypes.c_size_t),
                    ("nFinishedRegions", ctypes.c_size_t), ("status", ctypes.c_int),
                    ("lastPhase", ctypes.c_int), ("chi_sq", ctypes.c_double),
                    ("iters", ctypes.c_size_t)]
    lib.reftest_integrate.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(_Res)]
    try:
        for name, eps in (("F_4_5D", 1e-6), ("F_2_6D", 1e-4), ("F_6_6D", 1e-4), ("XY", 1e-8),
                          ("Fun6", 1e-4), ("BoxIntegral8_15", 1e-5)):
            got = []
            for on in (0, 1):
                lib.reftest_set_device_loop(on)
                r = _Res()
                assert lib.reftest_integrate(name.encode(), eps, 1e-20, ctypes.byref(r)) == 0
                got.append((r.estimate, r.errorest, r.nregions, r.iters, r.status))
            assert got[0] == got[1], name
    finally:
        lib.reftest_set_device_loop(1)


def test_device_loop_round_graphs(pg, gpu):
    """Full-length rounds replayed from HIP graphs give the launches' own
    results bit for bit: a run with many small iterations (2D discontinuous
    at 1e-14: about 50 iterations of at most 850 regions, two full rounds
    of 16), with graphs, without, and on the host loop."""
    fam, a, u, eps = "discontinuous", [5.0, 6.0], [0.4, 0.6], 1e-14
    runs = {}
    for name, device, graphs in (("graphs", True, True), ("launches", True, False),
                                 ("host", False, False)):
        with pg.Workspace(2, 0) as ws:
            ws.set_device_loop(device)
            ws.set_graphs(graphs)
            r = ws.integrate_genz(fam, a, u, epsrel=eps, epsabs=1e-300)
            runs[name] = (r, ws.trace(), ws.stats())
    key = lambda r: (r.estimate, r.errorest, r.nregions, r.iters, r.status)
    assert key(runs["graphs"][0]) == key(runs["launches"][0]) == key(runs["host"][0])
    for k in ("n", "m", "estimate", "errorest"):
        assert list(runs["graphs"][1][k]) == list(runs["launches"][1][k]) == \
            list(runs["host"][1][k]), k
    assert runs["graphs"][2]["graph_launches"] >= 2
    assert runs["launches"][2]["graph_launches"] == 0


def test_tuned_launch_shapes(pg, gpu):
    """The autotuner's table picks launch shapes only.  With it on and off
    the host loop and the device loop still agree bit for bit; between the
    table and the fixed rule a region's sums may round differently (another
    slice count splits a batch's points differently among waves, and one
    wave per batch takes the factored sums), so the two runs agree within
    their error estimates."""
    for fam, d, eps in (("gaussian", 5, 1e-6), ("corner_peak", 6, 1e-6)):
        a, u = _params(fam, d, seed=d + 5)
        out = {}
        for tuned in (True, False):
            for device in (True, False):
                with pg.Workspace(d, 0) as ws:
                    ws.set_tuned_slices(tuned)
                    ws.set_device_loop(device)
                    r = ws.integrate_genz(fam, a, u, epsrel=eps, epsabs=1e-20)
                    out[tuned, device] = (r.estimate, r.errorest, r.nregions, r.iters,
                                          r.status)
        for tuned in (True, False):
            assert out[tuned, True] == out[tuned, False], (fam, tuned)
        t, f = out[True, True], out[False, True]
        assert t[4] == f[4] == 0
        assert abs(t[0] - f[0]) <= t[1] + f[1]
