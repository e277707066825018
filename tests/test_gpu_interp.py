"""Interpolation tables on the device (include/pagani/interp.hpp) against the
oracle (oracle/interp_oracle.py, pinned by the reference's interpolation
tests in test_interp_cpu.py), and PAGANI over an integrand that holds a
table.

The device formulas are compiled without FMA contraction and in the
reference's operation order, so values agree with the oracle bit for bit.
"""
import ctypes
import gzip
import os
import random
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import interp_oracle as I  # noqa: E402

REFTESTS = os.path.join(HERE, "cxx", "libreftests.so")
HMF = os.path.join(HERE, "golden", "HMF_t.dump.gz")

_dp = ctypes.POINTER(ctypes.c_double)


class _Res(ctypes.Structure):
    _fields_ = [("estimate", ctypes.c_double), ("errorest", ctypes.c_double),
                ("neval", ctypes.c_size_t), ("nregions", ctypes.c_size_t),
                ("nFinishedRegions", ctypes.c_size_t), ("status", ctypes.c_int),
                ("lastPhase", ctypes.c_int), ("chi_sq", ctypes.c_double),
                ("iters", ctypes.c_size_t)]


def _arr(v):
    return (ctypes.c_double * max(len(v), 1))(*v)


@pytest.fixture(scope="module")
def lib(pg):
    return ctypes.CDLL(REFTESTS)


def dev1(lib, xs, zs, px, clamp=0):
    out = (ctypes.c_double * len(px))()
    rc = lib.reftest_interp1d(_arr(xs), _arr(zs), ctypes.c_size_t(len(xs)), _arr(px),
                              ctypes.c_size_t(len(px)), clamp, out)
    assert rc == 0
    return list(out)


def dev2(lib, xs, ys, zs, px, py, clamp=0):
    out = (ctypes.c_double * len(px))()
    rc = lib.reftest_interp2d(_arr(xs), _arr(ys), _arr(zs), ctypes.c_size_t(len(xs)),
                              ctypes.c_size_t(len(ys)), _arr(px), _arr(py),
                              ctypes.c_size_t(len(px)), clamp, out)
    assert rc == 0
    return list(out)


def dev3(lib, xs, ys, zs, vs, px, py, pz, clamp=0):
    out = (ctypes.c_double * len(px))()
    rc = lib.reftest_interp3d(_arr(xs), _arr(ys), _arr(zs), _arr(vs), ctypes.c_size_t(len(xs)),
                              ctypes.c_size_t(len(ys)), ctypes.c_size_t(len(zs)), _arr(px),
                              _arr(py), _arr(pz), ctypes.c_size_t(len(px)), clamp, out)
    assert rc == 0
    return list(out)


def test_1d_kats_and_random(lib, gpu):
    xs = [1., 2., 3., 4., 5., 6., 7., 8., 9.]
    import math
    ys = [2 * x * (3 - x) * math.cos(x) for x in xs]
    assert dev1(lib, xs, ys, xs) == ys
    sq = [v * v for v in [1., 2., 3., 4., 5.]]
    got = dev1(lib, [1., 2., 3., 4., 5.], sq, [1.41421, 2.41421, 3.41421, 4.41421])
    for g, w in zip(got, [2.24263, 6.07105, 11.89947, 19.72789]):
        assert g == pytest.approx(w, rel=1e-4)
    rng = random.Random(1)
    kx = sorted(rng.uniform(-3, 3) for _ in range(57))
    kz = [rng.uniform(-1, 1) for _ in kx]
    pts = [rng.uniform(kx[0], kx[-1]) for _ in range(3000)] + kx + [kx[0] - 1, kx[-1] + 2]
    assert dev1(lib, kx, kz, pts) == [I.interp1d(kx, kz, p) for p in pts]
    assert dev1(lib, kx, kz, pts, clamp=1) == [
        I.interp1d(kx, kz, I.clamp(p, kx[0], kx[-1])) for p in pts]


def test_2d_kats_and_random(lib, gpu):
    xs, ys = [1., 2., 3.], [4., 5.]
    zs = [3 * x * y + 2 * x + 4 * y for y in ys for x in xs]
    assert dev2(lib, xs, ys, zs, [2.5], [4.5]) == [56.75]
    px = [x for y in ys for x in xs]
    py = [y for y in ys for x in xs]
    assert dev2(lib, xs, ys, zs, px, py) == zs
    cl = dev2(lib, xs, ys, zs, [0., 1., 4., 3., 0., 1.], [4.5, 4.5, 4.5, 4.5, 6., 5.], clamp=1)
    assert cl[0] == cl[1] and cl[2] == cl[3] and cl[4] == cl[5]
    rng = random.Random(2)
    kx = sorted(rng.uniform(0, 10) for _ in range(41))
    ky = sorted(rng.uniform(-2, 2) for _ in range(23))
    kz = [rng.uniform(-5, 5) for _ in range(len(kx) * len(ky))]
    n = 4000
    qx = [rng.uniform(kx[0], kx[-1]) for _ in range(n)] + [kx[0] - 1, kx[-1] + 1]
    qy = [rng.uniform(ky[0], ky[-1]) for _ in range(n)] + [ky[-1] + 1, ky[0] - 3]
    assert dev2(lib, kx, ky, kz, qx, qy) == [I.interp2d(kx, ky, kz, a, b)
                                             for a, b in zip(qx, qy)]


def test_3d_kats_and_random(lib, gpu):
    xs, ys, zs = [1., 2., 3.], [4., 5.], [4., 5., 6., 7.]
    nx, ny = 3, 2

    def grid(f):
        vs = [0.0] * (3 * 2 * 4)
        for k, z in enumerate(zs):
            for i, x in enumerate(xs):
                for j, y in enumerate(ys):
                    vs[i + nx * j + nx * ny * k] = f(x, y, z)
        return vs
    lin = lambda x, y, z: 2 * x + 3 * y + 4 * z  # noqa: E731
    vl = grid(lin)
    got = dev3(lib, xs, ys, zs, vl, [2.3], [4.3], [6.3])[0]
    assert got == pytest.approx(lin(2.3, 4.3, 6.3), rel=1e-15)
    pts = [(x, y, z) for z in zs for x in xs for y in ys]
    got = dev3(lib, xs, ys, zs, vl, *zip(*pts))
    assert got == [lin(*p) for p in pts]
    rng = random.Random(3)
    kx = sorted(rng.uniform(0, 1) for _ in range(9))
    ky = sorted(rng.uniform(0, 2) for _ in range(7))
    kz = sorted(rng.uniform(-1, 0) for _ in range(5))
    kv = [rng.uniform(-1, 1) for _ in range(9 * 7 * 5)]
    q = [(rng.uniform(kx[0], kx[-1]), rng.uniform(ky[0], ky[-1]), rng.uniform(kz[0], kz[-1]))
         for _ in range(2000)]
    assert dev3(lib, kx, ky, kz, kv, *zip(*q)) == [I.interp3d(kx, ky, kz, kv, *p) for p in q]


def _hmf(tmp_path):
    lines = gzip.open(HMF, "rt").read().splitlines()
    path = str(tmp_path / "HMF_t.dump")
    open(path, "w").write("\n".join(lines) + "\n")
    xs, ys, zs = (I.str_to_doubles(lines[k]) for k in range(3))
    return path, xs, ys, zs


def test_table_file_on_device(lib, gpu, tmp_path):
    path, xs, ys, zs = _hmf(tmp_path)
    rng = random.Random(4)
    n = 5000
    px = [rng.uniform(xs[0], xs[-1]) for _ in range(n)] + xs[:50]
    py = [rng.uniform(ys[0], ys[-1]) for _ in range(n)] + ys[:50]
    out = (ctypes.c_double * len(px))()
    rc = lib.reftest_interp2d_file(path.encode(), _arr(px), _arr(py), ctypes.c_size_t(len(px)),
                                   out)
    assert rc == 0
    assert list(out) == [I.interp2d(xs, ys, zs, a, b) for a, b in zip(px, py)]


def test_integrate_table_integrand(lib, gpu, tmp_path):
    path, xs, ys, zs = _hmf(tmp_path)
    exact = I.bilinear_integral(xs, ys, zs)
    r = _Res()
    lows, highs = _arr([xs[0], ys[0]]), _arr([xs[-1], ys[-1]])
    eps = 1e-5
    rc = lib.reftest_integrate_table_file(path.encode(), lows, highs, ctypes.c_double(eps),
                                          ctypes.c_double(1e-20), ctypes.byref(r))
    assert rc == 0
    assert r.status == 0
    assert abs(r.estimate - exact) <= 2 * eps * abs(exact), (r.estimate, exact, r.errorest)


def test_table_copies_do_not_leak(lib, gpu):
    before, after = ctypes.c_size_t(0), ctypes.c_size_t(0)
    last = ctypes.c_double(0)
    rc = lib.reftest_interp_mem_usage(20, ctypes.byref(before), ctypes.byref(after),
                                      ctypes.byref(last))
    assert rc == 0
    assert before.value == after.value
    # the tables interpolate a(x) = 1 + 2x and b(x, y) = 1 + 2x + 3y exactly;
    # the integral of a*b over the unit square is 13/3 + 3 = 22/3
    assert last.value == pytest.approx(22.0 / 3.0, rel=1e-6)
