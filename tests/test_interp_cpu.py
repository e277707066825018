"""The interpolation oracle (oracle/interp_oracle.py) pinned against the
reference's own interpolation tests, and the C++ table-file reader
(include/pagani/str_to_doubles.hh) against Python's hexfloat parser.

Known answers from test/cuda/common/Interpolation1D.cu, Interpolation2D.cu
and Interpolation3D.cu; the hexfloat fixture is the reference's
test/cuda/pagani/data/HMF_t.dump (tests/golden/HMF_t.dump.gz).
"""
import ctypes
import gzip
import math
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import interp_oracle as I  # noqa: E402

HMF = os.path.join(HERE, "golden", "HMF_t.dump.gz")
REFTESTS = os.path.join(HERE, "cxx", "libreftests.so")


def hmf_lines():
    return gzip.open(HMF, "rt").read().splitlines()


# ---- 1D (Interpolation1D.cu) ------------------------------------------------

def test_1d_exact_at_knots():
    xs = [1., 2., 3., 4., 5., 6., 7., 8., 9.]
    ys = [2 * x * (3 - x) * math.cos(x) for x in xs]
    for x, y in zip(xs, ys):
        assert I.interp1d(xs, ys, x) == y


@pytest.mark.parametrize("x,want", [(1.41421, 2.24263), (2.41421, 6.07105),
                                    (3.41421, 11.89947), (4.41421, 19.72789)])
def test_1d_on_quadratic(x, want):
    xs = [1., 2., 3., 4., 5.]
    ys = [v * v for v in xs]
    assert I.interp1d(xs, ys, x) == pytest.approx(want, rel=1e-4)


# ---- 2D (Interpolation2D.cu) ------------------------------------------------

def _grid2(xs, ys, f):
    return [f(x, y) for y in ys for x in xs]  # z[j*nx + i]


def test_2d_interpolation_and_knots():
    xs, ys = [1., 2., 3.], [4., 5.]
    f = lambda x, y: 3 * x * y + 2 * x + 4 * y  # noqa: E731
    zs = _grid2(xs, ys, f)
    assert I.interp2d(xs, ys, zs, 2.5, 4.5) == 56.75
    for x in xs:
        for y in ys:
            assert I.interp2d(xs, ys, zs, x, y) == f(x, y)


def test_2d_on_bilinear():
    xs, ys = [1., 2., 3.], [1., 2., 3., 4.]
    zs = _grid2(xs, ys, lambda x, y: 2 * x + 3 * y - 5)
    assert I.interp2d(xs, ys, zs, 2.5, 1.5) == 4.5


def test_2d_clamp():
    xs, ys = [1., 2., 3.], [4., 5.]
    zs = _grid2(xs, ys, lambda x, y: 3 * x * y + 2 * x + 4 * y)

    def clamp(x, y):
        return I.interp2d(xs, ys, zs, I.clamp(x, 1., 3.), I.clamp(y, 4., 5.))
    for a, b in [((0., 4.5), (1., 4.5)), ((4., 4.5), (3., 4.5)), ((2., 3.), (2., 4.)),
                 ((2., 5.5), (2., 5.)), ((0., 0.), (1., 4.)), ((4., 3.), (3., 4.)),
                 ((0., 6.), (1., 5.)), ((4., 6.), (3., 5.))]:
        assert clamp(*a) == clamp(*b)


# ---- 3D (Interpolation3D.cu) ------------------------------------------------

XS3, YS3, ZS3 = [1., 2., 3.], [4., 5.], [4., 5., 6., 7.]


def _grid3(f):
    nx, ny = len(XS3), len(YS3)
    vs = [0.0] * (nx * ny * len(ZS3))
    for k, z in enumerate(ZS3):
        for i, x in enumerate(XS3):
            for j, y in enumerate(YS3):
                vs[i + nx * j + nx * ny * k] = f(x, y, z)
    return vs


def _quad3(x, y, z):
    return (2 * x ** 2 + 3 * y ** 2 + 4 * z ** 2 + 5 * x + 6 * y + 7 * z + x * y - x * z
            + z * y)


def _lin3(x, y, z):
    return 2 * x + 3 * y + 4 * z


def test_3d_quadratic_and_linear():
    vq = _grid3(_quad3)
    assert I.interp3d(XS3, YS3, ZS3, vq, 2.1, 4.9, 6.1) == pytest.approx(
        _quad3(2.1, 4.9, 6.1), rel=1e-2)
    vl = _grid3(_lin3)
    assert I.interp3d(XS3, YS3, ZS3, vl, 2.3, 4.3, 6.3) == pytest.approx(
        _lin3(2.3, 4.3, 6.3), rel=1e-15)


@pytest.mark.parametrize("f", [_lin3, _quad3])
def test_3d_exact_at_knots(f):
    vs = _grid3(f)
    for z in ZS3:
        for x in XS3:
            for y in YS3:
                assert I.interp3d(XS3, YS3, ZS3, vs, x, y, z) == f(x, y, z)


# ---- exact integral of a bilinear interpolant --------------------------------

def test_bilinear_integral_closed_form():
    # a bilinear function is its own interpolant on any grid
    xs = [0.0, 0.3, 0.35, 1.0, 2.5]
    ys = [-1.0, 0.0, 0.7, 2.0]
    a, b, c, d = 1.5, -0.25, 2.0, 0.75
    zs = _grid2(xs, ys, lambda x, y: a + b * x + c * y + d * x * y)
    x0, x1, y0, y1 = xs[0], xs[-1], ys[0], ys[-1]
    exact = (a * (x1 - x0) * (y1 - y0) + b * (x1 ** 2 - x0 ** 2) / 2 * (y1 - y0)
             + c * (y1 ** 2 - y0 ** 2) / 2 * (x1 - x0)
             + d * (x1 ** 2 - x0 ** 2) / 2 * (y1 ** 2 - y0 ** 2) / 2)
    assert I.bilinear_integral(xs, ys, zs) == pytest.approx(exact, rel=1e-14)


# ---- the table-file format ----------------------------------------------------

def test_fixture_shape():
    lines = hmf_lines()
    xs, ys, zs = (I.str_to_doubles(lines[k]) for k in range(3))
    assert (len(xs), len(ys), len(zs)) == (969, 100, 969 * 100)
    assert all(a < b for a, b in zip(xs, xs[1:])) and all(a < b for a, b in zip(ys, ys[1:]))


@pytest.fixture(scope="module")
def parse():
    if not os.path.exists(REFTESTS):
        pytest.skip("tests/cxx/libreftests.so not built")
    import gpuintegration_amd as pg
    pg.load()
    lib = ctypes.CDLL(REFTESTS)
    fn = lib.reftest_parse_doubles
    fn.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t,
                   ctypes.POINTER(ctypes.c_size_t)]

    def run(line, cap=200000):
        out = (ctypes.c_double * cap)()
        n = ctypes.c_size_t(0)
        rc = fn(line.encode(), out, cap, ctypes.byref(n))
        return rc, list(out[:min(n.value, cap)])
    return run


def test_cpp_reader_matches_hexfloat_parse(parse):
    lines = hmf_lines()
    for k in (0, 1, 3):
        rc, got = parse(lines[k])
        assert rc == 0
        assert got == I.str_to_doubles(lines[k])
    toks = lines[2].split()[:5000]
    rc, got = parse(" ".join(toks))
    assert rc == 0 and got == [float.fromhex(t) for t in toks]


def test_cpp_reader_decimal_and_errors(parse):
    rc, got = parse("  1.5 -2e-3\t0x1p-2 7 ")
    assert rc == 0 and got == [1.5, -2e-3, 0.25, 7.0]
    assert parse("")[1] == []
    assert parse("1.0 abc")[0] != 0
