"""Interpolation tables of the reference's stateful integrands, restated on
the CPU.  TEST INFRASTRUCTURE ONLY: the tests (and nothing in the product)
use these functions as the checker for the device interpolators of
include/pagani/interp.hpp.

Restated from the reference (operation order kept, so results are the
IEEE-double values of the formulas as written, without FMA contraction):

- hexfloat text lines -> doubles: str_to_doubles, common/cuda/str_to_doubles.hh:9-21
  (std::hexfloat extraction of whitespace-separated tokens, std::stod each)
- Interp1D: _find_smallest__index_range (bisection over index ranges) and
  the two-point formula, common/cuda/Interp1D.cuh:75-95, 186-213
- Interp2D: FindNeighbourIndices and the bilinear formula, table
  z[row * cols + col] with rows = y, cols = x, common/cuda/Interp2D.cuh:
  209-268
- Interp3D: the same search per axis and the trilinear formula, table
  v[x + nx*y + nx*ny*z], common/cuda/Interp3D.cuh:171-236
- clamp(): each coordinate clamped to the table's range before
  evaluation, Interp1D.cuh:232-243, Interp2D.cuh:281-296,
  Interp3D.cuh:277-296

The reference's searches assume the value lies inside the table (outside
it, Interp2D/3D loop forever and Interp1D divides by zero).  Values outside
use the end cell here, as the device code does: linear extrapolation.
"""
import math


def str_to_doubles(line):
    """Whitespace-separated decimal or C99 hexfloat tokens."""
    out = []
    for tok in line.split():
        t = tok.strip()
        try:
            out.append(float.fromhex(t) if "0x" in t.lower() else float(t))
        except ValueError as e:
            raise ValueError(f"not a number: {tok!r}") from e
    return out


def _cell_1d(xs, val):
    """Interp1D::_find_smallest__index_range for values inside the table."""
    n = len(xs)
    if not (xs[0] <= val):
        return 0, 1
    if not (val <= xs[n - 1]):
        return n - 2, n - 1
    left, right = 0, n - 1
    while left < right:
        mid = int((left + right) * 0.5)
        if xs[mid] <= val <= xs[mid + 1]:
            return mid, mid + 1
        if xs[mid] > val:
            right = mid
        else:
            left = mid + 1
    return left, right


def _cell(arr, val):
    """FindNeighbourIndices (Interp2D / Interp3D) for values inside the table."""
    n = len(arr)
    if not (arr[0] <= val):
        return 0, 1
    if not (val <= arr[n - 1]):
        return n - 2, n - 1
    lo, hi = 0, n - 1
    while lo <= hi:
        mid = int((hi + lo) * 0.5)
        if arr[mid] <= val <= arr[mid + 1]:
            return mid, mid + 1
        if arr[mid] > val:
            hi = mid
        else:
            lo = mid
    raise AssertionError("unreachable for values inside the table")


def interp1d(xs, zs, x):
    i0, i1 = _cell_1d(xs, x)
    y0, y1, x0, x1 = zs[i0], zs[i1], xs[i0], xs[i1]
    return (y0 * (x1 - x) + y1 * (x - x0)) / (x1 - x0)


def interp2d(xs, ys, zs, x, y):
    """zs[row * cols + col]: rows follow ys, columns follow xs."""
    cols = len(xs)
    y1, y2 = _cell(ys, y)
    x1, x2 = _cell(xs, x)
    q11 = zs[y1 * cols + x1]
    q12 = zs[y2 * cols + x1]
    q21 = zs[y1 * cols + x2]
    q22 = zs[y2 * cols + x2]
    x1v, x2v, y1v, y2v = xs[x1], xs[x2], ys[y1], ys[y2]
    f_x_y1 = q11 * (x2v - x) / (x2v - x1v) + q21 * (x - x1v) / (x2v - x1v)
    f_x_y2 = q12 * (x2v - x) / (x2v - x1v) + q22 * (x - x1v) / (x2v - x1v)
    return f_x_y1 * (y2v - y) / (y2v - y1v) + f_x_y2 * (y - y1v) / (y2v - y1v)


def interp3d(xs, ys, zs, vs, x, y, z):
    """vs[x + nx*y + nx*ny*z]."""
    nx, ny = len(xs), len(ys)

    def idx(i, j, k):
        return i + nx * j + nx * ny * k

    y0, y1 = _cell(ys, y)
    x0, x1 = _cell(xs, x)
    z0, z1 = _cell(zs, z)
    xd = (x - xs[x0]) / (xs[x1] - xs[x0])
    yd = (y - ys[y0]) / (ys[y1] - ys[y0])
    zd = (z - zs[z0]) / (zs[z1] - zs[z0])
    c00 = vs[idx(x0, y0, z0)] * (1 - xd) + vs[idx(x1, y0, z0)] * xd
    c01 = vs[idx(x0, y0, z1)] * (1 - xd) + vs[idx(x1, y0, z1)] * xd
    c10 = vs[idx(x0, y1, z0)] * (1 - xd) + vs[idx(x1, y1, z0)] * xd
    c11 = vs[idx(x0, y1, z1)] * (1 - xd) + vs[idx(x1, y1, z1)] * xd
    c0 = c00 * (1 - yd) + c10 * yd
    c1 = c01 * (1 - yd) + c11 * yd
    return c0 * (1 - zd) + c1 * zd


def clamp(v, lo, hi):
    return lo if v < lo else (hi if hi < v else v)


def bilinear_integral(xs, ys, zs):
    """Exact integral of the bilinear interpolant over the table's rectangle:
    each cell contributes its area times the mean of its four corners."""
    cols = len(xs)
    total = math.fsum(
        (xs[i + 1] - xs[i]) * (ys[j + 1] - ys[j]) * 0.25 *
        (zs[j * cols + i] + zs[j * cols + i + 1] + zs[(j + 1) * cols + i] +
         zs[(j + 1) * cols + i + 1])
        for j in range(len(ys) - 1) for i in range(cols - 1))
    return total
