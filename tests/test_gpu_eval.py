"""Per-region rule evaluation on the GPU vs the CPU oracle.

Replaces the reference's RegionSampling.cu checks (constant integrands) with
a direct comparison of every region's estimate, error estimate and split
axis against the oracle's restatement of Sample.cuh / Phases.cuh, for all
six Genz families and d = 2..8.

Tolerances (written here, used below).  The GPU forms each rule sum per
symmetric set (sum_s W[s][r] * sum_{p in s} f_p) while the oracle accumulates
w * f point by point, so the sums round differently.  A_r = sum_p |w_r f_p|
(reported per region by the oracle) bounds that rounding:
  estimate:  |gpu - oracle| <= 1e-13 * vol * A_0
  error:     the null-rule errors e_1..e_3 are re-derived with every rule sum
             perturbed by 1e-12 * A_r (oracle_api.err_bounds); the GPU error
             must match the oracle's branch of Sample.cuh:282-288 within that
             bound, or either branch where the branch condition itself lies
             within the rounding.
  split axis: identical wherever the two largest fourth differences differ
             by more than 1e-12 times the magnitude of their terms (a fourth
             difference cancels; closer ones are ties decided by rounding)
  FP32 integrand: estimate within 2e-5 * vol * A_0 (+1e-35 * vol for values
             below FP32's normal range) of the FP64 oracle.
"""
import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

FAMILIES = ["oscillatory", "product_peak", "corner_peak", "gaussian", "c0", "discontinuous"]


def genz_params(family, d, seed):
    """Seeded Genz parameters (mt19937_64 draw order a_1..a_d, u_1..u_d)."""
    rng = np.random.default_rng(seed)
    if family == "oscillatory":
        a = 1.0 + 0.5 * rng.random(d)
        u = rng.random(d)
    elif family == "gaussian":
        a = 10.0 * (1 + 0.05 * (2 * rng.random(d) - 1))
        u = 0.2 + 0.6 * rng.random(d)
    elif family == "corner_peak":
        a = 0.5 + rng.random(d)
        u = np.zeros(d)
    elif family == "discontinuous":
        a = 5.0 + 10.0 * rng.random(d) * 0.3
        u = 0.3 + 0.4 * rng.random(d)
    else:
        a = 5.0 + 10.0 * rng.random(d)
        u = rng.random(d)
    return a, u


def random_regions(d, n, seed):
    rng = np.random.default_rng(seed)
    H = 10.0 ** rng.uniform(-3, -0.3, size=(d, n))
    L = rng.random((d, n)) * (1.0 - H)
    return np.ascontiguousarray(L.ravel()), np.ascontiguousarray(H.ravel())


def to_dev(a, dev, dtype=None):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def gpu_eval(ws, dev, family, a, u, L, H, n, precision=64, lows=None, highs=None):
    import torch
    Ld, Hd = to_dev(L, dev), to_dev(H, dev)
    est = torch.empty(n, dtype=torch.float64, device=dev)
    err = torch.empty_like(est)
    dim = torch.empty(n, dtype=torch.uint8, device=dev)
    ws.eval_genz(family, a, u, n, Ld, Hd, est, err, dim, lows=lows, highs=highs,
                 precision=precision)
    return est.cpu().numpy(), err.cpu().numpy(), dim.cpu().numpy().astype(np.int32)


EPS = 2.0 ** -53


def centred_cond(family, a):
    """Relative difference the centred integrand forms may add.  The device
    forms t = x - u in one FMA from (centre - u), the oracle as (centre - g h)
    - u, so t differs by about 2 eps (|t| <= 1); f inherits
    sum_j |d log f / d t_j| |dt_j|: a_j for C0, at most a_j for the product
    peak.  The Gaussian takes v = a (x - u) from the centre offset and
    half-width each scaled by a (three roundings: |dv| <= eps a (|c - u| +
    |g h| + |t|) <= 2.5 eps a), so v^2 moves by at most 5 eps a^2; with the
    oracle's own t error (2 eps, 4 eps a^2 on a^2 t^2) and the roundings of
    the squares and the sum, 12 eps a_j^2 per axis."""
    if family == "gaussian":
        return 12 * EPS * sum(v * v for v in a)
    if family in ("c0", "product_peak"):
        return 4 * EPS * sum(abs(v) for v in a)
    if family == "discontinuous":
        # exponent formed as sum a u + sum a (x - u): 2 eps sum a_j (|u_j| + |x_j|)
        return 8 * EPS * sum(abs(v) for v in a)
    return 0.0


def check_region_parity(d, family, ge, gerr, gdim, oe, oerr, odim, fd, rs, cond=0.0,
                        floor=False):
    """floor: add the FP64 evaluation floor of the kernel-point oracle,
    (d + 2) eps times the region's sum of |w_0 J| x f's sensitivity to one
    rounding of its inner sum (oracle_api.eval_floor), on top of 1e-13."""
    scale = O.est_scale(rs)
    extra = (d + 2) * EPS * O.eval_floor(rs) if floor else 0.0
    bad = np.abs(ge - oe) > (1e-13 + cond) * scale + extra + 1e-300
    assert not bad.any(), (
        f"{family} d={d}: {bad.sum()} estimates off; worst "
        f"{np.max(np.abs(ge - oe) / np.maximum(scale, 1e-300)):.3e} of scale")
    va, vb, ta, tb, cond, amb = O.err_bounds(d, rs)
    # the oracle's own error must be its branch value (checks the record)
    assert np.allclose(oerr, np.where(cond, va, vb), rtol=1e-15, atol=0)
    ok_a = np.abs(gerr - va) <= ta + 1e-300
    ok_b = np.abs(gerr - vb) <= tb + 1e-300
    ok = np.where(amb, ok_a | ok_b, np.where(cond, ok_a, ok_b))
    assert ok.all(), (f"{family} d={d}: {(~ok).sum()} error estimates off "
                      f"({amb.sum()} ambiguous branches)")
    srt = np.sort(fd, axis=1)
    clear = (srt[:, -1] - srt[:, -2]) > 1e-12 * rs[:, 14]
    mism = (gdim != odim) & clear
    assert not mism.any(), f"{family} d={d}: {mism.sum()} split axes differ on non-ties"


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("d", [2, 3, 5, 8])
@pytest.mark.parametrize("slices", [0, 1])
def test_eval_random_regions_fp64(pg, gpu, family, d, slices):
    """slices 0: the library's choice (several waves per batch at these
    sizes: the table-driven point loop); 1: one wave per batch, where the
    separable families take their term-difference evaluation."""
    n = 3000 if d < 8 else 1000
    a, u = genz_params(family, d, seed=100 + d)
    L, H = random_regions(d, n, seed=7 * d)
    with pg.Workspace(d, 0) as ws:
        if slices:
            ws.set_slices(slices)
        ge, gerr, gdim = gpu_eval(ws, gpu, family, a, u, L, H, n)
    prm = O.params(a, u)
    oe, oerr, odim, fd, rs = O.eval_regions(d, family, prm, L, H, n, want_extra=True)
    check_region_parity(d, family, ge, gerr, gdim, oe, oerr, odim, fd, rs, centred_cond(family, a))


@pytest.mark.parametrize("family,scale", [("gaussian", 20.0), ("c0", 200.0),
                                          ("discontinuous", 100.0), ("oscillatory", 2e5)])
def test_eval_factored_range_fallback(pg, gpu, family, scale):
    """Parameters large enough that part of the regions leave the range of
    the factored sums (a Gaussian or C0 centre sum above 650, a
    discontinuous exponent beyond +-550, oscillatory angles beyond 2^20):
    in one wave, lanes take the factored sums or the point loop region by
    region, and every region must still match the oracle."""
    d, n = 8, 1500
    a, u = genz_params(family, d, seed=77)
    a = [scale * (1.0 + 0.1 * k) for k in range(d)]
    L, H = random_regions(d, n, seed=79)
    with pg.Workspace(d, 0) as ws:
        ws.set_slices(1)
        ge, gerr, gdim = gpu_eval(ws, gpu, family, a, u, L, H, n)
    oe, oerr, odim, fd, rs = O.eval_regions(d, family, O.params(a, u), L, H, n,
                                            want_extra=True)
    if family == "oscillatory":
        # the angle sums (about 1e6 here) carry 2^-53 of themselves on both
        # sides: the estimates agree within that conditioning; the error
        # estimates (differences of nearly equal rule sums) and the split
        # axes amplify it and are not compared
        scale_r = O.est_scale(rs)
        cond = 8 * EPS * sum(a)
        assert np.all(np.abs(ge - oe) <= (1e-13 + cond) * scale_r + 1e-300)
        assert np.all(np.isfinite(gerr)) and np.all(gerr >= 0)
        return
    check_region_parity(d, family, ge, gerr, gdim, oe, oerr, odim, fd, rs,
                        centred_cond(family, a))


@pytest.mark.parametrize("d", [4, 6, 7])
def test_eval_all_families_remaining_dims(pg, gpu, d):
    n = 2000
    L, H = random_regions(d, n, seed=11 * d)
    with pg.Workspace(d, 0) as ws:
        for family in FAMILIES:
            a, u = genz_params(family, d, seed=200 + d)
            ge, gerr, gdim = gpu_eval(ws, gpu, family, a, u, L, H, n)
            oe, oerr, odim, fd, rs = O.eval_regions(d, family, O.params(a, u), L, H, n,
                                                    want_extra=True)
            check_region_parity(d, family, ge, gerr, gdim, oe, oerr, odim, fd, rs,
                                centred_cond(family, a))


@pytest.mark.parametrize("slices", [1, 2, 4, 8])
def test_eval_slices_agree(pg, gpu, slices):
    """Every launch shape (waves sharing a region batch) gives the oracle's sums."""
    d, n = 8, 777
    a, u = genz_params("gaussian", d, seed=5)
    L, H = random_regions(d, n, seed=3)
    with pg.Workspace(d, 0) as ws:
        ws.set_slices(slices)
        ge, gerr, gdim = gpu_eval(ws, gpu, "gaussian", a, u, L, H, n)
    oe, oerr, odim, fd, rs = O.eval_regions(d, "gaussian", O.params(a, u), L, H, n,
                                            want_extra=True)
    check_region_parity(d, "gaussian", ge, gerr, gdim, oe, oerr, odim, fd, rs,
                        centred_cond("gaussian", a))


def test_eval_uniform_split_8d_config(pg, gpu):
    """The 8D Gaussian workload's first iteration: 4^8 = 65,536 regions."""
    import torch
    d, k = 8, 4
    n = k ** d
    a, u = genz_params("gaussian", d, seed=2024)
    with pg.Workspace(d, 0) as ws:
        Ld = torch.empty(n * d, dtype=torch.float64, device=gpu)
        Hd = torch.empty_like(Ld)
        ws.uniform_split(k, Ld, Hd)
        L, H = Ld.cpu().numpy(), Hd.cpu().numpy()
        Lo, Ho, _ = O.uniform_split(d, k)
        assert np.array_equal(L, Lo) and np.array_equal(H, Ho)
        ge, gerr, gdim = gpu_eval(ws, gpu, "gaussian", a, u, L, H, n)
    oe, oerr, odim, fd, rs = O.eval_regions(d, "gaussian", O.params(a, u), L, H, n,
                                            want_extra=True)
    check_region_parity(d, "gaussian", ge, gerr, gdim, oe, oerr, odim, fd, rs,
                        centred_cond("gaussian", a))


def test_eval_general_volume(pg, gpu):
    """Non-unit integration box: the Jacobian and the affine map."""
    d, n = 5, 500
    lows, highs = [-1.0, 0.0, 2.0, -0.5, 0.25], [1.0, 3.0, 2.5, 0.5, 0.75]
    a, u = genz_params("c0", d, seed=9)
    L, H = random_regions(d, n, seed=19)
    with pg.Workspace(d, 0) as ws:
        ge, gerr, gdim = gpu_eval(ws, gpu, "c0", a, u, L, H, n, lows=lows, highs=highs)
        ws.set_slices(1)
        ge1, gerr1, gdim1 = gpu_eval(ws, gpu, "c0", a, u, L, H, n, lows=lows, highs=highs)
    oe, oerr, odim, fd, rs = O.eval_regions(d, "c0", O.params(a, u), L, H, n,
                                            lows=np.array(lows), highs=np.array(highs),
                                            want_extra=True)
    check_region_parity(d, "c0", ge, gerr, gdim, oe, oerr, odim, fd, rs)
    check_region_parity(d, "c0", ge1, gerr1, gdim1, oe, oerr, odim, fd, rs)


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("slices", [0, 1])
def test_eval_fp32_integrand(pg, gpu, family, slices):
    """Mixed precision: FP32 integrand, FP64 sums, vs the all-FP64 oracle."""
    d, n = 6, 2000
    a, u = genz_params(family, d, seed=31)
    L, H = random_regions(d, n, seed=37)
    with pg.Workspace(d, 0) as ws:
        if slices:
            ws.set_slices(slices)
        ge, gerr, gdim = gpu_eval(ws, gpu, family, a, u, L, H, n, precision=32)
    oe, oerr, odim, fd, rs = O.eval_regions(d, family, O.params(a, u), L, H, n,
                                            want_extra=True)
    scale = O.est_scale(rs) + 1e-35 * rs[:, 10]
    if family == "discontinuous":
        # FP32 coordinates can fall on the other side of a step: compare only
        # regions whose points all lie away from the discontinuity
        far = np.ones(n, dtype=bool)
        Lm, Hm = L.reshape(d, n), H.reshape(d, n)
        for j in range(d):
            far &= (Lm[j] > u[j] + 1e-5) | (Lm[j] + Hm[j] < u[j] - 1e-5)
        ge, oe, scale = ge[far], oe[far], scale[far]
    assert np.all(np.abs(ge - oe) <= 2e-5 * scale + 1e-300)


def test_eval_constant_kat(pg, gpu):
    """RegionSampling.cu: constant c over a k^2 split gives c/n per region (1e-6)."""
    import ctypes
    import os
    lib_path = os.path.join(os.path.dirname(__file__), "cxx", "libreftests.so")
    lib = ctypes.CDLL(lib_path)
    lib.reftest_sample.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_size_t]
    for name, c in [(b"PTest", 15.37), (b"NTest", -15.37), (b"ZTest", 0.0)]:
        for k in range(5, 15):
            n = k * k
            est = np.zeros(n)
            err = np.zeros(n)
            got = lib.reftest_sample(name, k, est.ctypes.data, err.ctypes.data, n)
            assert got == n
            assert np.allclose(est, c / n, rtol=1e-6, atol=1e-300)


# ---------------------------------------------------------------------------
# Parity at the kernel's own points.  The tests above compare with the
# reference's point map and allow for the conditioning of the steep
# integrands (centred_cond: up to 42x the 1e-13 bound for the a = 25
# Gaussian).  Here the oracle forms every point exactly as the kernel does
# (centre c = fma(.5 (lo + up), range, lows), minus the origin and times the
# axis scale the functor declares, v = fma(-g, h, c)) and evaluates f and
# the sums in long double: what is left is the kernel's own rounding.
#   - The headline's first iteration (the 4^8 split of the 8D Gaussian, at
#     one wave per batch and at the library's launch shape) and F_4_5D's
#     parameters are held to 1e-13 * vol * A_0 with no allowance.
#   - Elsewhere the bound adds the floor any FP64 evaluation of f has at
#     those points (check_region_parity's `floor`): f's sensitivity to one
#     rounding of its inner sum, point by point.  It matters only where that
#     sum is large against f's scale -- a cosine near a zero (measured: one
#     2D oscillatory region at 1.02e-13), a Gaussian exponent of several
#     hundred on a stretched box (1.67e-13) -- and is below 1e-14 of the
#     scale elsewhere.

KFORM_DIMS = [2, 3, 4, 5, 6, 7, 8]


def _kform(d, family, a, u, L, H, n, plain=False, lows=None, highs=None):
    return O.eval_regions_kform(d, family, O.params(a, u), L, H, n, lows=lows, highs=highs,
                                plain=plain)


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("d", KFORM_DIMS)
@pytest.mark.parametrize("slices", [0, 1])
def test_eval_kernel_points_fp64(pg, gpu, family, d, slices):
    n = 2000 if d < 8 else 1000
    a, u = genz_params(family, d, seed=300 + d)
    L, H = random_regions(d, n, seed=13 * d + 1)
    with pg.Workspace(d, 0) as ws:
        if slices:
            ws.set_slices(slices)
        ge, gerr, gdim = gpu_eval(ws, gpu, family, a, u, L, H, n)
    oe, oerr, odim, fd, rs = _kform(d, family, a, u, L, H, n)
    check_region_parity(d, family, ge, gerr, gdim, oe, oerr, odim, fd, rs, floor=True)


@pytest.mark.parametrize("slices", [0, 1])
def test_eval_kernel_points_config2_split(pg, gpu, slices):
    """The 8D Gaussian workload's first iteration (4^8 regions) at one wave
    per batch (the factored set sums of the headline run) and at the
    library's launch shape."""
    import torch
    d, k = 8, 4
    n = k ** d
    from gpuintegration_amd.workloads import gaussian_8d
    _, a, u = gaussian_8d(2024)
    with pg.Workspace(d, 0) as ws:
        if slices:
            ws.set_slices(slices)
        Ld = torch.empty(n * d, dtype=torch.float64, device=gpu)
        Hd = torch.empty_like(Ld)
        ws.uniform_split(k, Ld, Hd)
        L, H = Ld.cpu().numpy(), Hd.cpu().numpy()
        ge, gerr, gdim = gpu_eval(ws, gpu, "gaussian", a, u, L, H, n)
    oe, oerr, odim, fd, rs = _kform(d, "gaussian", a, u, L, H, n)
    check_region_parity(d, "gaussian", ge, gerr, gdim, oe, oerr, odim, fd, rs)


@pytest.mark.parametrize("slices", [0, 1])
def test_eval_kernel_points_f4_5d(pg, gpu, slices):
    """F_4_5D's parameters (a = 25, u = 0.5; the steepest Gaussian of the
    reference's tests): a 3^5 split and random sub-boxes."""
    d = 5
    a, u = [25.0] * d, [0.5] * d
    L3, H3, n3 = O.uniform_split(d, 3)
    Lr, Hr = random_regions(d, 3000, seed=45)
    with pg.Workspace(d, 0) as ws:
        if slices:
            ws.set_slices(slices)
        for L, H, n in ((L3, H3, n3), (Lr, Hr, 3000)):
            ge, gerr, gdim = gpu_eval(ws, gpu, "gaussian", a, u, L, H, n)
            oe, oerr, odim, fd, rs = _kform(d, "gaussian", a, u, L, H, n)
            check_region_parity(d, "gaussian", ge, gerr, gdim, oe, oerr, odim, fd, rs)


def test_eval_kernel_points_general_volume(pg, gpu):
    d, n = 5, 800
    lows, highs = [-1.0, 0.0, 2.0, -0.5, 0.25], [1.0, 3.0, 2.5, 0.5, 0.75]
    L, H = random_regions(d, n, seed=21)
    with pg.Workspace(d, 0) as ws:
        for family in FAMILIES:
            a, u = genz_params(family, d, seed=41)
            ge, gerr, gdim = gpu_eval(ws, gpu, family, a, u, L, H, n, lows=lows, highs=highs)
            oe, oerr, odim, fd, rs = _kform(d, family, a, u, L, H, n, lows=np.array(lows),
                                            highs=np.array(highs))
            check_region_parity(d, family, ge, gerr, gdim, oe, oerr, odim, fd, rs, floor=True)


@pytest.fixture(scope="module")
def plainlib():
    import ctypes
    import os
    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx",
                                   "libreftests.so"))
    lib.reftest_sample_plain_genz.argtypes = [
        ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_void_p]
    return lib


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("d", KFORM_DIMS)
@pytest.mark.parametrize("slices", [0, 1])
def test_eval_plain_functor(pg, gpu, plainlib, family, d, slices):
    """A user's functor with only operator() (the reference's formula of each
    family, in tests/cxx/reftests.hip) takes the kernel's plain path: every
    point x = fma(-g, h, c), f called point by point."""
    n = 1500 if d < 8 else 700
    a, u = genz_params(family, d, seed=500 + d)
    L, H = random_regions(d, n, seed=17 * d + 3)
    a = np.ascontiguousarray(a, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    est, err = np.zeros(n), np.zeros(n)
    dim = np.zeros(n, dtype=np.int32)
    rc = plainlib.reftest_sample_plain_genz(pg.FAMILY[family], d, a.ctypes.data, u.ctypes.data,
                                            n, L.ctypes.data, H.ctypes.data, slices,
                                            est.ctypes.data, err.ctypes.data, dim.ctypes.data)
    assert rc == 0
    oe, oerr, odim, fd, rs = _kform(d, family, a, u, L, H, n, plain=True)
    check_region_parity(d, family, est, err, dim, oe, oerr, odim, fd, rs, floor=True)
