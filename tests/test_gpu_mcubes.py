"""m-Cubes on the GPU against the CPU restatement (oracle/mcubes_oracle.c)
and against the reference's demo integrals (cuda/mcubes/demos/*.cu).

Parity bar.  One sampling launch evaluates the same points as the oracle
(same streams, same point mapping), so per-cube values agree to the
integrand's rounding; the sums differ only in their order (the GPU folds
per-thread, per-workgroup, then across workgroups; the oracle sums cubes in
order) and the histograms also in the order of the LDS additions.  Measured
relative differences are ~1e-15 for ti and the histograms and ~1e-13 for
tsi, whose per-cube terms (sqrt(npg S2) - S1)(sqrt(npg S2) + S1) cancel;
the tolerances below leave two to three orders of magnitude of room.  Whole
runs feed those sums through the grid refinement, which is continuous in
them but not contractive: a grid edge that moves by e moves the sample points
by e, which changes f by about 2 a^2 |x - u| e for a Gaussian of width 1/a,
and the next refinement's logarithms and 3/2 powers pass that on.  Over 12
iterations of the steepest case (a = 25, 5D) the measured traces drift apart
by 1e-11; whole-run comparisons allow 1e-8, and the stepwise test checks
every iteration's launch at the 1e-13 bar on the grids the run adapts.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

pg = pytest.importorskip("gpuintegration_amd")
NB = O.MC_BINS
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASES = [
    ("gaussian", [25.0] * 5, [0.5] * 5),
    ("oscillatory", [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0], [0.0] * 8),
    ("product_peak", [50.0] * 6, [0.5] * 6),
    ("corner_peak", [3.0, 2.0, 1.0], None),
    ("c0", [10.0] * 8, [0.5] * 8),
    ("discontinuous", [5.0, 6.0, 7.0, 8.0, 9.0, 10.0], [0.9, 0.8, 0.7, 0.6, 0.5, 0.4]),
    ("gaussian", [10.0, 12.0], [0.3, 0.6]),
    ("product_peak", [7.0] * 7, [0.45] * 7),
    ("c0", [4.0] * 4, [0.6] * 4),
]


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


@pytest.mark.parametrize("fam,a,u", CASES)
@pytest.mark.parametrize("gen", ["lcg", "default"])
def test_sampling_pass_matches_oracle(fam, a, u, gen):
    d = len(a)
    prm = O.params(a, u)
    ncall = 3e4
    key = O.mc_iter_key(11, 3)
    # a non-uniform grid: one refinement from a peaked histogram
    xi = O.mc_grid_init(d)
    hist0 = np.tile(np.exp(-((np.arange(NB) - 200.0) / 60.0) ** 2) + 0.01, d)
    _, xi = O.mc_refine(d, d, hist0, xi)
    with pg.MCubes(d) as mc:
        ti, tsi, h = mc.sample_genz(fam, a, u, ncall, xi, key, generator=gen, hist_dims=d)
        ti0, tsi0, _ = mc.sample_genz(fam, a, u, ncall, xi, key, generator=gen, hist_dims=0)
    s = O.mc_setup(d, ncall)
    oti, otsi, oh = O.mc_sample(fam, d, prm, s, xi, key, gen, d)
    assert abs(ti - oti) <= 1e-13 * abs(oti) + 1e-300
    assert abs(tsi - otsi) <= 1e-10 * otsi
    assert np.max(np.abs(h - oh)) <= 1e-13 * np.max(oh)
    # the non-adjusting kernel draws the same points: identical sums
    assert ti0 == ti and tsi0 == tsi


@pytest.mark.parametrize("fam,a,u", CASES[:6])
@pytest.mark.parametrize("gen", ["lcg", "default"])
def test_integrate_matches_oracle(fam, a, u, gen):
    d = len(a)
    prm = O.params(a, u)
    kw = dict(ncall=2e4, total_iters=12, adjust_iters=8, skip_iters=3, generator=gen, seed=5)
    with pg.MCubes(d) as mc:
        r = mc.integrate_genz(fam, a, u, epsrel=1e-9, epsabs=1e-30, **kw)
        ti, tsi = mc.trace()
        xi = mc.grid()
    o = O.mc_integrate(fam, d, prm, 1e-9, 1e-30, kw["ncall"], kw["total_iters"],
                       kw["adjust_iters"], kw["skip_iters"], gen, seed=5, trace_cap=64)
    assert r["iters"] == o["iters"] and r["status"] == o["status"] and r["runs"] == 1
    assert _rel(ti, o["ti"]) <= 1e-8
    assert _rel(tsi, o["tsi"]) <= 1e-7
    assert np.max(np.abs(xi - o["xi"])) <= 1e-8
    assert abs(r["estimate"] - o["estimate"]) <= 1e-8 * abs(o["estimate"])
    assert abs(r["errorest"] - o["errorest"]) <= 1e-7 * o["errorest"]
    assert abs(r["chi_sq"] - o["chi_sq"]) <= 1e-5 * max(o["chi_sq"], 1e-3)
    s = O.mc_setup(d, kw["ncall"])
    assert r["neval"] == s.calls * r["iters"]


@pytest.mark.parametrize("fam,a,u", [CASES[0], CASES[2], CASES[5]])
@pytest.mark.parametrize("gen", ["lcg", "default"])
def test_stepwise_iterations_match_oracle(fam, a, u, gen):
    # every adjusting iteration's launch against the oracle's on the same
    # (adapted) grid; the next grid is refined from the GPU's histograms
    d = len(a)
    prm = O.params(a, u)
    ncall = 2e4
    s = O.mc_setup(d, ncall)
    xi = O.mc_grid_init(d)
    with pg.MCubes(d) as mc:
        for it in range(1, 9):
            key = O.mc_iter_key(5, it)
            ti, tsi, h = mc.sample_genz(fam, a, u, ncall, xi, key, generator=gen, hist_dims=d)
            oti, otsi, oh = O.mc_sample(fam, d, prm, s, xi, key, gen, d)
            assert abs(ti - oti) <= 1e-13 * abs(oti), it
            assert abs(tsi - otsi) <= 1e-10 * otsi, it
            assert np.max(np.abs(h - oh)) <= 1e-13 * np.max(oh), it
            _, xi = O.mc_refine(d, d, h, xi)


@pytest.mark.parametrize("one_axis,lows,highs", [
    (False, [-1.0, 0.0, 0.5, -2.0], [1.0, 0.5, 1.0, 2.0]),
    (True, [-1.0] * 4, [1.0] * 4),
])
def test_one_axis_and_volume_match_oracle(one_axis, lows, highs):
    # the one-axis grid (vegasT1D.cuh) is meant for integrands alike on every
    # axis: a shared grid on axes whose peaks sit elsewhere starves them of
    # samples, and such runs are too ill-conditioned to compare
    d = 4
    a, u = [9.0] * d, [0.25] * d
    prm = O.params(a, u)
    with pg.MCubes(d) as mc:
        r = mc.integrate_genz("gaussian", a, u, epsrel=1e-9, epsabs=1e-30, ncall=2e4,
                              total_iters=10, adjust_iters=6, skip_iters=2, generator="default",
                              one_axis=one_axis, lows=lows, highs=highs)
        ti, _ = mc.trace()
        xi = mc.grid().reshape(d, NB + 1)
    o = O.mc_integrate("gaussian", d, prm, 1e-9, 1e-30, 2e4, 10, 6, 2, "default",
                       one_axis=one_axis, lows=lows, highs=highs, trace_cap=32)
    assert _rel(ti, o["ti"]) <= 1e-8
    if one_axis:
        assert all(np.array_equal(xi[j], xi[0]) for j in range(d))
    assert np.max(np.abs(xi.ravel() - o["xi"])) <= 1e-8
    assert abs(r["estimate"] - o["estimate"]) <= 1e-8 * abs(o["estimate"])
    if one_axis:
        # closed form on [-1, 1]^4: (sqrt(pi)/(2a) (erf(a(1-u)) + erf(a(1+u))))^4
        from math import erf, pi, sqrt
        exact = (sqrt(pi) / (2 * a[0]) * (erf(a[0] * (1 - u[0])) + erf(a[0] * (1 + u[0])))) ** d
        assert abs(r["estimate"] - exact) <= 5 * r["errorest"]


def test_adjust_params_matches_oracle():
    d = 3
    a, u = [20.0] * d, [0.5] * d
    prm = O.params(a, u)
    with pg.MCubes(d) as mc:
        r = mc.integrate_genz("gaussian", a, u, epsrel=2e-3, epsabs=1e-30, ncall=1e3,
                              total_iters=8, adjust_iters=5, skip_iters=2, generator="default",
                              adjust_params=True)
    o = O.mc_integrate("gaussian", d, prm, 2e-3, 1e-30, 1e3, 8, 5, 2, "default",
                       adjust_params=True)
    assert (r["runs"], r["ncall"], r["iters"], r["status"]) == (o["runs"], o["ncall"],
                                                                o["iters"], o["status"])
    assert abs(r["estimate"] - o["estimate"]) <= 1e-8 * abs(o["estimate"])


def test_full_size_8d_properties():
    # 1e9 calls per iteration (429 981 696 cubes, 2 points each): the sums
    # of a launch do not depend on anything but the streams, and every
    # axis' histogram holds the same total (each sample's f^2 once per axis)
    d = 8
    a, u = [10.0] * d, [0.5] * d
    xi = O.mc_grid_init(d)
    key = O.mc_iter_key(3, 1)
    with pg.MCubes(d) as mc:
        t1, s1, h = mc.sample_genz("gaussian", a, u, 1e9, xi, key, generator="default",
                                   hist_dims=d)
        t2, s2, _ = mc.sample_genz("gaussian", a, u, 1e9, xi, key, generator="default",
                                   hist_dims=0)
    assert t1 == t2 and s1 == s2
    tot = h.reshape(d, NB).sum(axis=1)
    assert np.max(np.abs(tot - tot[0])) <= 1e-10 * tot[0]
    # the uniform-grid estimate of the 8D Gaussian (a=10, u=.5)
    exact = O.exact("gaussian", d, O.params(a, u))
    s = O.mc_setup(d, 1e9)
    sd = np.sqrt(s1 * s.dv2g)
    assert abs(t1 - exact) <= 6 * sd


def test_errors():
    with pytest.raises(pg.PaganiError):
        pg.MCubes(0)
    with pytest.raises(pg.PaganiError):
        pg.MCubes(17)
    with pg.MCubes(3) as mc:
        with pytest.raises(pg.PaganiError):
            mc.integrate_genz("gaussian", [1.0] * 4, [0.5] * 4)  # ndim mismatch
        with pytest.raises(pg.PaganiError):
            mc.integrate_genz("gaussian", [1.0] * 3, [0.5] * 3, ncall=0.5)
        with pytest.raises(pg.PaganiError):
            mc.integrate_genz("gaussian", [1.0] * 3, [0.5] * 3, generator=7)
        with pytest.raises(pg.PaganiError):
            mc.integrate_genz("gaussian", [1.0] * 3, [0.5] * 3, lows=[0, 0, 1], highs=[1, 1, 1])
        # still usable after the errors
        r = mc.integrate_genz("gaussian", [1.0] * 3, [0.5] * 3, ncall=1e4)
        assert r["status"] == 0


# --- the reference's C++ surface (include/pagani/mcubes.hpp) -----------------

def _reftests():
    lib = ctypes.CDLL(os.path.join(ROOT, "tests", "cxx", "libreftests.so"))
    lib.reftest_mcubes.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(pg._Result)]
    return lib


# (name, generator, mode, epsrel, ncall, titer, itmax, skip, true value): the
# demos' settings (Genz*_*.cu, SinSum6D.cu, Gauss9D.cu), at the ncall their
# required_ncall tables give for epsrel 1e-3
DEMOS = [
    ("F_4_5D", 1, 0, 1e-3, 1e7, 100, 20, 5, 1.79132603674879e-06),
    ("F_2_6D", 1, 0, 1e-3, 1e7, 100, 40, 10, 1.286889807581113e13),
    ("F_3_3D", 1, 0, 1e-3, 1e7, 100, 20, 0, 0.010846560846560846561),
    ("F_5_8D", 1, 0, 1e-3, 1e6, 100, 20, 5, 2.425217625641885e-06),
    ("F_6_6D", 1, 0, 1e-3, 1e7, 100, 20, 5, 1.5477367885091207413e8),
    # SinSum6D.cu prints its result without a convergence check: sin(sum x)
    # over [0, 10]^6 does not reach 1e-3 in 10 iterations of 2e9 calls
    ("SinSum6D", 0, 0, 0.0, 2e9, 10, 0, 0, -49.165073),
    ("Gauss9D", 0, 0, 1e-3, 1e8, 15, 10, 10, 1.0),
    ("F_4_5D", 0, 1, 1e-3, 1e6, 100, 20, 5, 1.79132603674879e-06),
    ("XY", 0, 3, 1e-3, 1e6, 70, 20, 0, 0.25),
    ("XY", 1, 4, 1e-3, 1e6, 15, 10, 5, 0.25),
    ("F_4_5D", 0, 2, 1e-3, 1e6, 100, 20, 5, 1.79132603674879e-06),
    # the dimension range beyond the demos: 1D, and 12D / 16D, whose LDS
    # (grid + histograms, 96 KB / 128 KB) exceeds the default 64 KB
    ("Gauss1D", 0, 0, 1e-4, 1e5, 20, 10, 2, 0.25032445820538396),
    ("CosProd12", 0, 0, 1e-3, 1e7, 20, 10, 3, 0.8414709848078965 ** 12),
    ("CosProd16", 1, 0, 1e-3, 1e7, 20, 10, 3, 0.8414709848078965 ** 16),
]


@pytest.mark.parametrize("name,gen,mode,eps,ncall,titer,itmax,skip,truth", DEMOS)
def test_reference_demos_through_cpp_surface(name, gen, mode, eps, ncall, titer, itmax, skip,
                                             truth):
    lib = _reftests()
    r = pg._Result()
    rc = lib.reftest_mcubes(name.encode(), gen, mode, eps, ncall, titer, itmax, skip,
                            ctypes.byref(r))
    assert rc == 0
    if eps > 0:
        assert r.status == 0, (r.estimate, r.errorest)
        assert r.errorest <= eps * abs(r.estimate)
        assert r.iters >= 5
    else:
        assert r.status == 1 and r.iters == titer
    assert abs(r.estimate - truth) <= 5 * r.errorest + 2e-4 * abs(truth), (r.estimate, truth)
