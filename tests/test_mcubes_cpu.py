"""CPU pins of the m-Cubes oracle (oracle/mcubes_oracle.c).

The reference ships no m-Cubes tests; its demos (cuda/mcubes/demos/*.cu)
hold the integrands, parameters and true values.  The oracle is pinned here
by published generator vectors, by two independent restatements of the same
reference step agreeing, by exact properties of the algorithm (a constant
integrand, the inverse-CDF meaning of rebin) and statistically by the Genz
closed forms.
"""
import itertools

import numpy as np
import pytest

import oracle_api as O

NB = O.MC_BINS


def test_lcg_is_the_c_standard_sample_rand():
    # Custom_generator (vegas_utils.cuh:24-48) is the LCG of the C standard's
    # sample rand(); seeded with 1, (next/65536) % 32768 gives the published
    # sequence 16838, 5758, 10113, ...
    s, out = 1, []
    for _ in range(10):
        s = O.mc_lcg_step(s)
        out.append((s // 65536) % 32768)
    assert out == [16838, 5758, 10113, 17515, 31051, 5627, 23010, 7419, 16212, 4086]


def test_lcg_stream_values_are_state_over_2_31():
    # reseeded with the cube index: the first value of cube c is lcg(c) / 2^31
    for cube in (0, 1, 12345, 2**31 - 1, 2**32 - 1):
        v = O.mc_randoms("lcg", 0, cube, 3)
        s = cube & 0xFFFFFFFF
        for x in v:
            s = O.mc_lcg_step(s)
            assert x == s / 2.0**31


def test_splitmix64_published_vectors():
    # SplitMix64 (Steele, Lea, Flood 2014) outputs for seeds 0 and 1234567
    assert O.mc_splitmix64(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert O.mc_splitmix64(1234567, 3) == [6457827717110365317, 3203168211198807973,
                                           9817491932198370423]


def test_default_stream_is_uniform_on_the_open_unit_interval():
    key = O.mc_iter_key(0, 1)
    x = np.concatenate([O.mc_randoms("default", key, c, 200) for c in range(200)])
    assert np.all(x > 0) and np.all(x < 1)
    # cell midpoints: odd multiples of 2^-32
    k = x * 2.0**32
    assert np.all(k == np.round(k)) and np.all(k.astype(np.int64) % 2 == 1)
    n = x.size
    assert abs(x.mean() - 0.5) < 4 * np.sqrt(1 / 12 / n)
    counts, _ = np.histogram(x, bins=50, range=(0, 1))
    chi2 = np.sum((counts - n / 50) ** 2 / (n / 50))
    assert chi2 < 49 + 5 * np.sqrt(2 * 49)
    # streams of neighbouring cubes and iterations differ
    a = O.mc_randoms("default", key, 5, 8)
    b = O.mc_randoms("default", key, 6, 8)
    c = O.mc_randoms("default", O.mc_iter_key(0, 2), 5, 8)
    assert not np.any(a == b) and not np.any(a == c)


@pytest.mark.parametrize("d,ng", [(1, 7), (2, 5), (3, 4), (5, 3)])
def test_get_indx_matches_the_kernels_odometer_and_mixed_radix(d, ng):
    # get_indx (vegasT.cuh:137-156) and the per-cube odometer step
    # (:386-391) are two statements of the cube numbering; both must give
    # the mixed-radix digits (last axis fastest) plus one
    kg = [1] * d
    for m in range(ng ** d):
        digits = [int(v) + 1 for v in np.unravel_index(m, (ng,) * d)]
        assert O.mc_get_indx(m, d, ng) == digits == kg
        for k in range(d - 1, -1, -1):
            kg[k] %= ng
            kg[k] += 1
            if kg[k] != 1:
                break


def test_grid_init_is_uniform():
    xi = O.mc_grid_init(3).reshape(3, NB + 1)
    ref = np.arange(NB + 1) / NB
    assert np.all(xi[:, 0] == 0) and np.all(xi[:, NB] == 1)
    assert np.max(np.abs(xi - ref)) < 1e-13
    assert np.all(np.diff(xi, axis=1) > 0)


def test_rebin_is_the_inverse_cdf_of_the_weights():
    # rebin puts rc = sum(r)/nd of the weight in each new bin, the weight of
    # an old bin spread uniformly over it: the new edges are the inverse of
    # the piecewise-linear cumulative weight
    rng = np.random.default_rng(3)
    nd = 40
    old = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, nd - 1)), [1.0]])
    r = np.concatenate([[0.0], rng.uniform(0.1, 3.0, nd)])
    rc = r[1:].sum() / nd
    new = O.mc_rebin(rc, r, old)
    cum = np.concatenate([[0.0], np.cumsum(r[1:])])
    want = np.interp(rc * np.arange(1, nd), cum, old)
    assert np.max(np.abs(new[1:nd] - want)) < 1e-12
    assert new[nd] == 1.0
    # uniform weights leave a uniform grid in place
    uni = np.arange(nd + 1) / nd
    again = O.mc_rebin(1.0, np.concatenate([[0.0], np.ones(nd)]), uni)
    assert np.max(np.abs(again - uni)) < 1e-14


@pytest.mark.parametrize("d,ncall", [(2, 1e4), (5, 1e6), (8, 1e9), (3, 1e7)])
def test_setup_constants(d, ncall):
    s = O.mc_setup(d, ncall)
    # ng = floor((ncall/2 + 1/4)^(1/d)), found by integer search
    target = ncall / 2 + 0.25
    ng = 1
    while (ng + 1) ** d <= target:
        ng += 1
    assert s.ng == ng and s.ncubes == ng ** d
    npg = max(int(ncall / ng ** d), 2)
    assert s.npg == npg and s.calls == npg * ng ** d
    assert s.dxg == pytest.approx(NB / ng, rel=1e-15)
    assert s.xjac == pytest.approx(1 / (npg * ng ** d), rel=1e-15)
    assert s.dv2g == pytest.approx(1.0 / (npg - 1), rel=1e-13)


def test_setup_constants_of_a_demo():
    # Genz4_5D.cu at ncall 1e6: 13 strata per axis, 371293 cubes, 2 points each
    s = O.mc_setup(5, 1e6)
    assert (s.ng, s.ncubes, s.npg, s.calls) == (13, 371293.0, 2, 742586.0)


@pytest.mark.parametrize("gen", ["lcg", "default"])
def test_constant_integrand_is_exact(gen):
    # f = c: on the uniform grid every weight is xjac * prod(xo * nd) = V /
    # calls, so ti = c V up to rounding and every cube's variance is zero
    d = 3
    lows, highs = np.array([0.0, -1.0, 2.0]), np.array([1.0, 1.0, 5.0])
    c = 2.5
    s = O.mc_setup(d, 2e4, lows, highs)
    xi = O.mc_grid_init(d)
    ti, tsi, hist = O.mc_sample("const", d, O.params([c] * d), s, xi, O.mc_iter_key(0, 1), gen, d)
    V = np.prod(highs - lows)
    assert ti == pytest.approx(c * V, rel=1e-12)
    assert tsi <= s.ncubes * 1e-20
    # every sample's f^2 lands once per axis
    per = (c * V / s.calls) ** 2
    for j in range(d):
        assert hist[j * NB:(j + 1) * NB].sum() == pytest.approx(per * s.calls, rel=1e-12)


def test_histogram_totals_agree_across_axes():
    d = 4
    prm = O.params([5.0] * d, [0.3, 0.4, 0.5, 0.6])
    s = O.mc_setup(d, 5e4)
    xi = O.mc_grid_init(d)
    _, _, hist = O.mc_sample("gaussian", d, prm, s, xi, O.mc_iter_key(1, 1), "default", d)
    tot = hist.reshape(d, NB).sum(axis=1)
    assert np.max(np.abs(tot - tot[0])) <= 1e-12 * tot[0]


FAMS = [("oscillatory", [1.1, 1.3, 1.2], [0.3, 0, 0]),
        ("product_peak", [5.0, 6.0, 7.0], [0.4, 0.5, 0.6]),
        ("corner_peak", [1.0, 2.0, 3.0], None),
        ("gaussian", [8.0, 9.0, 7.0], [0.45, 0.5, 0.55]),
        ("c0", [3.0, 4.0, 5.0], [0.5, 0.5, 0.5]),
        ("discontinuous", [2.0, 3.0, 4.0], [0.7, 0.8, 0.9])]


@pytest.mark.parametrize("fam,a,u", FAMS)
@pytest.mark.parametrize("gen", ["lcg", "default"])
def test_estimates_agree_with_closed_forms(fam, a, u, gen):
    d = len(a)
    prm = O.params(a, u)
    r = O.mc_integrate(fam, d, prm, 1e-3, 1e-20, 2e4, total_iters=15, adjust_iters=10,
                       skip_iters=3, generator=gen)
    exact = O.exact(fam, d, prm)
    assert r["errorest"] > 0
    assert abs(r["estimate"] - exact) <= 5 * r["errorest"] + 1e-3 * abs(exact), (r, exact)


def test_grid_concentrates_on_the_peak():
    d = 2
    prm = O.params([30.0, 30.0], [0.5, 0.5])
    r = O.mc_integrate("gaussian", d, prm, 1e-12, 1e-30, 1e4, total_iters=6, adjust_iters=6,
                       skip_iters=0, generator="default")
    w = np.diff(r["xi"].reshape(d, NB + 1), axis=1)
    centre = w[:, NB // 2 - 5:NB // 2 + 5].mean()
    edge = np.concatenate([w[:, :10], w[:, -10:]], axis=1).mean()
    assert centre < edge / 10


def test_one_axis_grid_is_shared():
    d = 3
    prm = O.params([8.0] * d, [0.5] * d)
    r = O.mc_integrate("gaussian", d, prm, 1e-12, 1e-30, 1e4, total_iters=4, adjust_iters=4,
                       skip_iters=0, generator="default", one_axis=True)
    xi = r["xi"].reshape(d, NB + 1)
    assert np.array_equal(xi[1], xi[0]) and np.array_equal(xi[2], xi[0])
    assert not np.allclose(xi[0], np.arange(NB + 1) / NB)


def test_status_needs_five_iterations_and_skip_defers_the_estimate():
    d = 2
    prm = O.params([1.0] * d)
    # an easy integrand converges at once, but GetStatus waits for it >= 5
    r = O.mc_integrate("const", d, prm, 1e-3, 1e-20, 1e4, total_iters=20, adjust_iters=20,
                       skip_iters=0, generator="lcg", trace_cap=30)
    assert r["status"] == 0 and r["iters"] == 5
    # all iterations skipped: no estimate, status stays 1
    r = O.mc_integrate("const", d, prm, 1e-3, 1e-20, 1e4, total_iters=3, adjust_iters=3,
                       skip_iters=5, generator="lcg")
    assert r["status"] == 1 and r["estimate"] == 0 and r["iters"] == 3


def test_adjust_params_retries_with_ten_times_the_calls():
    # simple_integrate (vegasT.cuh:1082-1112): AdjustParams multiplies ncall
    # by 10 below 1e9 and reruns while the status is 1
    d = 3
    prm = O.params([20.0] * d, [0.5] * d)
    r = O.mc_integrate("gaussian", d, prm, 2e-3, 1e-30, 1e3, total_iters=8, adjust_iters=5,
                       skip_iters=2, generator="default", adjust_params=True)
    assert r["runs"] >= 2
    assert r["ncall"] == 1e3 * 10 ** (r["runs"] - 1)
    assert r["status"] == 0
