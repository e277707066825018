"""CPU tests of the oracle (oracle/pagani_oracle.c) -- what pins it.

Each function of the oracle is checked against something other than its own
formula: the reference's known answers (its tests' constants and fixtures),
its compiled closed forms (oracle/_ref, when built), mathematical invariants
of the degree-9 rule, brute-force quadrature, and hand-worked cases.
"""
import itertools
import math

import numpy as np
import pytest

import oracle_api as O

# ---------------------------------------------------------------------------
# rule tables (Rule.cuh)


def test_feval_counts():
    # quad.h:68-75; 1105 points per region at 8D
    assert [O.feval(d) for d in range(2, 9)] == [33, 77, 153, 273, 453, 717, 1105]


def _monomials(d, deg):
    for c in itertools.combinations_with_replacement(range(d), deg):
        k = np.zeros(d, dtype=int)
        for i in c:
            k[i] += 1
        yield k


def _centred_exact(k):
    """integral of x^k over [-1/2, 1/2]^d"""
    return np.prod([0.0 if ki % 2 else 2 * 0.5 ** (ki + 1) / (ki + 1) for ki in k])


def _exact_degree(W, S, G, r, d, maxdeg=10):
    X = -G  # points of the rule on the centred unit cube
    w = W[S, r]
    best = -1
    for deg in range(maxdeg + 1):
        worst = 0.0
        for k in _monomials(d, deg):
            v = np.prod(X ** k, axis=1)
            ex = _centred_exact(k) if r == 0 else 0.0
            worst = max(worst, abs((w * v).sum() - ex) / max(np.abs(w * v).sum(), 1e-300))
        if worst < 1e-12:
            best = deg
        else:
            break
    return best


@pytest.mark.parametrize("d", [2, 3, 4, 5, 6])
def test_rule_degrees(d):
    """Column 0 is a degree-9 rule (exact through 9, not 10); null rules 1, 2
    annihilate degree <= 7, null rule 3 degree <= 5, null rule 4 degree <= 3."""
    W, lam, sc, nm, cnt = O.rule_tables(d)
    G, S = O.generators(d)
    assert [_exact_degree(W, S, G, r, d) for r in range(5)] == [9, 7, 7, 5, 3]


def test_rule_degree_8d_spot():
    """8D: degree 9 exact on random monomials, degree 10 (x0^10) not."""
    d = 8
    L, H = np.zeros(d), np.ones(d)
    rng = np.random.default_rng(0)
    for deg in range(10):
        k = np.zeros(d)
        for _ in range(deg):
            k[rng.integers(d)] += 1
        prm = O.params(k)
        est, *_ = O.eval_regions(d, "monomial", prm, L, H, 1)
        assert abs(est[0] - O.exact("monomial", d, prm)) <= 1e-13
    prm = O.params([10] + [0] * 7)
    est, *_ = O.eval_regions(d, "monomial", prm, L, H, 1)
    assert abs(est[0] - 1 / 11) > 1e-9  # (3.7e-7 here; rounding is ~1e-16)


def test_planted_weight_error_is_caught():
    """The degree check fails when one weight is perturbed (checks the checker)."""
    d = 3
    W, lam, sc, nm, cnt = O.rule_tables(d)
    G, S = O.generators(d)
    W2 = W.copy()
    W2[5, 0] *= 1 + 1e-6
    assert _exact_degree(W2, S, G, 0, d) < 9


@pytest.mark.parametrize("d", range(2, 9))
def test_generator_orbits(d):
    """Every point is a signed permutation of its set's seed (Rule.cuh:217-249);
    set sizes are the reference's counts; the first 1+4d points are the centre
    and the +-lambda1, +-lambda2 axis points in the fourth difference's order."""
    W, lam, sc, nm, cnt = O.rule_tables(d)
    G, S = O.generators(d)
    seeds = {0: [], 1: [lam[0]], 2: [lam[1]], 3: [lam[2]], 4: [lam[3]], 5: [lam[0]] * 2,
             6: sorted([lam[0], lam[1]]), 7: [lam[0]] * 3, 8: [lam[4]] * d}
    assert list(np.bincount(S, minlength=9)) == list(cnt)
    for p in range(len(S)):
        nz = sorted(abs(G[p][G[p] != 0]))
        assert np.allclose(nz, sorted(seeds[S[p]]), rtol=0, atol=0)
    assert len({tuple(g) for g in G}) == len(G)  # no duplicate points
    for j in range(d):
        e = np.zeros(d)
        e[j] = lam[0]
        assert np.array_equal(G[1 + 2 * j], e) and np.array_equal(G[2 + 2 * j], -e)
        e[j] = lam[1]
        assert np.array_equal(G[1 + 2 * d + 2 * j], e) and np.array_equal(G[2 + 2 * d + 2 * j], -e)


def test_weights_sum_to_one():
    for d in range(2, 9):
        W, lam, sc, nm, cnt = O.rule_tables(d)
        assert abs((cnt * W[:, 0]).sum() - 1.0) < 1e-14


# ---------------------------------------------------------------------------
# integrands and closed forms (compute_genz_integrals.cuh, integrands.cuh)

REF_CONSTANTS = [
    # family, d, a, u, value, rel tolerance (digits given by the reference)
    ("oscillatory", 5, list(range(1, 6)), None, 0.020242422119901896863, 1e-14),
    ("oscillatory", 6, list(range(1, 7)), None, -0.0013062949651908022873, 1e-14),
    ("oscillatory", 7, list(range(1, 8)), None, -0.00003764562579508779, 1e-14),
    ("oscillatory", 8, list(range(1, 9)), None, 3.439557952183252e-05, 1e-14),
    ("corner_peak", 5, [3, 2, 1, 4, 5], None, 0.00002602538279621613, 1e-14),
    ("corner_peak", 6, [3, 2, 1, 4, 5, 6], None, 7.1790160638199853886e-7, 1e-14),
    ("corner_peak", 7, [3, 2, 1, 4, 5, 6, 7], None, 1.459429e-08, 1e-6),
    ("corner_peak", 8, [3, 2, 1, 4, 5, 6, 7, 8], None, 2.2751965817917756076e-10, 1e-14),
    ("gaussian", 5, [25.0] * 5, [0.5] * 5, 1.79132603674879e-06, 1e-13),
    ("c0", 8, [10.0] * 8, [0.5] * 8, 2.425217625641885e-06, 1e-14),
    ("discontinuous", 6, [10, 9, 8, 7, 6, 5], [.9, .8, .7, .6, .5, .4], 1.5477367885091207413e8,
     1e-14),
    # Easy_Integrals.cu's F_2_6D constant is 1.4e-6 off the closed form (its
    # test tolerance is 1e-3)
    ("product_peak", 6, [50.0] * 6, [0.5] * 6, 1.286889807581113e13, 2e-6),
]


@pytest.mark.parametrize("case", REF_CONSTANTS, ids=lambda c: f"{c[0]}-{c[1]}D")
def test_closed_forms_vs_reference_constants(case):
    fam, d, a, u, val, tol = case
    got = O.exact(fam, d, O.params(a, u))
    assert abs(got - val) <= tol * abs(val)


def test_closed_forms_vs_reference_code():
    """The reference's own closed forms, compiled from its source into oracle/_ref."""
    ref = O.ref_closed_form_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(42)
    for fam in ["product_peak", "gaussian", "c0", "discontinuous"]:
        for d in range(2, 9):
            a = 5 + 10 * rng.random(d)
            u = rng.random(d)
            A, U = np.ascontiguousarray(a), np.ascontiguousarray(u)
            r = ref.ref_genz_closed_form(O.FAM[fam], d, O.dptr(A), O.dptr(U))
            got = O.exact(fam, d, O.params(a, u))
            # the reference's Gaussian uses genz_phi, documented to 1e-7
            tol = 2e-7 if fam == "gaussian" else 1e-13
            assert abs(got - r) <= tol * abs(r), (fam, d, got, r)


def _gauss_legendre_box(f, d, lo, hi, n=24):
    x, w = np.polynomial.legendre.leggauss(n)
    grids = [0.5 * (h - l) * x + 0.5 * (h + l) for l, h in zip(lo, hi)]
    wts = [0.5 * (h - l) * w for l, h in zip(lo, hi)]
    total = 0.0
    for idx in itertools.product(range(n), repeat=d):
        pt = np.array([grids[j][i] for j, i in enumerate(idx)])
        total += np.prod([wts[j][i] for j, i in enumerate(idx)]) * f(pt)
    return total


@pytest.mark.parametrize("fam", ["oscillatory", "product_peak", "corner_peak", "gaussian", "c0",
                                 "discontinuous"])
def test_closed_forms_vs_brute_force(fam):
    """2D closed forms vs tensor Gauss-Legendre on the pieces where f is smooth."""
    d = 2
    a = [2.0, 3.0]
    u = [0.4, 0.7]
    prm = O.params(a, u)
    f = lambda x: O.integrand(fam, d, prm, x)
    cuts = [[0.0, 1.0], [0.0, 1.0]]
    if fam in ("c0", "discontinuous"):
        cuts = [[0.0, u[0], 1.0], [0.0, u[1], 1.0]]
    total = 0.0
    for i in range(len(cuts[0]) - 1):
        for j in range(len(cuts[1]) - 1):
            total += _gauss_legendre_box(f, d, [cuts[0][i], cuts[1][j]],
                                         [cuts[0][i + 1], cuts[1][j + 1]])
    assert abs(total - O.exact(fam, d, prm)) <= 1e-12 * max(1.0, abs(total))


def test_integrand_reference_forms():
    """Spot values of the reference's fixed functors (integrands.cuh)."""
    x = np.array([0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8])
    # F_1_8D: cos(s + 2t + ... + 8z)
    assert math.isclose(O.integrand("oscillatory", 8, O.params(range(1, 9)), x),
                        math.cos(sum((i + 1) * x[i] for i in range(8))), rel_tol=1e-15)
    # F_4: exp(-625 sum (x - .5)^2)
    assert math.isclose(O.integrand("gaussian", 8, O.params([25.0] * 8, [0.5] * 8), x),
                        math.exp(-625 * sum((xi - 0.5) ** 2 for xi in x)), rel_tol=1e-13)
    # F_6_6D with arguments (y, v, x, w, z, u): zero past the steps
    prm = O.params([10, 9, 8, 7, 6, 5], [.9, .8, .7, .6, .5, .4])
    assert O.integrand("discontinuous", 6, prm, np.array([.1, .1, .1, .1, .1, .45])) == 0.0


# ---------------------------------------------------------------------------
# per-region evaluation: hand-worked cases


def test_constant_region_sampling_kat():
    """RegionSampling.cu:48-172: constant c over k^2 splits gives c/n per region
    (1e-6); the null rules vanish to rounding.  The split axis is the first
    axis with the largest fourth difference (Sample.cuh:194-218; for c != 0
    every axis carries the same rounding residue, so axis 0), and the widest
    axis, upper - lower as in Phases.cuh:331-346, when all of them are 0."""
    for c in (15.37, -15.37, 0.0):
        for k in range(5, 15):
            L, H, n = O.uniform_split(2, k)
            est, err, dim, fd, rs = O.eval_regions(2, "const", O.params([c, 0.0]), L, H, n,
                                                   want_extra=True)
            assert np.allclose(est, c / n, rtol=1e-6, atol=0)
            assert np.all(err <= 1e-14 * abs(c) / n + 1e-300)
            assert np.all(fd <= 1e-14 * abs(c))
            Lm, Hm = L.reshape(2, n), H.reshape(2, n)
            widths = (Lm + Hm) - Lm
            want = np.where(fd.max(axis=1) > 0, np.argmax(fd, axis=1),
                            np.argmax(widths, axis=0))
            assert np.array_equal(dim, want)


def test_split_axis_of_axis_quartic():
    """f = x_j^4 varies along axis j only: the fourth difference picks j."""
    d = 5
    L, H = np.zeros(d), np.ones(d)
    for j in range(d):
        k = np.zeros(d)
        k[j] = 4
        _, _, dim = O.eval_regions(d, "monomial", O.params(k), L, H, 1)
        assert dim[0] == j


def test_null_rule_errors_of_a_quartic():
    """A degree-4 monomial: null rules 1-3 vanish (exact through degree 5),
    null rule 4 does not, so e1 and e2 are rounding noise and e3 > 0.  The
    error formula (Sample.cuh:282-288) then compares noise (5 e1 <= e2) and
    returns either vol*e1 (~0) or 5*vol*e3."""
    d = 3
    L, H = np.zeros(d), np.ones(d)
    k = np.array([2.0, 2.0, 0.0])
    est, err, dim, fd, rs = O.eval_regions(d, "monomial", O.params(k), L, H, 1, want_extra=True)
    s = rs[0, :5]
    A = rs[0, 5:10]
    assert abs(s[1]) <= 1e-14 * A[1] and abs(s[2]) <= 1e-14 * A[2]
    assert abs(s[3]) <= 1e-14 * A[3]
    assert abs(s[4]) > 1e-6 * A[4]
    e1, e2, e3 = rs[0, 11:14]
    assert e1 <= 1e-14 and e2 <= 1e-14 and e3 > 1e-6
    branch_a = 5 * e1 <= e2 and 5 * e2 <= e3
    assert err[0] == (e1 if branch_a else 5 * max(e1, e2, e3))


def test_volume_and_jacobian():
    """The Jacobian of a general box scales estimates (Phases.cuh:323-349)."""
    d = 3
    L, H, n = O.uniform_split(d, 2)
    lows, highs = np.array([-1.0, 0.0, 2.0]), np.array([1.0, 3.0, 2.5])
    est, err, dim = O.eval_regions(d, "const", O.params([2.0, 0, 0]), L, H, n, lows, highs)
    assert math.isclose(est.sum(), 2.0 * 2.0 * 3.0 * 0.5, rel_tol=1e-14)


# ---------------------------------------------------------------------------
# region-list phases


@pytest.mark.parametrize("d,k", [(3, 2), (5, 3), (8, 4)])
def test_uniform_split_kat(d, k):
    """Uniform_sub_division.cu:23-95: volumes 1/n summing to 1."""
    L, H, n = O.uniform_split(d, k)
    Lm, Hm = L.reshape(d, n), H.reshape(d, n)
    vol = np.prod(Hm, axis=0)
    assert np.allclose(vol, 1.0 / n, rtol=1e-12)
    assert math.isclose(vol.sum(), 1.0, rel_tol=1e-12)
    assert np.all(Lm >= 0) and np.all(Hm < 1) and np.all(Hm > 0)
    # every box distinct and inside the cube
    assert len({tuple(Lm[:, t]) for t in range(n)}) == n
    assert np.all(Lm + Hm <= 1.0 + 1e-15)


def test_refine_hand_case():
    """RefineError (Phases.cuh:219-268) on four regions, values worked by hand."""
    est = np.array([1.0, 2.0, 1.5, 2.5])
    err = np.array([0.1, 0.2, 0.3, 0.0])
    pest = np.array([2.0, 4.0])
    err2, act = O.refine(est, err, pest, epsrel=1e-3)
    # pair (0, 2): diff = |.25 (1 + 1.5 - 2)| = .125, E = .4, c = 1 + 2*.125/.4
    c = 1 + 2 * 0.125 / 0.4
    assert math.isclose(err2[0], 0.1 * c + 0.125, rel_tol=1e-15)
    assert math.isclose(err2[2], 0.3 * c + 0.125, rel_tol=1e-15)
    # pair (1, 3): diff = |.25 (2 + 2.5 - 4)| = .125, E = .2
    c = 1 + 2 * 0.125 / 0.2
    assert math.isclose(err2[1], 0.2 * c + 0.125, rel_tol=1e-15)
    assert math.isclose(err2[3], 0.0 * c + 0.125, rel_tol=1e-15)
    assert list(act) == [1, 1, 1, 1]
    # consistent children: no refinement beyond the raw error; tiny errors finish
    err2, act = O.refine(np.array([1.0, 1.0]), np.array([1e-9, 1e-3]), np.array([2.0]),
                         epsrel=1e-6)
    assert math.isclose(err2[0], 1e-9, rel_tol=1e-15) and list(act) == [0, 1]
    # relerr classification off: everything stays active
    _, act = O.refine(np.array([1.0, 1.0]), np.array([1e-9, 1e-9]), np.array([2.0]), 1e-6,
                      relerr=False)
    assert list(act) == [1, 1]


def _filter_case(active):
    d = 2
    L, H, n = O.uniform_split(d, 10)
    est = np.full(n, 1000.0)
    err = np.full(n, 10.0)
    dim = np.zeros(n, dtype=np.int32)
    for t in np.where(active == 0)[0]:
        est[t], err[t], dim[t] = 999.0, 1.0, 1
    return d, n, L, H, est, err, dim


@pytest.mark.parametrize("inactive", [[], [15], [0, 99], list(range(100))])
def test_filter_kat(inactive):
    """RegionFiltering.cu:42-494: active count, order-preserving compaction,
    parents get the pre-filter estimates, split axes carried."""
    active = np.ones(100, dtype=np.uint8)
    active[inactive] = 0
    d, n, L, H, est, err, dim = _filter_case(active)
    m, L2, H2, pe, pr, d2 = O.filter_regions(d, n, L, H, active, est, err, dim)
    assert m == n - len(inactive)
    if m == 0:
        return
    keep = np.where(active == 1)[0]
    assert np.array_equal(L2.reshape(d, m), L.reshape(d, n)[:, keep])
    assert np.array_equal(H2.reshape(d, m), H.reshape(d, n)[:, keep])
    assert np.all(pe == 1000.0) and np.all(pr == 10.0) and np.all(d2 == 0)


def test_split_kat():
    """RegionSplitting.cu:41-203: children at [i] and [m + i] with stride 2m;
    only the split axis is halved."""
    d = 2
    L, H, m = O.uniform_split(d, 5)
    dim = np.ones(m, dtype=np.int32)
    dim[0] = 0
    L2, H2 = O.split_regions(d, m, L, H, dim)
    Lp, Hp = L.reshape(d, m), H.reshape(d, m)
    Lc, Hc = L2.reshape(d, 2 * m), H2.reshape(d, 2 * m)
    for i in range(m):
        b = dim[i]
        for j in range(d):
            if j == b:
                assert Lc[j, i] == Lp[j, i] and Lc[j, m + i] == Lp[j, i] + Hp[j, i] / 2
                assert Hc[j, i] == Hp[j, i] / 2 == Hc[j, m + i]
            else:
                assert Lc[j, i] == Lc[j, m + i] == Lp[j, i]
                assert Hc[j, i] == Hc[j, m + i] == Hp[j, i]


def test_finished_estimates_kat():
    """finished_estimates.cpp (kokkos tests): 100 regions of 3.2 +- 1e-5,
    active 11..17 -> finished = total - 7 regions' worth."""
    n = 100
    est = np.full(n, 3.2)
    err = np.full(n, 1e-5)
    act = np.zeros(n, dtype=np.uint8)
    act[11:18] = 1
    fin_est = est.sum() - (act * est).sum()
    fin_err = err.sum() - (act * err).sum()
    assert math.isclose(fin_est, 3.2 * 100 - 3.2 * 7, rel_tol=1e-12)
    assert math.isclose(fin_err, 1e-5 * 100 - 1e-5 * 7, rel_tol=1e-9)


# ---------------------------------------------------------------------------
# heuristic classifier (heuristic_classifier.cuh)


def test_memory_model():
    # 8 (10 d + 8) + 16 bytes per region: 720 at 8D
    assert O.LIB.orc_mem_required(1, 8) == 720
    assert O.LIB.orc_mem_required(1000, 3) == 1000 * (8 * 38 + 16)
    assert O.LIB.orc_free_mem(10, 8, 100) == 1  # never 0
    assert O.LIB.orc_free_mem(10, 8, int(16e9)) == int(16e9) - 7200


def _clf(d=3, epsrel=1e-3, total=int(16e9)):
    c = O.OrcClassifier()
    O.LIB.orc_clf_init(c, d, epsrel, 1e-12, total)
    return c


def test_estimate_converged():
    c = _clf(epsrel=1e-3)
    assert c.required_digits == 3
    O.LIB.orc_clf_store(c, 7000.0)
    assert not O.LIB.orc_clf_estimate_converged(c)  # one estimate is not enough
    for _ in range(20):
        O.LIB.orc_clf_store(c, 7000.0)
    assert O.LIB.orc_clf_estimate_converged(c)
    c = _clf(epsrel=1e-3)
    for v in (1.2345e-6, 1.2349e-6, 1.2341e-6):  # 1.234 in all three
        O.LIB.orc_clf_store(c, v)
    assert O.LIB.orc_clf_sig_digits_same(c)
    O.LIB.orc_clf_store(c, 1.2356e-6)  # "1.23" still agrees: 3 digits compared
    assert O.LIB.orc_clf_sig_digits_same(c)
    O.LIB.orc_clf_store(c, 1.2456e-6)  # "1.24" breaks the third digit
    assert not O.LIB.orc_clf_sig_digits_same(c)
    c = _clf(epsrel=1e-6)
    assert c.required_digits == 6


def test_classifier_kat():
    """test_heuristic_classifier.cpp:32-104: errors {.075, .99, .079, 101.96,
    101.33, 1.93, 101.99}, estimates 1000 each, finished error 4.2,
    epsrel 1e-3: classification passes memory and budget, saves (30%, 50%]
    of the regions, and had to raise the budget and active percentages."""
    c = _clf(d=3, epsrel=1e-3)
    for _ in range(20):
        O.LIB.orc_clf_store(c, 7000.0)
    errs = np.array([.075, .99, .079, 101.96, 101.33, 1.93, 101.99])
    flags = np.zeros(7, dtype=np.uint8)
    res = O.OrcClassRes()
    O.LIB.orc_clf_classify(c, O.dptr(errs), 7, errs.sum(), 0.0, 4.2, O.uptr(flags),
                           res)
    saved = 1 - res.num_active / 7
    assert res.pass_mem and res.pass_errorest_budget
    assert 0.3 < saved <= 0.5
    assert res.max_budget_perc_to_cover > 0.25
    assert res.max_active_perc > 0.5
    assert int(flags.sum()) == res.num_active


def test_criteria_met():
    c = _clf(d=8, total=int(16e9))
    # the loop stores an estimate before every check (Workspace.cuh:457-463)
    O.LIB.orc_clf_store(c, 1.0)
    # ratio > 1 always classifies; ratio in (.1, 1] needs a converged estimate
    big = int(12e6)
    assert O.LIB.orc_clf_criteria_met(c, big)
    mid = int(3e6)
    assert not O.LIB.orc_clf_criteria_met(c, mid)
    for _ in range(3):
        O.LIB.orc_clf_store(c, 1.0)
    assert O.LIB.orc_clf_criteria_met(c, mid)
    assert not O.LIB.orc_clf_criteria_met(c, 1000)


# ---------------------------------------------------------------------------
# the full loop (Workspace.cuh) on the reference's known answers

INTEGRATION_KATS = [
    ("const", 2, [15.37, 0.0], None, 15.37, 1e-3),
    ("product", 2, [0.0, 0.0], None, 0.25, 1e-6),
    ("corner_peak", 8, [3, 2, 1, 4, 5, 6, 7, 8], None, 2.2751965817917756076e-10, 1e-3),
    ("discontinuous", 6, [10, 9, 8, 7, 6, 5], [.9, .8, .7, .6, .5, .4], 1.5477367885091207413e8,
     1e-3),
    ("box", 8, [15.0] + [0.0] * 7, None, 8879.851175413485, 1e-3),
    ("box", 8, [25.0] + [0.0] * 7, None, 14996089.096112404019, 1e-3),
    ("box", 8, [22.0] + [0.0] * 7, None, 1495369.283757217694, 1e-3),
]


@pytest.mark.parametrize("case", INTEGRATION_KATS, ids=lambda c: f"{c[0]}-{c[1]}D")
def test_integrate_kats(case):
    fam, d, a, u, true, eps = case
    r = O.integrate(d, fam, O.params(a, u), eps, 1e-12)
    assert r.status == 0
    assert abs(r.estimate - true) <= eps * abs(true)


@pytest.mark.slow
@pytest.mark.parametrize("case", [
    ("product_peak", 6, [50.0] * 6, [0.5] * 6, 1.286889807581113e13),
    ("gaussian", 5, [25.0] * 5, [0.5] * 5, 1.79132603674879e-06),
    ("c0", 8, [10.0] * 8, [0.5] * 8, 2.425217625641885e-06),
], ids=["F_2_6D", "F_4_5D", "F_5_8D"])
def test_integrate_easy_integrals_slow(case):
    """Easy_Integrals.cu:74-134 (the oracle needs a few million regions)."""
    fam, d, a, u, true = case
    r = O.integrate(d, fam, O.params(a, u), 1e-3, 1e-12)
    assert r.status == 0 and abs(r.estimate - true) <= 1e-3 * abs(true)


def test_integrate_nan():
    """NanIntegral.cu:16-37: NaN propagates to the estimate."""
    r = O.integrate(2, "nan", O.params([0.0, 0.0]), 1e-3, 1e-12)
    assert math.isnan(r.estimate)


def test_accuracy_improves_with_epsrel():
    """accuracy_improves_with_epsrel.cu:38-67 (Fun6, 6D, integral 1)."""
    epsrel, prev = 1e-3, 1.0
    while epsrel > 1e-5:
        r = O.integrate(6, "fun6", O.params([0.0] * 6), epsrel, 1e-40)
        assert r.status == 0 and r.errorest / r.estimate <= epsrel
        assert r.errorest <= prev and abs(r.estimate - 1.0) <= 3 * r.errorest
        prev = r.errorest
        epsrel /= 2.0


def test_memory_termination():
    """A tiny memory budget forces classification, then termination with
    status 1 and the leaves summed (Workspace.cuh:294-314)."""
    prm = O.params([25.0] * 5, [0.5] * 5)
    r, tr = O.integrate(5, "gaussian", prm, 1e-8, 1e-20, total_mem=2_000_000, trace_cap=100)
    assert r.status == 1
    assert tr["m"][-1] == 0
    exact = O.exact("gaussian", 5, prm)
    assert abs(r.estimate - exact) <= 0.05 * exact


def test_iteration_bookkeeping():
    """nregions counts every finished region plus the last list; the
    Sub_regions& overload counts iterations before the convergence check."""
    prm = O.params([3, 2, 1, 4, 5], None)
    r1, tr = O.integrate(5, "corner_peak", prm, 1e-4, 1e-20, trace_cap=100)
    assert r1.nregions == sum(int(n - m) for n, m in zip(tr["n"], tr["m"]))
    assert r1.iters == len(tr["n"]) - 1
    r2 = O.integrate(5, "corner_peak", prm, 1e-4, 1e-20, user_regions=True)
    assert r2.iters == len(tr["n"]) and r2.estimate == r1.estimate


def test_cuhre_sequential_baseline():
    """The CPU baseline (sequential Cuhre-equivalent) on the 5D oscillatory
    workload at epsrel 1e-3, max 1e6 evaluations."""
    from gpuintegration_amd import workloads as W
    fam, a, u = W.oscillatory_5d(2024)
    prm = O.params(a, u)
    r = O.cuhre_seq(5, fam, prm, 1e-3, 1e-20, 1_000_000)
    exact = O.exact(fam, 5, prm)
    assert r.status == 0 and r.neval <= 1_000_000
    assert abs(r.estimate - exact) <= max(r.errorest, 1e-3 * abs(exact))


def test_mt19937_64_standard_value():
    """[rand.predef]: the 10000th output of a default-seeded mt19937_64."""
    from gpuintegration_amd.workloads import MT19937_64
    g = MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


# ---------------------------------------------------------------------------
# Evaluation at the library kernel's own points (orc_eval_regions_kform)

def test_kform_rule_exactness():
    """At any point formation the rule is exact through degree 9: the
    kernel-point mode integrates random monomials over a sub-box exactly
    (plain form: x = fma(-g, h, c)), and fails at degree 10 as the
    reference-map mode does."""
    rng = np.random.default_rng(5)
    d = 4
    L = np.array([0.1, 0.2, 0.05, 0.3])
    H = np.array([0.5, 0.25, 0.7, 0.4])
    for _ in range(20):
        k = rng.multinomial(int(rng.integers(0, 10)), [1 / d] * d).astype(float)
        prm = O.params(k)
        exact = np.prod([((L[j] + H[j]) ** (k[j] + 1) - L[j] ** (k[j] + 1)) / (k[j] + 1)
                         for j in range(d)])
        est, *_ = O.eval_regions_kform(d, "monomial", prm, L, H, 1, plain=True)
        assert abs(est[0] - exact) <= 1e-15 * max(abs(exact), 1e-300) + 1e-18
    k = np.array([10.0, 0, 0, 0])
    exact = ((L[0] + H[0]) ** 11 - L[0] ** 11) / 11 * np.prod(H[1:])
    est, *_ = O.eval_regions_kform(d, "monomial", O.params(k), L, H, 1, plain=True)
    assert abs(est[0] - exact) > 1e-9 * exact


@pytest.mark.parametrize("fam", ["oscillatory", "product_peak", "corner_peak", "gaussian", "c0",
                                 "discontinuous"])
@pytest.mark.parametrize("plain", [False, True])
def test_kform_agrees_with_reference_map(fam, plain):
    """Both modes evaluate the same rule at points a few roundings apart: the
    estimates agree within that conditioning (a dropped origin, a wrong sign
    of the axis scale or a transposed centre would not), and the split axes
    agree away from ties."""
    d, n = 5, 400
    rng = np.random.default_rng(11)
    a = 5 + 10 * rng.random(d) if fam != "oscillatory" else 1 + 0.5 * rng.random(d)
    u = 0.2 + 0.6 * rng.random(d)
    H = (10.0 ** rng.uniform(-2.5, -0.5, size=(d, n)))
    L = rng.random((d, n)) * (1 - H)
    L, H = L.ravel().copy(), H.ravel().copy()
    prm = O.params(a, u)
    e1, _, d1, fd1, rs1 = O.eval_regions(d, fam, prm, L, H, n, want_extra=True)
    e2, _, d2, _, _ = O.eval_regions_kform(d, fam, prm, L, H, n, plain=plain)
    cond = 64 * 2.0 ** -53 * float(np.sum(np.asarray(a) ** 2 if fam == "gaussian" else a))
    scale = O.est_scale(rs1)
    assert np.all(np.abs(e1 - e2) <= (1e-13 + cond) * scale + 1e-300)
    srt = np.sort(fd1, axis=1)
    clear = (srt[:, -1] - srt[:, -2]) > 1e-8 * rs1[:, 14]
    assert np.all((d1 == d2) | ~clear)


def test_kform_discontinuous_mask_at_kernel_points():
    """The discontinuous family's mask is decided on the kernel's t = x - u:
    a region whose right edge is u_j exactly keeps its points on the valid
    side (t <= 0 exactly at the edge) and integrates the exponential."""
    d = 2
    a, u = [2.0, 3.0], [0.5, 0.75]
    L = np.array([0.25, 0.5])
    H = np.array([0.25, 0.25])
    est, *_ = O.eval_regions_kform(d, "discontinuous", O.params(a, u), L, H, 1)
    exact = (np.exp(2 * 0.5) - np.exp(2 * 0.25)) / 2 * (np.exp(3 * 0.75) - np.exp(3 * 0.5)) / 3
    assert abs(est[0] - exact) <= 1e-10 * exact


# ---------------------------------------------------------------------------
# Loop options of orc_integrate_ex

def test_budget_fix_branch_is_taken():
    """cos(sum (i+1) x_i): the finished regions' errors overflow the budget in
    5D and 6D, so fix_error_budget_overflow (Workspace.cuh:121-146) runs; the
    runs still converge to the closed form."""
    for d, eps, least in ((5, 1e-5, 1), (6, 1e-4, 1), (6, 1e-6, 1)):
        a, u = [float(i + 1) for i in range(d)], [0.0] * d
        prm = O.params(a, u)
        r = O.integrate(d, "oscillatory", prm, eps, 1e-20)
        assert r.budget_fixes >= least
        assert r.status == 0
        assert abs(r.estimate - O.exact("oscillatory", d, prm)) <= r.errorest
    # an integrand whose finished errors stay within budget never takes it
    r = O.integrate(3, "gaussian", O.params([5.0] * 3, [0.5] * 3), 1e-6, 1e-20)
    assert r.budget_fixes == 0


def test_predict_split_rule():
    """Under the predict_split rule no region finishes by relerr before
    iteration 15 while nothing has finished (every region of those
    iterations is split again: m == n), and relerr is on from iteration 15,
    where regions finish."""
    d, a, u, eps = 3, [10.0] * 3, [0.3, 0.5, 0.7], 1e-6
    r, tr = O.integrate(d, "c0", O.params(a, u), eps, 1e-20, trace_cap=800,
                        predict_split=True)
    assert len(tr["n"]) > 16
    assert all(int(m) == int(n) for n, m in zip(tr["n"][:15], tr["m"][:15]))
    assert int(tr["m"][15]) < int(tr["n"][15])
    assert r.status == 0
    assert abs(r.estimate - O.exact("c0", d, O.params(a, u))) <= 2 * r.errorest
    # with the caller's flag instead, regions finish from the first iterations
    r2, tr2 = O.integrate(d, "c0", O.params(a, u), eps, 1e-20, trace_cap=800)
    assert any(int(m) < int(n) for n, m in zip(tr2["n"][1:15], tr2["m"][1:15]))


def test_iteration_limit_and_final_list():
    """max_iters stops before that iteration with the list it would evaluate
    (the split of the previous one); the full run's final list was evaluated
    and its estimates sum, with the finished regions', to the result."""
    d, a, u, eps = 4, [10.0] * 4, [0.4, 0.5, 0.6, 0.45], 1e-5
    prm = O.params(a, u)
    full, tr = O.integrate(d, "gaussian", prm, eps, 1e-20, trace_cap=800, keep_list=True)
    lst = full.final_list
    assert lst["n"] == int(tr["n"][-1]) and lst["est"] is not None
    for k in (1, 3, len(tr["n"]) - 1):
        part = O.integrate(d, "gaussian", prm, eps, 1e-20, max_iters=k, keep_list=True)
        assert part.suspended
        pl = part.final_list
        assert pl["n"] == int(tr["n"][k]) and pl["est"] is None
        # the list is the oracle's split of iteration k-1: boxes inside the
        # unit cube whose volumes sum to what was still active
        Lm, Hm = pl["L"].reshape(d, -1), pl["H"].reshape(d, -1)
        assert np.all(Lm >= 0) and np.all(Lm + Hm <= 1 + 1e-15)
    # the evaluated final list: its estimates are those of an evaluation
    e, r_, _ = O.eval_regions(d, "gaussian", prm, lst["L"], lst["H"], lst["n"])
    assert np.array_equal(e, lst["est"])
