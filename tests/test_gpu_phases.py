"""Region-list phases on the GPU vs the oracle and the reference's KATs.

- refinement + relerr flags (RefineError, Phases.cuh:219-268): bit-identical
  to the oracle (plain IEEE double, same operation order; the phases are
  built without FMA contraction); its sums within 1e-15 relative
- filter / split / fused filter+split (RegionFiltering.cu, RegionSplitting.cu):
  bit-identical
- reductions (Reduction.cu, prefix_sum.cu, min_max.cu): exact on integer
  data, min/max positions of the reference test
- heuristic classification (heuristic_classifier.cuh:362-451): the device
  version evaluates each run of threshold raises in one pass; every result
  field equals the oracle's sequential restatement (thresholds, counts and
  flags exactly; the finished error within the rounding of its sum)
"""
import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu


def dev(a, gpu):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(gpu)


def host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("n", [2, 64, 4096, 12346, 1 << 20])
@pytest.mark.parametrize("relerr", [True, False])
def test_refine_bit_identical(pg, gpu, n, relerr):
    import torch
    rng = np.random.default_rng(n)
    est = rng.normal(1.0, 0.3, n) * 10.0 ** rng.uniform(-6, 0, n)
    err = np.abs(est) * 10.0 ** rng.uniform(-9, -1, n)
    err[rng.random(n) < 0.01] = 0.0
    pest = (est[: n // 2] + est[n // 2:]) * (1 + 1e-4 * rng.normal(size=n // 2))
    epsrel = 1e-4
    with pg.Workspace(3, 0) as ws:
        e_out = torch.empty(n, dtype=torch.float64, device=gpu)
        act = torch.empty(n, dtype=torch.uint8, device=gpu)
        sums, na = ws.refine(n, dev(est, gpu), dev(err, gpu), e_out, dev(pest, gpu), epsrel,
                             relerr, act)
        g_err, g_act = host(e_out), host(act)
    o_err, o_act = O.refine(est, err, pest, epsrel, relerr)
    assert np.array_equal(g_err, o_err)
    assert np.array_equal(g_act, o_act)
    assert na == int(o_act.sum())
    want = [est.sum(), o_err.sum(), (o_act * est).sum(), (o_act * o_err).sum()]
    for g, w, mag in zip(sums, want, [np.abs(est).sum(), o_err.sum(),
                                      np.abs(est).sum(), o_err.sum()]):
        assert abs(g - w) <= 1e-14 * mag


def test_refine_first_iteration(pg, gpu):
    """No parents: every region active, errors unchanged."""
    import torch
    n = 5000
    rng = np.random.default_rng(1)
    est, err = rng.random(n), rng.random(n) * 1e-3
    with pg.Workspace(3, 0) as ws:
        e_in = dev(err, gpu)
        act = torch.empty(n, dtype=torch.uint8, device=gpu)
        sums, na = ws.refine(n, dev(est, gpu), e_in, e_in, None, 1e-3, True, act)
        assert np.array_equal(host(e_in), err) and host(act).all() and na == n


def _filter_inputs(d, k, inactive, seed=0):
    L, H, n = O.uniform_split(d, k)
    rng = np.random.default_rng(seed)
    est, err = rng.random(n), rng.random(n)
    dim = rng.integers(0, d, n).astype(np.int32)
    act = np.ones(n, dtype=np.uint8)
    act[inactive] = 0
    return L, H, n, est, err, dim, act


@pytest.mark.parametrize("inactive", ["none", "15", "first_last", "all", "random"])
def test_filter_and_split_bit_identical(pg, gpu, inactive):
    import torch
    d, k = 2, 10
    n = k ** d
    sel = {"none": [], "15": [15], "first_last": [0, n - 1], "all": list(range(n)),
           "random": list(np.random.default_rng(5).choice(n, 37, replace=False))}[inactive]
    L, H, n, est, err, dim, act = _filter_inputs(d, k, sel)
    m_o, L2o, H2o, peo, pro, d2o = O.filter_regions(d, n, L, H, act, est, err, dim)
    with pg.Workspace(d, 0) as ws:
        mm = max(m_o, 1)
        L2 = torch.empty(mm * d, dtype=torch.float64, device=gpu)
        H2 = torch.empty_like(L2)
        pe = torch.empty(mm, dtype=torch.float64, device=gpu)
        pr = torch.empty_like(pe)
        d2 = torch.empty(mm, dtype=torch.uint8, device=gpu)
        Ld, Hd, ad = dev(L, gpu), dev(H, gpu), dev(act, gpu)
        ed, rd, dd = dev(est, gpu), dev(err, gpu), dev(dim.astype(np.uint8), gpu)
        m = ws.filter(n, Ld, Hd, ad, ed, rd, dd, L2, H2, pe, pr, d2)
        assert m == m_o
        if m == 0:
            return
        assert np.array_equal(host(L2)[: m * d], L2o) and np.array_equal(host(H2)[: m * d], H2o)
        assert np.array_equal(host(pe)[:m], peo) and np.array_equal(host(pr)[:m], pro)
        assert np.array_equal(host(d2)[:m].astype(np.int32), d2o)
        # split of the filtered list, and the fused kernel, against the oracle
        Ls, Hs = O.split_regions(d, m, L2o, H2o, d2o)
        C1 = torch.empty(2 * m * d, dtype=torch.float64, device=gpu)
        C2 = torch.empty_like(C1)
        ws.split(m, L2, H2, d2, C1, C2)
        assert np.array_equal(host(C1), Ls) and np.array_equal(host(C2), Hs)
        F1, F2 = torch.empty_like(C1), torch.empty_like(C1)
        pe2, pr2 = torch.empty_like(pe), torch.empty_like(pe)
        m2 = ws.filter_split(n, Ld, Hd, ad, ed, rd, dd, F1, F2, pe2, pr2)
        assert m2 == m
        assert np.array_equal(host(F1), Ls) and np.array_equal(host(F2), Hs)
        assert np.array_equal(host(pe2)[:m], peo) and np.array_equal(host(pr2)[:m], pro)


def test_filter_split_large_random(pg, gpu):
    """Fused filter+split on 3M 8D regions with a random 40% active pattern."""
    import torch
    d, n = 8, 3_000_001
    rng = np.random.default_rng(11)
    L = rng.random(d * n)
    H = rng.random(d * n) * 0.1
    act = (rng.random(n) < 0.4).astype(np.uint8)
    est, err = rng.random(n), rng.random(n)
    dim = rng.integers(0, d, n).astype(np.int32)
    m_o, L2o, H2o, peo, pro, d2o = O.filter_regions(d, n, L, H, act, est, err, dim)
    Ls, Hs = O.split_regions(d, m_o, L2o, H2o, d2o)
    with pg.Workspace(d, 0) as ws:
        F1 = torch.empty(2 * m_o * d, dtype=torch.float64, device=gpu)
        F2 = torch.empty_like(F1)
        pe = torch.empty(m_o, dtype=torch.float64, device=gpu)
        pr = torch.empty_like(pe)
        m = ws.filter_split(n, dev(L, gpu), dev(H, gpu), dev(act, gpu), dev(est, gpu),
                            dev(err, gpu), dev(dim.astype(np.uint8), gpu), F1, F2, pe, pr)
        assert m == m_o
        assert np.array_equal(host(F1), Ls) and np.array_equal(host(F2), Hs)
        assert np.array_equal(host(pe), peo) and np.array_equal(host(pr), pro)


@pytest.mark.parametrize("start", [-1000.0, 1000.0, 1000.5, 1.0])
def test_reduction_kat(pg, gpu, start):
    """Reduction.cu:49-157: iota arrays of 2^1..2^25 and 11..1009 elements."""
    with pg.Workspace(2, 0) as ws:
        sizes = [2 ** i for i in range(1, 26)] + list(range(11, 1010, 97))
        for size in sizes:
            arr = start + np.arange(size, dtype=np.float64)
            got = ws.sum(dev(arr, gpu))
            assert got == arr.sum() or abs(got - arr.sum()) <= 1e-15 * np.abs(arr).sum()


def test_dot_scan_minmax_kat(pg, gpu):
    import torch
    with pg.Workspace(2, 0) as ws:
        # prefix_sum.cu:19-104: the scan of set flags equals the running count
        for n in (8, 10000, 10001, 3_000_017):
            rng = np.random.default_rng(n)
            f = (rng.random(n) < 0.37).astype(np.uint8)
            off = torch.empty(n, dtype=torch.int64, device=gpu)
            tot = ws.scan_flags(dev(f, gpu), off)
            want = np.concatenate([[0], np.cumsum(f)[:-1]])
            assert tot == f.sum() and np.array_equal(host(off), want)
            x = rng.random(n)
            assert abs(ws.dot(dev(f, gpu), dev(x, gpu)) - (f * x).sum()) <= 1e-14 * x.sum()
        # min_max.cu: extremes at 0, n/4, n/2, n-1
        n = 1 << 20
        for pos in (0, n // 4, n // 2, n - 1):
            x = np.random.default_rng(pos).random(n) + 10.0
            x[pos] = -5.0
            x[(pos + 7) % n] = 1e6
            assert ws.minmax(dev(x, gpu)) == (-5.0, 1e6)


def _oracle_classify(d, err, iter_err, iter_fin, tot_fin, hist, epsrel):
    c = O.OrcClassifier()
    O.LIB.orc_clf_init(c, d, epsrel, 1e-12, int(16e9))
    for h in hist:
        O.LIB.orc_clf_store(c, h)
    flags = np.zeros(err.size, dtype=np.uint8)
    res = O.OrcClassRes()
    O.LIB.orc_clf_classify(c, O.dptr(err), err.size, iter_err, iter_fin, tot_fin, O.uptr(flags),
                           res)
    return res, flags


def _check_classify(pg, gpu, d, err, iter_fin, tot_fin, hist, epsrel, device_search=True):
    import torch
    iter_err = err.sum()
    ores, oflags = _oracle_classify(d, err, iter_err, iter_fin, tot_fin, hist, epsrel)
    with pg.Workspace(d, 0) as ws:
        ws.set_device_search(device_search)
        flags = torch.empty(err.size, dtype=torch.uint8, device=gpu)
        g = ws.classify(dev(err, gpu), err.size, epsrel, hist, iter_err, iter_fin, tot_fin, flags)
        gflags = host(flags)
    for key in ("pass_mem", "pass_errorest_budget", "num_active"):
        assert g[key] == getattr(ores, key), key
    for key in ("threshold", "percent_mem_active", "range_low", "range_high",
                "max_budget_perc_to_cover", "max_active_perc"):
        assert g[key] == getattr(ores, key), key
    assert abs(g["finished_errorest"] - ores.finished_errorest) <= 1e-12 * iter_err
    assert g["threshold_passes"] == ores.threshold_passes
    assert np.array_equal(gflags, oflags)


@pytest.mark.parametrize("device_search", [True, False])
def test_classifier_kat(pg, gpu, device_search):
    """test_heuristic_classifier.cpp:32-104 on the device, the threshold
    search run on the device and on the host."""
    errs = np.array([.075, .99, .079, 101.96, 101.33, 1.93, 101.99])
    _check_classify(pg, gpu, 3, errs, 0.0, 4.2, [7000.0] * 3, 1e-3, device_search)


@pytest.mark.parametrize("device_search", [True, False])
@pytest.mark.parametrize("seed", range(6))
def test_classifier_random(pg, gpu, seed, device_search):
    """Log-normal error lists of 10^4..10^6 regions under varied budgets."""
    rng = np.random.default_rng(seed)
    n = int(10 ** rng.uniform(4, 6))
    err = 10.0 ** rng.normal(-8, 1.5, n)
    iter_err = err.sum()
    est = iter_err / float(rng.choice([1e-3, 1e-4, 1e-5]))
    tot_fin = iter_err * rng.uniform(0, 2)
    iter_fin = iter_err * rng.uniform(0, 0.3)
    _check_classify(pg, gpu, 8, err, iter_fin, tot_fin, [est] * 3, 1e-4, device_search)


def test_classifier_search_device_vs_host_in_the_loop(pg, gpu):
    """A loop that classifies (the 8D Gaussian of the benchmark at 1e-6 from
    4^8 regions classifies in about ten iterations): the device search and
    the host search give the same trajectory and result, bit for bit."""
    from gpuintegration_amd import workloads as W
    fam, a, u = W.gaussian_8d(2024)
    out = []
    for device_search in (True, False):
        with pg.Workspace(8, 0) as ws:
            ws.set_device_search(device_search)
            ws.set_timing(True)
            r = ws.integrate_genz(fam, a, u, epsrel=1e-6, epsabs=1e-20, init_k=4)
            out.append((r, ws.trace(), ws.stats()))
    (r1, t1, s1), (r2, t2, s2) = out
    assert s1["classify_passes"] > 0 and s1["classify_passes"] == s2["classify_passes"]
    assert (r1.estimate, r1.errorest, r1.nregions, r1.iters, r1.status) == \
        (r2.estimate, r2.errorest, r2.nregions, r2.iters, r2.status)
    for k in ("n", "m", "estimate", "errorest"):
        assert np.array_equal(t1[k], t2[k]), k
