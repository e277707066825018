"""Batches of independent integrals (pagani_integrate_batch) on the GPU.

Each item's result must be the single-integral loop's result for that item,
bit for bit, on a workspace with the batch's memory-model total: the batch
changes only which host thread and stream run an integral and when.
"""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _items(count=48, seed=11):
    from gpuintegration_amd import workloads as W
    out = []
    for fam, a, u, eps in W.batch(count, seed=seed):
        if len(a) > 6:
            a, u = a[:6], u[:6]
        out.append((fam, a, u, max(eps, 1e-4)))
    return out


def _hbm():
    import torch
    return torch.cuda.get_device_properties(0).total_memory


def _single(pg, items, mem_total):
    wss = {}
    out = []
    for fam, a, u, eps in items:
        d = len(a)
        if d not in wss:
            wss[d] = pg.Workspace(d, 0)
            wss[d].set_device_mem(mem_total)
        out.append(wss[d].integrate_genz(fam, a, u, epsrel=eps, epsabs=1e-20))
    for ws in wss.values():
        ws.close()
    return out


def _single_in_turn(pg, items, mem_total):
    """Lone runs one dimension at a time (highest first, as the batch takes
    them), each dimension's workspace closed before the next is made, so at
    most one pool holds device memory."""
    out = [None] * len(items)
    for d in sorted({len(it[1]) for it in items}, reverse=True):
        with pg.Workspace(d, 0) as ws:
            ws.set_device_mem(mem_total)
            for i, (fam, a, u, eps) in enumerate(items):
                if len(a) == d:
                    out[i] = ws.integrate_genz(fam, a, u, epsrel=eps, epsabs=1e-20)
    return out


@pytest.mark.parametrize("workers", [1, 3])
def test_batch_matches_single_integrals(pg, gpu, workers):
    items = _items()
    with pg.Batch(0, workers) as b:
        res, codes = b.integrate(items)
        st = b.stats()
    assert len(res) == len(items)
    assert all(c >= 0 for c in codes), codes
    want = _single(pg, items, _hbm() // workers)
    for i, (r, w) in enumerate(zip(res, want)):
        assert (r.estimate, r.errorest, r.nregions, r.iters, r.status) == \
               (w.estimate, w.errorest, w.nregions, w.iters, w.status), (i, items[i][0])
    assert st["integrals"] == len(items)


def test_batch_results_are_accurate(pg, gpu):
    from gpuintegration_amd import workloads as W
    items = [it for it in _items(64, seed=5) if it[0] in ("gaussian", "product_peak",
                                                         "oscillatory", "corner_peak")]
    with pg.Batch(0, 2) as b:
        res, codes = b.integrate(items)
    for (fam, a, u, eps), r, c in zip(items, res, codes):
        assert c >= 0
        exact = W.exact(fam, a, u)
        if r.status == 0:
            # converged: the error estimate bounds the true error with margin
            assert abs(r.estimate - exact) <= max(10 * r.errorest, 1e-15), (fam, len(a), eps)


def test_batch_explicit_mem_total_and_repeat(pg, gpu):
    items = _items(24, seed=3)
    with pg.Batch(0, 2) as b:
        b.set_device_mem(16_000_000_000)
        r1, _ = b.integrate(items)
        r2, _ = b.integrate(items[::-1])
    want = _single(pg, items, 16_000_000_000)
    for x, y, w in zip(r1, r2[::-1], want):
        assert (x.estimate, x.errorest, x.nregions) == (w.estimate, w.errorest, w.nregions)
        assert (y.estimate, y.errorest, y.nregions) == (w.estimate, w.errorest, w.nregions)


def test_batch_bad_item_fails_alone(pg, gpu):
    items = _items(8, seed=9)
    lib = pg.load()
    with pg.Batch(0, 2) as b:
        arr = (pg._BatchItem * len(items))()
        for i, (fam, a, u, eps) in enumerate(items):
            arr[i].f = pg.genz(fam, a, u)
            arr[i].epsrel = eps
            arr[i].epsabs = 1e-20
            arr[i].precision = pg.FP64
            arr[i].relerr_classification = 1
        arr[3].f.family = 99
        res = (pg._Result * len(items))()
        codes = (ctypes.c_int * len(items))()
        rc = lib.pagani_integrate_batch(b._b, arr, len(items), res, codes)
        assert rc == pg.EBADARG
        assert codes[3] == pg.EBADARG
        assert b"family" in lib.pagani_batch_last_error(b._b)
        assert all(codes[i] >= 0 for i in range(len(items)) if i != 3)
        assert res[3].status == 1


def test_trim_releases_pool(pg, gpu):
    fam, a, u, eps = _items(4, seed=13)[0]
    with pg.Workspace(len(a), 0) as ws:
        r1 = ws.integrate_genz(fam, a, u, epsrel=eps, epsabs=1e-20)
        assert ws.footprint() > 0
        ws.trim()
        assert ws.footprint() == 0
        r2 = ws.integrate_genz(fam, a, u, epsrel=eps, epsabs=1e-20)
        # a suspended run keeps its pool
        ws.set_iteration_limit(1)
        ws.integrate_genz(fam, a, u, epsrel=eps, epsabs=1e-20)
        if ws.suspended():
            with pytest.raises(pg.PaganiError):
                ws.trim()
    assert (r1.estimate, r1.errorest, r1.nregions) == (r2.estimate, r2.errorest, r2.nregions)


def test_config4_full_batch(pg, gpu):
    """The 512-integral batch exactly as benchmarked (families and d in
    {3..8} drawn, epsrel log-uniform in [1e-6, 1e-3], seed 2024), one call:
    - every item bit for bit equal to its lone run on a workspace with the
      batch's memory total;
    - converged smooth items (oscillatory, product peak, corner peak,
      Gaussian) within their epsrel of the analytic value (the accuracy the
      caller asked for); how many are also within their own error estimate
      is printed (not all are: measured, a 5D oscillatory item at epsrel
      2.7e-5 is 5.7e-10 off, more than its error estimate);
    - C0 and discontinuous items the oracle's loop can run in seconds (d <= 4,
      or d = 5 at epsrel >= 4e-5): the oracle's status, region count and
      iteration count, estimates within the combined error estimates."""
    import oracle_api as O
    from gpuintegration_amd import workloads as W
    items = W.batch(512, 2024)
    hbm = _hbm()
    with pg.Batch(0, 1) as b:
        res, codes = b.integrate(items)
    assert all(c >= 0 for c in codes), [c for c in codes if c < 0]
    want = _single_in_turn(pg, items, hbm)
    bad = [(i, items[i][0], len(items[i][1]), items[i][3],
            (r.estimate, r.errorest, r.nregions, r.iters, r.status),
            (w.estimate, w.errorest, w.nregions, w.iters, w.status))
           for i, (r, w) in enumerate(zip(res, want))
           if (r.estimate, r.errorest, r.nregions, r.iters, r.status) !=
              (w.estimate, w.errorest, w.nregions, w.iters, w.status)]
    assert not bad, bad[:5]
    smooth = ("oscillatory", "product_peak", "corner_peak", "gaussian")
    counts = {}
    for i, ((fam, a, u, eps), r) in enumerate(zip(items, res)):
        if fam in smooth and r.status == 0:
            exact = W.exact(fam, a, u)
            counts.setdefault(fam, [0, 0, 0])
            counts[fam][0] += 1
            counts[fam][1] += abs(r.estimate - exact) <= r.errorest
            counts[fam][2] += 1
            assert abs(r.estimate - exact) <= eps * abs(exact), \
                (i, fam, len(a), eps, r.estimate, exact, r.errorest)
    print("converged smooth items [count, within errorest, within epsrel]:", counts)
    checked = 0
    for i, ((fam, a, u, eps), r) in enumerate(zip(items, res)):
        if fam not in ("c0", "discontinuous"):
            continue
        d = len(a)
        if not (d <= 4 or (d == 5 and eps >= 4e-5)):
            continue
        o = O.integrate(d, fam, O.params(a, u), eps, 1e-20, total_mem=hbm)
        assert o.status == r.status and o.nregions == r.nregions and o.iters == r.iters, i
        assert abs(r.estimate - o.estimate) <= r.errorest + o.errorest, i
        checked += 1
    assert checked >= 40
