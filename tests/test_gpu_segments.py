"""The batch's persistent kernel (pagani_batch_set_segments) on the GPU.

One workgroup runs an item's whole loop in its own segment of one pool.  The
iterations are the single-integral loop's (Workspace.cuh:362-494 with the
device-resident loop's decisions), but every region is evaluated with one
wave per 64-region batch and the iteration's sums are reduced in another
order, so results agree with the regular path to rounding, not to the bit.
Checked here against the regular path (the same batch with the kernel off),
against the oracle's loop (status, region and iteration counts) and against
the closed forms.  An item whose list outgrows its segment is handed over
with its loop state and continues on the regular loop from the iteration
it stopped at; items the kernel does not take (FP32) come out bit for bit
as the regular path gives them.
"""
import pytest

pytestmark = pytest.mark.gpu

SEG = 65536


def _items(count, seed, max_d=8, min_eps=1e-6):
    from gpuintegration_amd import workloads as W
    out = []
    for fam, a, u, eps in W.batch(count, seed=seed):
        if len(a) > max_d:
            a, u = a[:max_d], u[:max_d]
        out.append((fam, a, u, max(eps, min_eps)))
    return out


def _hbm():
    import torch
    return torch.cuda.get_device_properties(0).total_memory


def _run(pg, items, seg, workers=1, mem=None, **kw):
    with pg.Batch(0, workers) as b:
        if mem:
            b.set_device_mem(mem)
        if seg:
            b.set_segments(seg)
        res, codes = b.integrate(items, **kw)
        st = b.stats()
    assert all(c >= 0 for c in codes), codes
    return res, st


def _key(r):
    return (r.estimate, r.errorest, r.nregions, r.iters, r.status)


def test_segments_agree_with_the_regular_path(pg, gpu):
    items = _items(160, seed=31, max_d=7, min_eps=1e-5)
    seg, st = _run(pg, items, SEG)
    reg, _ = _run(pg, items, 0)
    assert st["segment_integrals"] >= len(items) // 2, st["segment_integrals"]
    assert st["integrals"] == len(items) and st["segment_ms"] > 0
    same_path = 0
    for i, (s, r) in enumerate(zip(seg, reg)):
        assert s.status == r.status, i
        # within the combined error estimates always
        assert abs(s.estimate - r.estimate) <= s.errorest + r.errorest, i
        if (s.nregions, s.iters) == (r.nregions, r.iters):
            same_path += 1
            # the same iterations: the totals differ by the rounding of the
            # per-region arithmetic and the order of the sums.  A region's
            # error is a difference of rule sums of the estimate's size, so
            # its rounding is measured against the estimate (measured:
            # 9e-18 on an estimate of 0.58, 8e-9 of the error itself)
            assert abs(s.estimate - r.estimate) <= 1e-11 * abs(r.estimate) + 1e-300, i
            assert abs(s.errorest - r.errorest) <= \
                1e-12 * abs(r.estimate) + 1e-9 * r.errorest + 1e-300, i
    # a different path needs a decision that rounding tips over
    assert same_path >= len(items) - 2, same_path


def test_segments_are_deterministic(pg, gpu):
    items = _items(96, seed=17, max_d=6, min_eps=1e-4)
    a, sa = _run(pg, items, SEG)
    # the other order; two workers (the regular path's items then run on two
    # streams), with the first run's memory-model total
    b, sb = _run(pg, items[::-1], SEG, workers=2, mem=_hbm())
    assert sa["segment_integrals"] == sb["segment_integrals"] > 0
    assert [_key(x) for x in a] == [_key(x) for x in b[::-1]]


def test_items_beyond_the_segment_continue_on_the_regular_path(pg, gpu):
    # 256 regions: the default split of d >= 5 (2^d) fits, most lists
    # outgrow it within a few iterations and are handed over
    items = _items(48, seed=5, max_d=7, min_eps=1e-4)
    seg, st = _run(pg, items, 256)
    reg, st0 = _run(pg, items, 0)
    assert 0 < st["segment_integrals"] < len(items)
    # handed over, an item continues from its state: the same iterations as
    # the regular path, totals to rounding
    for i, (s, r) in enumerate(zip(seg, reg)):
        assert s.status == r.status and (s.nregions, s.iters) == (r.nregions, r.iters), i
        assert abs(s.estimate - r.estimate) <= 1e-11 * abs(r.estimate) + 1e-300, i
        assert abs(s.errorest - r.errorest) <= \
            1e-12 * abs(r.estimate) + 1e-9 * r.errorest + 1e-300, i
    # no iteration is done twice: the kernel's finished iterations and the
    # regular loop's after the hand-over evaluate what the regular path does
    assert st["regions_evaluated"] == st0["regions_evaluated"]
    assert st["iterations"] == st0["iterations"]


def test_fp32_items_skip_the_kernel(pg, gpu):
    items = _items(12, seed=3, max_d=5, min_eps=1e-3)
    seg, st = _run(pg, items, SEG, precision=pg.FP32)
    reg, _ = _run(pg, items, 0, precision=pg.FP32)
    assert st["segment_integrals"] == 0
    assert [_key(x) for x in seg] == [_key(x) for x in reg]


def test_segments_match_the_oracle_loop(pg, gpu):
    """C0 and discontinuous items (no closed-form check needed) at small d:
    the oracle's status, region count and iteration count; estimates within
    the combined error estimates."""
    import oracle_api as O
    hbm = _hbm()
    items = [it for it in _items(200, seed=23, max_d=4, min_eps=1e-5)
             if it[0] in ("c0", "discontinuous")][:24]
    seg, st = _run(pg, items, SEG)
    # measured: 21 of these 24 stay within 65536 regions and the budget
    assert st["segment_integrals"] >= len(items) * 3 // 4, st["segment_integrals"]
    for i, ((fam, a, u, eps), r) in enumerate(zip(items, seg)):
        o = O.integrate(len(a), fam, O.params(a, u), eps, 1e-20, total_mem=hbm)
        assert (o.status, o.nregions, o.iters) == (r.status, r.nregions, r.iters), i
        assert abs(r.estimate - o.estimate) <= r.errorest + o.errorest, i


def test_segments_are_accurate(pg, gpu):
    from gpuintegration_amd import workloads as W
    items = [it for it in _items(120, seed=41, max_d=6, min_eps=1e-5)
             if it[0] in ("oscillatory", "product_peak", "corner_peak", "gaussian")]
    seg, st = _run(pg, items, SEG)
    assert st["segment_integrals"] > 0
    for (fam, a, u, eps), r in zip(items, seg):
        if r.status == 0:
            exact = W.exact(fam, a, u)
            assert abs(r.estimate - exact) <= eps * abs(exact), (fam, len(a), eps)


def test_segment_size_bounds(pg, gpu):
    with pg.Batch(0, 1) as b:
        for bad in (1, 255, (1 << 22) + 1):
            with pytest.raises(pg.PaganiError):
                b.set_segments(bad)
        b.set_segments(0)
        b.set_segments(1 << 22)


@pytest.mark.parametrize("relerr,init_k", [(False, 0), (True, 3)])
def test_segments_up_to_ten_dimensions_and_loop_options(pg, gpu, relerr, init_k):
    """d = 2..10 (every dimension the Genz descriptors cover), the relerr
    classification off, and an explicit initial split: the same iterations
    as the regular path, totals to rounding."""
    from gpuintegration_amd import workloads as W
    items = []
    for i, (fam, a, u, eps) in enumerate(W.batch(60, seed=77)):
        d = 2 + i % 9
        a = (list(a) * 4)[:d]
        u = (list(u) * 4)[:d]
        items.append((fam, a, u, max(eps, 1e-3 if d >= 9 else 1e-4)))
    if init_k == 3:  # 3^d regions must fit the segment
        items = [it for it in items if 3 ** len(it[1]) <= SEG]
    seg, st = _run(pg, items, SEG, relerr=relerr, init_k=init_k)
    reg, st0 = _run(pg, items, 0, relerr=relerr, init_k=init_k)
    assert st["segment_integrals"] > 0
    assert st["regions_evaluated"] == st0["regions_evaluated"]
    for i, (s, r) in enumerate(zip(seg, reg)):
        assert s.status == r.status and (s.nregions, s.iters) == (r.nregions, r.iters), i
        assert abs(s.estimate - r.estimate) <= 1e-11 * abs(r.estimate) + 1e-300, i


def test_segments_larger_than_the_pool_cap(pg, gpu):
    """Segments of 2^22 10D regions (1.4 GB each) for 64 integrals: 90 GB
    of segments against a pool capped at a quarter of the device, so the
    launch runs fewer workgroups that take more items each; results as with
    the kernel off."""
    import numpy as np
    rng = np.random.default_rng(3)
    items = [("gaussian", list(rng.uniform(3, 4, 10)), list(rng.uniform(0.3, 0.7, 10)), 1e-3)
             for _ in range(64)]
    seg, st = _run(pg, items, 1 << 22)
    reg, st0 = _run(pg, items, 0)
    assert st["segment_integrals"] >= len(items) // 2
    for i, (s, r) in enumerate(zip(seg, reg)):
        assert s.status == r.status and (s.nregions, s.iters) == (r.nregions, r.iters), i
        assert abs(s.estimate - r.estimate) <= 1e-11 * abs(r.estimate) + 1e-300, i


def test_bad_and_mixed_items_with_segments(pg, gpu):
    """With the kernel on, an item of an unknown family still fails alone
    (EBADARG, status 1), FP32 items run on the regular loop beside FP64
    items in the kernel, and every good item's result is the one the batch
    gives with the kernel off, to rounding."""
    import ctypes
    items = _items(12, seed=9, max_d=6, min_eps=1e-3)
    lib = pg.load()

    def run(seg):
        with pg.Batch(0, 1) as b:
            if seg:
                b.set_segments(seg)
            arr = (pg._BatchItem * len(items))()
            for i, (fam, a, u, eps) in enumerate(items):
                arr[i].f = pg.genz(fam, a, u)
                arr[i].epsrel = eps
                arr[i].epsabs = 1e-20
                arr[i].precision = pg.FP32 if i % 3 == 1 else pg.FP64
                arr[i].relerr_classification = 1
            arr[4].f.family = 99
            res = (pg._Result * len(items))()
            codes = (ctypes.c_int * len(items))()
            rc = lib.pagani_integrate_batch(b._b, arr, len(items), res, codes)
            st = b.stats()
        return rc, list(codes), [res[i] for i in range(len(items))], st

    rc, codes, res, st = run(SEG)
    rc0, codes0, res0, _ = run(0)
    assert rc == rc0 == pg.EBADARG
    assert codes == codes0 and codes[4] == pg.EBADARG and res[4].status == 1
    assert st["segment_integrals"] > 0
    for i, (s, r) in enumerate(zip(res, res0)):
        if i == 4:
            continue
        assert codes[i] >= 0
        assert s.status == r.status and (s.nregions, s.iters) == (r.nregions, r.iters), i
        assert abs(s.estimate - r.estimate) <= 1e-11 * abs(r.estimate) + 1e-300, i
        if i % 3 == 1:  # FP32: never in the kernel, bit for bit
            assert (s.estimate, s.errorest) == (r.estimate, r.errorest), i
