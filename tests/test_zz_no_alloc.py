"""pagani_refine_dev_noalloc against pagani_refine_dev, through the C ABI.

The no-allocation variant keeps its chunk, fold and result records in a
fixed device array of the library instead of workspace buffers.  Its
refined errors and flags must equal those of pagani_refine_dev bit for bit,
and its sums and active count must be the same numbers (the same kernels in
the same order).  Two calls in a row, the second on a smaller list, show
that what the first call left in the shared array does not leak into the
second; a list above PAGANI_NOALLOC_MAX_REGIONS is refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_REGIONS = 1 << 26


def _case(n, seed):
    rng = np.random.default_rng(seed)
    est = rng.normal(1.0, 0.3, n) * 10.0 ** rng.uniform(-6, 0, n)
    err = np.abs(est) * 10.0 ** rng.uniform(-9, -1, n)
    err[rng.random(n) < 0.01] = 0.0
    pest = (est[: n // 2] + est[n // 2:]) * (1 + 1e-4 * rng.normal(size=n // 2))
    return est, err, pest


def _run_pair(pg, ws, gpu, n, seed, parents, relerr):
    import torch
    est, err, pest = _case(n, seed)
    d_est = torch.from_numpy(est).to(gpu)
    d_err = torch.from_numpy(err).to(gpu)
    d_pest = torch.from_numpy(pest).to(gpu) if parents else None
    e_ref = torch.empty(n, dtype=torch.float64, device=gpu)
    a_ref = torch.empty(n, dtype=torch.uint8, device=gpu)
    sums_ref, na_ref = ws.refine(n, d_est, d_err, e_ref, d_pest, 1e-4, relerr, a_ref)
    # outputs prefilled with garbage, so a value the call fails to write shows
    e_new = torch.full((n,), -7.0, dtype=torch.float64, device=gpu)
    a_new = torch.full((n,), 9, dtype=torch.uint8, device=gpu)
    sums_new = torch.full((4,), -7.0, dtype=torch.float64, device=gpu)
    na_new = torch.zeros(1, dtype=torch.int64, device=gpu)
    torch.cuda.synchronize()
    rc = ws.lib.pagani_refine_dev_noalloc(
        ws._ws, n, pg._ptr(d_est), pg._ptr(d_err), pg._ptr(e_new), pg._ptr(d_pest), 1e-4,
        1 if relerr else 0, pg._ptr(a_new), pg._ptr(sums_new), pg._ptr(na_new))
    ws._check(rc)
    pg_stream_sync(ws)
    assert torch.equal(e_new, e_ref)
    assert torch.equal(a_new, a_ref)
    assert sums_new.cpu().tolist() == sums_ref
    assert int(na_new.item()) == na_ref
    return na_ref


def pg_stream_sync(ws):
    """Wait for the workspace's own stream, where the entry point enqueues."""
    import torch
    torch.cuda.ExternalStream(ws.lib.pagani_ws_stream(ws._ws)).synchronize()


@pytest.mark.parametrize("parents,relerr", [(True, True), (True, False), (False, True)])
def test_noalloc_matches_refine_two_calls(pg, gpu, parents, relerr):
    with pg.Workspace(3, 0) as ws:
        first = _run_pair(pg, ws, gpu, 3 * 4096 * 256 + 2 * 123, 11, parents, relerr)
        second = _run_pair(pg, ws, gpu, 4096 + 18, 12, parents, relerr)
        assert first > second > 0


def test_noalloc_largest_list(pg, gpu):
    """A list of the maximum size (every chunk record of the array in use),
    then a small one."""
    with pg.Workspace(3, 0) as ws:
        _run_pair(pg, ws, gpu, MAX_REGIONS, 21, True, True)
        _run_pair(pg, ws, gpu, 2, 22, True, True)


def test_noalloc_rejects_oversize(pg, gpu):
    import torch
    with pg.Workspace(3, 0) as ws:
        x = torch.zeros(4, dtype=torch.float64, device=gpu)
        a = torch.zeros(4, dtype=torch.uint8, device=gpu)
        # the size check comes before any pointer is touched
        rc = ws.lib.pagani_refine_dev_noalloc(ws._ws, MAX_REGIONS + 2, pg._ptr(x), pg._ptr(x),
                                              pg._ptr(x), None, 1e-4, 1, pg._ptr(a), None, None)
        assert rc == pg.EBADARG
        assert b"PAGANI_NOALLOC_MAX_REGIONS" in ws.lib.pagani_ws_last_error(ws._ws)
