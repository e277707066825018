"""Suspend / checkpoint / resume, the CSV region dump and predict_split
snapshots on the GPU.

A run stopped at an iteration boundary and continued -- in the same
workspace, or from a checkpoint file in a new one -- must give exactly the
result of the uninterrupted run: the file carries the whole loop state.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REFTESTS = os.path.join(HERE, "cxx", "libreftests.so")

# 6D Gaussian of the benchmark's parameter family; about a dozen iterations
# at this tolerance, with heuristic classification under the 16e9 model
FAM = "gaussian"
A = [9.6, 10.3, 10.1, 9.8, 10.4, 9.9]
U = [0.31, 0.52, 0.47, 0.66, 0.28, 0.55]
EPS = 1e-6
MEM = int(16e9)


def _key(r):
    return (r.estimate, r.errorest, r.nregions, r.iters, r.status)


def _uninterrupted(pg):
    with pg.Workspace(len(A), 0) as ws:
        ws.set_device_mem(MEM)
        r = ws.integrate_genz(FAM, A, U, epsrel=EPS, epsabs=1e-20)
        tr = ws.trace()
    return r, tr


def test_suspend_and_resume_in_place(pg, gpu):
    r0, tr0 = _uninterrupted(pg)
    assert len(tr0["n"]) > 6
    with pg.Workspace(len(A), 0) as ws:
        ws.set_device_mem(MEM)
        ws.set_iteration_limit(3)
        r = ws.integrate_genz(FAM, A, U, epsrel=EPS, epsabs=1e-20)
        assert ws.suspended()
        assert r.status == 1
        rounds = 1
        again = True
        while again:
            r, again = ws.resume_genz(FAM, A, U)
            rounds += 1
            assert rounds < 100
        assert not ws.suspended()
        tr = ws.trace()
    assert rounds >= 3
    assert _key(r) == _key(r0)
    for k in ("n", "m", "estimate", "errorest"):
        assert list(tr[k]) == list(tr0[k]), k


def test_checkpoint_file_round_trip(pg, gpu, tmp_path):
    r0, tr0 = _uninterrupted(pg)
    path = str(tmp_path / "run.pagani")
    with pg.Workspace(len(A), 0) as ws:
        ws.set_device_mem(MEM)
        ws.set_iteration_limit(5)
        ws.integrate_genz(FAM, A, U, epsrel=EPS, epsabs=1e-20)
        assert ws.suspended()
        ws.save_state(path)
    assert os.path.getsize(path) > 0
    # a fresh workspace (the memory model total travels in the file)
    with pg.Workspace(len(A), 0) as ws:
        ws.load_state(path)
        assert ws.suspended()
        r, again = ws.resume_genz(FAM, A, U)
        assert not again
        tr = ws.trace()
    assert _key(r) == _key(r0)
    assert list(tr["n"]) == list(tr0["n"])


def test_checkpoint_rejects_damage_and_mismatch(pg, gpu, tmp_path):
    path = str(tmp_path / "run.pagani")
    with pg.Workspace(len(A), 0) as ws:
        ws.set_device_mem(MEM)
        ws.set_iteration_limit(4)
        ws.integrate_genz(FAM, A, U, epsrel=EPS, epsabs=1e-20)
        ws.save_state(path)
        # another integrand cannot continue this run
        with pytest.raises(pg.PaganiError):
            ws.resume_genz(FAM, [a * 1.01 for a in A], U)
    blob = bytearray(open(path, "rb").read())
    # one flipped bit in the region bounds
    bad = bytearray(blob)
    bad[len(bad) - 64] ^= 0x10
    p_bad = str(tmp_path / "bad.pagani")
    open(p_bad, "wb").write(bytes(bad))
    p_short = str(tmp_path / "short.pagani")
    open(p_short, "wb").write(bytes(blob[:-100]))
    p_magic = str(tmp_path / "magic.pagani")
    open(p_magic, "wb").write(b"NOTPAGAN" + bytes(blob[8:]))
    with pg.Workspace(len(A), 0) as ws:
        for p, why in ((p_bad, "checksum"), (p_short, "size"), (p_magic, "not a PAGANI")):
            with pytest.raises(pg.PaganiError, match=why):
                ws.load_state(p)
            assert not ws.suspended()
        with pytest.raises(pg.PaganiError):
            ws.resume_genz(FAM, A, U)  # nothing held
    with pg.Workspace(5, 0) as ws5:
        with pytest.raises(pg.PaganiError, match="ndim"):
            ws5.load_state(path)


def test_failed_load_drops_a_held_run(pg, gpu, tmp_path):
    """A load that fails after it has begun to overwrite the region pool (a
    checksum mismatch is found only after the arrays) must not leave the
    workspace holding the earlier suspended run on damaged bounds: resume
    refuses, and a good file then continues to the uninterrupted result."""
    r0, _ = _uninterrupted(pg)
    path = str(tmp_path / "run.pagani")
    with pg.Workspace(len(A), 0) as ws:
        ws.set_device_mem(MEM)
        ws.set_iteration_limit(4)
        ws.integrate_genz(FAM, A, U, epsrel=EPS, epsabs=1e-20)
        assert ws.suspended()
        ws.save_state(path)
        bad = bytearray(open(path, "rb").read())
        bad[len(bad) - 64] ^= 0x10
        p_bad = str(tmp_path / "bad.pagani")
        open(p_bad, "wb").write(bytes(bad))
        with pytest.raises(pg.PaganiError, match="checksum"):
            ws.load_state(p_bad)
        assert not ws.suspended()
        with pytest.raises(pg.PaganiError, match="no suspended run"):
            ws.resume_genz(FAM, A, U)
        ws.load_state(path)
        assert ws.suspended()
        again, rounds = True, 0
        while again:
            r, again = ws.resume_genz(FAM, A, U)
            rounds += 1
            assert rounds < 100
    assert _key(r) == _key(r0)


def test_regions_csv(pg, gpu, tmp_path):
    d = 4
    a, u = [8.0, 9.0, 10.0, 11.0], [0.4, 0.5, 0.6, 0.3]
    with pg.Workspace(d, 0) as ws:
        r = ws.integrate_genz(FAM, a, u, epsrel=1e-4, epsabs=1e-20)
        assert r.status == 0
        tr = ws.trace()
        p = str(tmp_path / "regions.csv")
        ws.write_regions_csv(p)
    rows = np.loadtxt(p, delimiter=",", ndmin=2)
    n_last = int(tr["n"][-1])
    assert rows.shape == (n_last, 2 + 2 * d + 2)
    up, lo = rows[:, 2:2 + 2 * d:2], rows[:, 3:3 + 2 * d:2]
    assert np.all(lo >= 0) and np.all(up <= 1) and np.all(up > lo)
    assert np.all((rows[:, -2] >= 0) & (rows[:, -2] < d))
    assert np.all(rows[:, -1] == -1)
    assert np.all(rows[:, 1] >= 0)
    # the regions of the last iteration are disjoint boxes of the unit cube
    total_vol = float(np.sum(np.prod(up - lo, axis=1)))
    assert 0 < total_vol <= 1 + 1e-12
    # the estimates are the integrand's: positive for the Gaussian, and at
    # most the whole integral
    assert np.all(rows[:, 0] >= 0) and rows[:, 0].sum() <= r.estimate * (1 + 1e-12)


def test_snapshot_iteration(pg, gpu):
    with pg.Workspace(len(A), 0) as ws:
        ws.set_device_mem(MEM)
        ws.set_snapshot_iter(2)
        ws.integrate_genz(FAM, A, U, epsrel=EPS, epsabs=1e-20, relerr=False)
        tr = ws.trace()
        n = ws.snapshot_size()
    finished_before = sum(int(a) - int(b) for a, b in zip(tr["n"][:2], tr["m"][:2]))
    assert len(tr["n"]) > 2
    assert n == (int(tr["n"][2]) if finished_before == 0 else 0)


def test_predict_split_cpp_overload(pg, gpu):
    class _Res(ctypes.Structure):
        _fields_ = [("estimate", ctypes.c_double), ("errorest", ctypes.c_double),
                    ("neval", ctypes.c_size_t), ("nregions", ctypes.c_size_t),
                    ("nFinishedRegions", ctypes.c_size_t), ("status", ctypes.c_int),
                    ("lastPhase", ctypes.c_int), ("chi_sq", ctypes.c_double),
                    ("iters", ctypes.c_size_t)]
    lib = ctypes.CDLL(REFTESTS)
    cap = 800
    tn = (ctypes.c_uint64 * cap)()
    tm = (ctypes.c_uint64 * cap)()
    r1, r2 = _Res(), _Res()
    snap = ctypes.c_size_t(0)
    ntr = ctypes.c_size_t(0)
    rc = lib.reftest_predict_split(b"F_4_5D", 2, ctypes.c_double(1e-7), ctypes.c_double(1e-20),
                                   0, ctypes.byref(r1), ctypes.byref(snap), tn, tm,
                                   ctypes.c_size_t(cap), ctypes.byref(ntr), ctypes.byref(r2))
    assert rc == 0
    n = ntr.value
    finished = sum(tn[i] - tm[i] for i in range(min(15, n)))
    if n > 15 and finished == 0:
        assert snap.value == tn[15]
        # restarting from the snapshot reaches the same answer within errors
        assert r2.status == 0
        assert abs(r2.estimate - r1.estimate) <= r1.errorest + r2.errorest
    else:
        assert snap.value == 0


def _csv_regions(path, d):
    rows = np.loadtxt(path, delimiter=",", ndmin=2)
    est, err = rows[:, 0], rows[:, 1]
    up, lo = rows[:, 2:2 + 2 * d:2], rows[:, 3:3 + 2 * d:2]
    return est, err, lo, up, rows[:, -2].astype(int)


def test_suspended_list_matches_oracle(pg, gpu, tmp_path):
    """A run suspended after k iterations holds the oracle's list for
    iteration k (Sub_regions.cuh:220-239 keeps such a list; quad.h:107-118
    writes it): the same regions in the same order, bounds bit for bit (the
    CSV carries 17 significant digits, which round-trip a double); and the
    list a finished run ends with carries the oracle's per-region estimates
    at the kernel's points within 1e-13 of their scale."""
    import oracle_api as O
    # distinct scales and centres per axis: no region's split axis is a tie
    # decided by rounding (the oracle maps points as the reference does, the
    # kernel a few roundings away)
    d, a, u, eps = 4, [8.0, 9.5, 11.0, 12.5], [0.41, 0.53, 0.62, 0.37], 1e-5
    prm = O.params(a, u)
    full, tr = O.integrate(d, "gaussian", prm, eps, 1e-20, trace_cap=800, keep_list=True)
    iters = len(tr["n"])
    assert iters > 4
    for k in (1, 3, iters - 1):
        o = O.integrate(d, "gaussian", prm, eps, 1e-20, max_iters=k, keep_list=True)
        lst = o.final_list
        with pg.Workspace(d, 0) as ws:
            ws.set_device_mem(MEM)
            ws.set_iteration_limit(k)
            ws.integrate_genz("gaussian", a, u, epsrel=eps, epsabs=1e-20)
            assert ws.suspended()
            p = str(tmp_path / f"susp{k}.csv")
            ws.write_regions_csv(p)
        est, err, lo, up, dim = _csv_regions(p, d)
        assert len(est) == lst["n"]
        assert np.all(np.isnan(est)) and np.all(dim == -1)
        oL = lst["L"].reshape(d, -1).T
        oU = oL + lst["H"].reshape(d, -1).T
        assert np.array_equal(lo, oL), f"iteration {k}: lower bounds differ"
        assert np.array_equal(up, oU), f"iteration {k}: upper bounds differ"
    # the final, evaluated list of the uninterrupted run
    with pg.Workspace(d, 0) as ws:
        ws.set_device_mem(MEM)
        r = ws.integrate_genz("gaussian", a, u, epsrel=eps, epsabs=1e-20)
        p = str(tmp_path / "final.csv")
        ws.write_regions_csv(p)
    assert r.status == full.status and r.nregions == full.nregions
    lst = full.final_list
    est, err, lo, up, dim = _csv_regions(p, d)
    oL = lst["L"].reshape(d, -1).T
    assert np.array_equal(lo, oL) and np.array_equal(up, oL + lst["H"].reshape(d, -1).T)
    oe, oerr, odim, fd, rs = O.eval_regions_kform(d, "gaussian", prm, lst["L"], lst["H"],
                                                  lst["n"])
    assert np.all(np.abs(est - oe) <= 1e-13 * O.est_scale(rs) + 1e-300)
