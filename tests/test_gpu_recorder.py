"""The debug recorders (pagani_ws_set_debug; the reference's
Workspace<T, ndim, debug, use_custom, collect_mult_runs>, Workspace.cuh:
171-179, 199-352, 377-431, and Cubature_rules, PaganiUtils.cuh:50-241).

The files are views of the loop, so each is checked against something
computed apart from it: the per-iteration rows against the workspace's trace
and the oracle's loop (region counts per iteration, running totals), the
per-region rows against the evaluation kernel run on the same list, and the
rule-point dump against the rule: every region's estimate is the region's
volume times the weighted sum of its recorded point values (the weights of
pagani_rule_tables, the sets of the oracle's generator list), which pins the
point order, the points and their values together.  Recording must leave
the result bit-identical.
"""
import csv
import ctypes
import math
import os

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REFTESTS = os.path.join(HERE, "cxx", "libreftests.so")

PHASES = ["apply_cubature_rules", "two_level_errorest", "compute_finished_estimates",
          "fix_error_budget_overflow", "heuristic_classify", "region_filtering",
          "region_splitting"]

# F_2_6D (Easy_Integrals / Genz KATs): eight iterations at 1e-3
KAT = ("product_peak", 6, [50.0] * 6, [0.5] * 6)


def read_csv(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return [c.strip() for c in rows[0]], rows[1:]


def run(pg, d, fam, a, u, epsrel, debug=None, init_k=0):
    with pg.Workspace(d, 0) as ws:
        ws.set_device_mem(int(16e9))
        if debug:
            ws.set_debug(**debug)
        r = ws.integrate_genz(fam, a, u, epsrel=epsrel, epsabs=1e-12, init_k=init_k)
        tr = ws.trace()
    return r, tr


def test_iteration_and_phase_files(pg, gpu, tmp_path):
    fam, d, a, u = KAT
    r0, tr0 = run(pg, d, fam, a, u, 1e-3)
    r, tr = run(pg, d, fam, a, u, 1e-3, debug=dict(level=1, directory=str(tmp_path), id="kat"))
    # recording changes nothing
    assert (r.estimate, r.errorest, r.nregions, r.iters, r.status) == \
        (r0.estimate, r0.errorest, r0.nregions, r0.iters, r0.status)
    assert list(tr["n"]) == list(tr0["n"])
    head, rows = read_csv(tmp_path / "cuda_pagani_iters.csv")
    assert head == ["it", "estimate", "errorest", "festimate", "ferrorest", "nregions",
                    "nfregions"]
    k = len(tr["n"])
    assert len(rows) == k and k >= 3
    it = [int(x[0]) for x in rows]
    est = np.array([float(x[1]) for x in rows])
    err = np.array([float(x[2]) for x in rows])
    fest = np.array([float(x[3]) for x in rows])
    n = [int(x[5]) for x in rows]
    nf = [int(x[6]) for x in rows]
    assert it == list(range(k))
    assert n == list(tr["n"])
    # the running totals are the trace's, to the bit
    assert np.array_equal(est, tr["estimate"]) and np.array_equal(err, tr["errorest"])
    # regions finished before each iteration: the previous lists minus what
    # they kept
    assert nf[0] == 0
    for i in range(1, k):
        assert nf[i] == nf[i - 1] + int(tr["n"][i - 1]) - int(tr["m"][i - 1])
    assert r.nregions == nf[-1] + n[-1]
    assert fest[0] == 0.0 and est[-1] == r.estimate
    # the oracle's loop: same list sizes, totals within the combined errors
    o, otr = O.integrate(d, fam, O.params(a, u), 1e-3, 1e-12, trace_cap=800)
    assert list(otr["n"]) == n
    assert np.all(np.abs(est - np.asarray(otr["estimate"])) <= err + np.asarray(otr["errorest"]))

    head, rows = read_csv(tmp_path / "cuda_pagani_time_breakdown.csv")
    assert head == ["id", "ndim", "epsrel", "it", "name", "time"]
    by_it = {}
    for row in rows:
        assert row[0] == "kat" and int(row[1]) == d and float(row[2]) == 1e-3
        assert float(row[5]) >= 0.0
        by_it.setdefault(int(row[3]), []).append((row[4], float(row[5])))
    assert sorted(by_it) == list(range(k))
    for i in range(k - 1):
        assert [p for p, _ in by_it[i]] == PHASES
        assert dict(by_it[i])["region_splitting"] == 0.0  # folded into the filter's kernel
        assert dict(by_it[i])["apply_cubature_rules"] > 0.0
    # accuracy reached after the refinement of the last iteration
    assert [p for p, _ in by_it[k - 1]] == PHASES[:2]
    assert not (tmp_path / "cuda_regions.csv").exists()


def test_plain_overload_file_and_append(pg, gpu, tmp_path):
    fam, d, a, u = KAT
    dbg = dict(level=1, directory=str(tmp_path), id=None)
    r, tr = run(pg, d, fam, a, u, 1e-3, debug=dbg)
    run(pg, d, fam, a, u, 1e-3, debug=dict(dbg, append=True))
    head, rows = read_csv(tmp_path / "cuda_iters.csv")
    assert head == ["it", "estimate", "errorest", "nregions"]
    k = len(tr["n"])
    assert len(rows) == 2 * k  # the second call appended, without a header
    for rr in (rows[:k], rows[k:]):
        assert [int(x[3]) for x in rr] == list(tr["n"])
        assert np.array_equal([float(x[1]) for x in rr], tr["estimate"])
    assert not (tmp_path / "cuda_pagani_time_breakdown.csv").exists()
    assert not (tmp_path / "cuda_pagani_iters.csv").exists()


def test_region_rows_are_the_kernels_output(pg, gpu, tmp_path):
    import torch
    fam, d, a, u = "gaussian", 3, [5.0, 6.0, 7.0], [0.4, 0.5, 0.6]
    r, tr = run(pg, d, fam, a, u, 1e-4, init_k=4,
                debug=dict(level=2, directory=str(tmp_path), id="r"))
    _, rows = read_csv(tmp_path / "cuda_pagani_iters.csv")
    with open(tmp_path / "cuda_regions.csv") as f:
        reg = [line.strip().split(",") for line in f if line.strip()]
    # no header line (Print_region_evals writes data only); one row per region
    # per iteration, in list order
    assert len(reg) == int(np.sum(tr["n"]))
    first = [x for x in reg if x[0] == "0"]
    assert [int(x[1]) for x in first] == list(range(4 ** d))
    # iteration 0 is the uniform 4^3 split: the same kernel on the same list
    dev = torch.device("cuda:0")
    n = 4 ** d
    L = torch.empty(n * d, dtype=torch.float64, device=dev)
    H = torch.empty_like(L)
    est = torch.empty(n, dtype=torch.float64, device=dev)
    err = torch.empty_like(est)
    dim = torch.empty(n, dtype=torch.uint8, device=dev)
    with pg.Workspace(d, 0) as ws:
        ws.uniform_split(4, L, H)
        ws.eval_genz(fam, a, u, n, L, H, est, err, dim)
    torch.cuda.synchronize()
    assert np.array_equal([float(x[2]) for x in first], est.cpu().numpy())
    assert np.array_equal([float(x[3]) for x in first], err.cpu().numpy())


def gaussian(x, a, u):
    return math.exp(-sum(ai * ai * (xi - ui) ** 2 for xi, ai, ui in zip(x, a, u)))


def test_point_dump_reproduces_every_estimate(pg, gpu, tmp_path):
    fam, d, a, u = "gaussian", 3, [5.0, 6.0, 7.0], [0.4, 0.5, 0.6]
    r, tr = run(pg, d, fam, a, u, 1e-3, init_k=2,
                debug=dict(level=3, directory=str(tmp_path), id="p"))
    assert int(np.max(tr["n"])) < 1024  # every list is dumped
    w = pg.rule_tables(d)[0]
    gens, sets = O.generators(d)
    F = len(gens)
    # the generator table: i = feval * axis + point
    head, rows = read_csv(tmp_path / "cuda_generators.csv")
    assert head == ["i", "gen"] and len(rows) == F * d
    for i, g in rows:
        assert float(g) == gens[int(i) % F][int(i) // F]
    head, rows = read_csv(tmp_path / "cuda_fevals.csv")
    assert head[:2] == ["reg", "fid"] and head[-3:] == ["feval", "estimate", "errorest"]
    assert len(head) == 2 + 5 * d + 3
    assert len(rows) == F * int(np.sum(tr["n"]))
    checked = 0
    for s in range(0, len(rows), F):
        block = rows[s:s + F]
        assert [int(x[1]) for x in block] == list(range(F))
        vals = np.array([float(x[2 + 5 * d]) for x in block])
        lo = np.array([float(block[0][2 + 2 * j]) for j in range(d)])
        hi = np.array([float(block[0][3 + 2 * j]) for j in range(d)])
        est = float(block[0][-2])
        vol = float(np.prod(hi - lo))
        terms = np.array([w[sets[p]][0] * vals[p] for p in range(F)])
        recon = vol * terms.sum()
        assert abs(recon - est) <= 1e-13 * vol * np.abs(terms).sum() + 1e-300, (s, recon, est)
        # each value is the integrand at the recorded point (box: the unit cube)
        for p in (0, 1, 2 * d + 1, F - 1):
            x = [float(block[p][2 + 4 * d + j]) for j in range(d)]
            assert [float(block[p][2 + 2 * d + 2 * j + 1]) for j in range(d)] == [1.0] * d
            assert math.isclose(vals[p], gaussian(x, a, u), rel_tol=1e-12, abs_tol=1e-300)
            # the point lies in its region
            assert all(lo[j] <= x[j] <= hi[j] for j in range(d))
        checked += 1
    assert checked == int(np.sum(tr["n"]))


class _Res(ctypes.Structure):
    _fields_ = [("estimate", ctypes.c_double), ("errorest", ctypes.c_double),
                ("neval", ctypes.c_size_t), ("nregions", ctypes.c_size_t),
                ("nFinishedRegions", ctypes.c_size_t), ("status", ctypes.c_int),
                ("lastPhase", ctypes.c_int), ("chi_sq", ctypes.c_double),
                ("iters", ctypes.c_size_t)]


def test_header_template_parameters_turn_recording_on(pg, gpu, tmp_path, monkeypatch):
    lib = ctypes.CDLL(REFTESTS)
    lib.reftest_debug_header.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(_Res)]
    monkeypatch.chdir(tmp_path)
    r = _Res()
    # Sub_regions& overload, debug 1, optional "hdr"
    assert lib.reftest_debug_header(0, 1e-6, ctypes.byref(r)) == 0
    assert r.status == 0 and abs(r.estimate - 0.25) <= 1e-6 * 0.25
    _, rows = read_csv(tmp_path / "cuda_pagani_iters.csv")
    assert len(rows) >= 1 and int(rows[0][5]) == 16  # 4 per axis
    _, times = read_csv(tmp_path / "cuda_pagani_time_breakdown.csv")
    assert {t[0] for t in times} == {"hdr"} and {t[4] for t in times} <= set(PHASES)
    # the plain overload with collect_mult_runs: cuda_iters.csv, appended
    for _ in range(2):
        assert lib.reftest_debug_header(1, 1e-6, ctypes.byref(r)) == 0
    head, rows = read_csv(tmp_path / "cuda_iters.csv")
    assert head[0] != "it"  # collect_mult_runs writes no header
    assert len(rows) % 2 == 1  # the first data row stands where the header would
    # debug 3: regions and rule points
    assert lib.reftest_debug_header(2, 1e-6, ctypes.byref(r)) == 0
    assert (tmp_path / "cuda_regions.csv").stat().st_size > 0
    head, rows = read_csv(tmp_path / "cuda_fevals.csv")
    assert len(head) == 2 + 5 * 2 + 3 and len(rows) > 0
