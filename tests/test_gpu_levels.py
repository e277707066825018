"""Widths held as levels (pagani_ws_set_level_form; pagani_eval_args).

The library's uniform start keeps every region's widths as the number of
halvings of 1/k, two bytes per axis, instead of doubles.  A width is then
h0 * 2^-level, the same double that repeated halving gives, so the form is
invisible in the results: every integral, trajectory, region list and
checkpoint must be the same bits with levels on and off (the form with
double widths is the one the oracle comparisons elsewhere pin).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("gaussian", 8, None, 1e-4, 4),  # the benchmark's integral, shorter
    ("discontinuous", 6, ([10, 9, 8, 7, 6, 5], [.9, .8, .7, .6, .5, .4]), 1e-3, 0),
    ("oscillatory", 5, ([1.1, 1.3, 1.2, 1.4, 1.0], [0.3] * 5), 1e-6, 0),
    ("product_peak", 3, ([5.0, 6.0, 7.0], [0.4, 0.5, 0.6]), 1e-7, 0),
    ("c0", 10, ([2.0] * 10, [0.5] * 10), 1e-2, 0),
]


def _params(fam, d, prm):
    if prm is None:
        from gpuintegration_amd import workloads as W
        _, a, u = W.gaussian_8d(2024)
        return a, u
    return prm


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1]}D" for c in CASES])
@pytest.mark.parametrize("device_loop", [True, False])
def test_levels_give_the_same_bits(pg, gpu, case, device_loop):
    fam, d, prm, epsrel, k = case
    a, u = _params(fam, d, prm)
    out = []
    for levels in (True, False):
        with pg.Workspace(d, 0) as ws:
            ws.set_device_mem(int(16e9))
            ws.set_level_form(levels)
            ws.set_device_loop(device_loop)
            r = ws.integrate_genz(fam, a, u, epsrel=epsrel, epsabs=1e-20, init_k=k)
            out.append((r, ws.trace()))
    (r1, t1), (r2, t2) = out
    assert (r1.estimate, r1.errorest, r1.nregions, r1.iters, r1.status) == \
        (r2.estimate, r2.errorest, r2.nregions, r2.iters, r2.status)
    for key in ("n", "m", "estimate", "errorest"):
        assert np.array_equal(t1[key], t2[key]), key


def test_suspended_lists_and_checkpoints_are_the_same(pg, gpu, tmp_path):
    """A run suspended after 6 iterations writes the same region list, and its
    checkpoint (which holds double widths either way) resumes to the result
    of the uninterrupted run."""
    from gpuintegration_amd import workloads as W
    _, a, u = W.gaussian_8d(2024)
    csv = {}
    for levels in (True, False):
        with pg.Workspace(8, 0) as ws:
            ws.set_level_form(levels)
            ws.set_iteration_limit(6)
            ws.integrate_genz("gaussian", a, u, epsrel=1e-4, epsabs=1e-20, init_k=4)
            assert ws.suspended()
            path = tmp_path / f"list_{levels}.csv"
            ws.write_regions_csv(str(path))
            csv[levels] = path.read_bytes()
            ws.save_state(str(tmp_path / f"state_{levels}.bin"))
    assert csv[True] == csv[False] and len(csv[True]) > 0
    with pg.Workspace(8, 0) as ws:
        full = ws.integrate_genz("gaussian", a, u, epsrel=1e-4, epsabs=1e-20, init_k=4)
    for levels in (True, False):
        with pg.Workspace(8, 0) as ws:
            ws.load_state(str(tmp_path / f"state_{levels}.bin"))
            r, again = ws.resume_genz("gaussian", a, u)
        assert not again
        assert (r.estimate, r.errorest, r.nregions) == (full.estimate, full.errorest, full.nregions)
    # and in memory: suspend in level form, resume in the same workspace
    with pg.Workspace(8, 0) as ws:
        ws.set_iteration_limit(6)
        ws.integrate_genz("gaussian", a, u, epsrel=1e-4, epsabs=1e-20, init_k=4)
        ws.set_iteration_limit(0)
        r, again = ws.resume_genz("gaussian", a, u)
    assert (r.estimate, r.errorest, r.nregions) == (full.estimate, full.errorest, full.nregions)
