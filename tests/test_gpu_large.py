"""The memory-bound workload at full size: 8D cos(sum (i+1) x_i) and the 8D
discontinuous Genz integrand at epsrel 1e-8 with the library's own memory
model, region pools of several hundred million regions (a few hundred GB of
HBM).

- cos(sum (i+1) x_i): the pool grows until the memory model stops the run
  (status 1): the last list is past the point where splitting it would need
  more than half the device (mem_required > total / 2, the model's ratio
  above 1, heuristic_classifier.cuh:96-133 with this library's per-region
  bytes); the estimate is within its error estimate of the closed form
  3.439557952183252e-05 (integrands.cuh:1322).
- discontinuous: at epsrel 1e-2 (the oracle's loop finishes in seconds),
  the same per-iteration region counts as the oracle and an estimate within
  the combined error estimates -- and both about 16 % below the closed
  form, while their error estimates say under 1 %: the degree-9 rule's
  error estimate does not see the jump, on the CPU restatement as on the
  GPU.  At epsrel 1e-8 the GPU run converges with the same deficit.
"""
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu

COS8_EXACT = 3.439557952183252e-05


def _hbm():
    import torch
    return torch.cuda.get_device_properties(0).total_memory


def _disc_params(seed=2024):
    from gpuintegration_amd.workloads import MT19937_64
    g = MT19937_64(seed)
    a = [5.0 + 10.0 * g.uniform() for _ in range(8)]
    u = [0.3 + 0.4 * g.uniform() for _ in range(8)]
    return a, u


def test_cos_sum_8d_full_pool(pg, gpu):
    d = 8
    a, u = [float(i + 1) for i in range(d)], [0.0] * d
    assert abs(O.exact("oscillatory", d, O.params(a, u)) - COS8_EXACT) <= 1e-15
    with pg.Workspace(d, 0) as ws:
        ws.set_mem_model("actual")
        r = ws.integrate_genz("oscillatory", a, u, epsrel=1e-8, epsabs=1e-20)
        tr = ws.trace()
        total = ws.device_mem()
    n = [int(v) for v in tr["n"]]
    print(f"cos8: status {r.status}, iterations {len(n)}, largest list {max(n)}, "
          f"nregions {r.nregions}, estimate {r.estimate!r}, errorest {r.errorest:.3g}")
    assert max(n) >= 300_000_000
    assert r.status == 1
    per_region = 24 * d + 64  # the library's footprint model (PAGANI_MEM_ACTUAL)
    assert per_region * n[-1] > total / 2
    assert abs(r.estimate - COS8_EXACT) <= r.errorest
    assert abs(r.estimate - COS8_EXACT) <= 1e-9 * COS8_EXACT


def test_discontinuous_8d_deficit(pg, gpu):
    d = 8
    a, u = _disc_params()
    exact = O.exact("discontinuous", d, O.params(a, u))
    hbm = _hbm()
    o, otr = O.integrate(d, "discontinuous", O.params(a, u), 1e-2, 1e-20, total_mem=hbm,
                         trace_cap=800)
    with pg.Workspace(d, 0) as ws:
        ws.set_device_mem(hbm)
        r = ws.integrate_genz("discontinuous", a, u, epsrel=1e-2, epsabs=1e-20)
        tr = ws.trace()
        ws.set_mem_model("actual")
        ws.set_device_mem(0)
        r8 = ws.integrate_genz("discontinuous", a, u, epsrel=1e-8, epsabs=1e-20)
    assert list(tr["n"]) == [int(v) for v in otr["n"]]
    assert (r.status, r.nregions, r.iters) == (o.status, o.nregions, o.iters)
    assert abs(r.estimate - o.estimate) <= r.errorest + o.errorest
    for e, est in (("gpu", r), ("oracle", o)):
        deficit = 1 - est.estimate / exact
        assert 0.14 < deficit < 0.18, (e, deficit)
        assert est.errorest < 0.01 * exact, e
    assert r8.status == 0
    assert 0.14 < 1 - r8.estimate / exact < 0.18
    print(f"discontinuous 8D: exact {exact!r}; oracle/gpu at 1e-2 "
          f"{o.estimate!r}/{r.estimate!r}; gpu at 1e-8 {r8.estimate!r} "
          f"(errorest {r8.errorest:.3g}, {r8.nregions} regions)")
