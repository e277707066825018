"""Benchmark: PAGANI on one MI355X.

Default (--config 2, the headline line): the seeded 8D Genz Gaussian
f(x) = exp(-sum a_i^2 (x_i - u_i)^2), a_i = 10 (1 + 0.05 (2U - 1)),
u_i = 0.2 + 0.6 U (std::mt19937_64, U = (x >> 11) 2^-53), integrated to
epsrel 1e-6 from an initial uniform split of 4 per axis (65,536 regions),
FP64.  One step = one full integration to convergence; inputs are synthetic
and resident on the device; the whole loop (evaluation, refinement,
classification, filtering, splitting) is inside the timed region.
value = regions evaluated per second.

Other configurations (each prints its own JSON line):
  --config 1  5D oscillatory, epsrel 1e-3: the sequential Cuhre-equivalent on
              one host core (max 1e6 evaluations) beside the GPU
  --config 3  mixed precision: 6D corner peak and product peak, epsrel
              1e-3..1e-7, FP32 integrand vs FP64, error vs analytic per run
  --config 5  memory-bound: 8D cos(sum (i+1) x_i) and 8D discontinuous at
              epsrel 1e-8 with the library's own memory model

The headline line also carries "mcubes": the m-Cubes VEGAS integrator on the
same 8D Gaussian (1e9 calls per iteration), samples/s and its kernel's
algorithmic FP64 rate (--no-mcubes leaves it out, --mcubes-only runs it
alone).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "regions/s + integrals/s to epsrel 1e-6 (8D Genz), mixed vs FP64, 1 GPU"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes per region of the 8D Gaussian evaluation kernel, from two
# rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over `bench.py
# --eval-only`, each scaled by what the counter reads for a 1 GiB device
# copy: 128.20 B read + 17.00 B written (algorithmic: 128 + 17).
EVAL_HBM_BYTES_PER_REGION_8D = 145.20
# FP64 work the 8D Gaussian evaluation kernel executes per region, from one
# rocprofv3 counter pass over `bench.py --eval-only` (SQ_INSTS_VALU_FMA_F64,
# _MUL_F64, _ADD_F64 and SQ_INSTS_VALU, one lane per region, every lane
# active): 1366 FMA + 233 MUL + 128 ADD = 3093 FLOP and 2221 VALU
# instructions.  The algorithmic count (90770 FLOP) charges every one of the
# 1105 points a full d-axis evaluation; the kernel evaluates the Gaussian as
# a product over axes -- the centre's value times one factor e^-(change of
# a_j^2 (x_j - u_j)^2) per changed axis, 81 exponentials per region instead
# of 1105, and each set's sum over its sign combinations factored per axis
# -- so the algorithmic rate would exceed the FP64 peak: the roofline uses
# the executed FLOPs.
EVAL_EXEC_FLOP_PER_REGION_8D = 3093
EVAL_VALU_PER_REGION_8D = 2221
VALU_ISSUE_PEAK = 256 * 2.4e9  # wave-instructions per second (one per CU per clock)


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def noneval_bytes(trace, d):
    """Algorithmic HBM bytes of refine + compaction + split over a run:
    n (16d + 33) + m (32d + 16) per iteration."""
    return sum(int(n) * (16 * d + 33) + int(m) * (32 * d + 16)
               for n, m in zip(trace["n"], trace["m"]))


def run_gpu(pg, fam, a, u, epsrel, steps, warmup, precision, init_k=0, mem_model=None,
            epsabs=1e-20):
    d = len(a)
    ws = pg.Workspace(d, 0)
    ws.set_timing(True)
    if mem_model is not None:
        ws.set_mem_model(mem_model)
    for _ in range(warmup):
        ws.integrate_genz(fam, a, u, epsrel=epsrel, epsabs=epsabs, precision=precision,
                          init_k=init_k)
    _sync()
    ws.reset_stats()
    t0 = time.perf_counter()
    res = None
    for _ in range(steps):
        res = ws.integrate_genz(fam, a, u, epsrel=epsrel, epsabs=epsabs, precision=precision,
                                init_k=init_k)
    _sync()
    dt = time.perf_counter() - t0
    st = ws.stats()
    tr = ws.trace()
    ws.close()
    return res, dt, st, tr


def phase_report(W, fam, d, res, dt, st, tr, steps):
    regions = st["regions_evaluated"]
    eval_s = st["eval_ms"] / 1e3
    # eval_ms times the host-driven evaluation launches only; the device
    # loop's iterations are timed as whole rounds (device_ms)
    flops = W.eval_flops_per_region(fam, d) * (regions - st["device_regions"])
    ne_bytes = noneval_bytes(tr, d) * steps
    ne_s = (st["refine_ms"] + st["compact_ms"]) / 1e3
    return {
        "regions_per_s": regions / dt,
        "integrals_per_s": steps / dt,
        "ms_per_integral": dt / steps * 1e3,
        "estimate": res.estimate,
        "errorest": res.errorest,
        "status": res.status,
        "nregions": res.nregions,
        "iters": res.iters,
        "regions_evaluated_per_integral": regions // steps,
        "max_regions": st["max_regions"],
        "eval_tflops": flops / eval_s / 1e12 if eval_s > 0 else None,
        "eval_frac_fp64_peak": flops / eval_s / 1e12 / FP64_PEAK_TFLOPS if eval_s > 0 else None,
        "noneval_gbs": ne_bytes / ne_s / 1e9 if ne_s > 0 else None,
        "noneval_frac_hbm": ne_bytes / ne_s / 1e9 / HBM_PEAK_GBS if ne_s > 0 else None,
        "phase_ms_per_integral": {k: st[k] / steps for k in
                                  ("eval_ms", "refine_ms", "compact_ms", "classify_ms")},
        "classify_passes_per_integral": st["classify_passes"] / steps,
        "device_loop": {"ms": st["device_ms"] / steps,
                        "iterations": st["device_iterations"] / steps,
                        "rounds": st["device_rounds"] / steps},
        "alloc_ms_per_integral": st["alloc_ms"] / steps,
        "allocations_per_integral": st["allocations"] / steps,
    }


def cpu_baseline(fam, a, u, budget_s):
    """The oracle's region evaluation on the host cores: repeated passes over
    the workload's first iteration (the 4^8 uniform split) for ~budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as O
    d = len(a)
    L, H, n = O.uniform_split(d, 4)
    prm = O.params(a, u)
    O.eval_regions(d, fam, prm, L, H, 4096)  # page in
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        O.eval_regions(d, fam, prm, L, H, n)
        done += n
    dt = time.perf_counter() - t0
    return {
        "value": done / dt,
        "unit": "regions/s",
        "cores": int(O.LIB.orc_num_threads()),
        "kind": "port",
        "sample": f"oracle rule evaluation (OpenMP, {O.LIB.orc_num_threads()} threads) of the "
                  f"workload's first iteration, {n} regions x {done // n} passes in {dt:.1f} s",
    }


def config2(args, pg, W):
    fam, a, u = W.gaussian_8d(args.seed)
    d = len(a)
    exact = W.gaussian_exact(a, u)
    res, dt, st, tr = run_gpu(pg, fam, a, u, args.epsrel, args.steps, args.warmup, pg.FP64,
                              init_k=4)
    rep = phase_report(W, fam, d, res, dt, st, tr, args.steps)
    regions = st["regions_evaluated"]
    # the roofline covers the host-driven evaluation launches (timed one by
    # one); the device loop's small iterations are reported beside it
    host_regions = regions - st["device_regions"]
    host_launches = st["eval_launches"] - st["device_iterations"]
    eval_s = st["eval_ms"] / 1e3
    achieved = EVAL_EXEC_FLOP_PER_REGION_8D * host_regions / eval_s / 1e12 if eval_s else 0.0
    valu_frac = (EVAL_VALU_PER_REGION_8D / 64 * host_regions / eval_s / VALU_ISSUE_PEAK
                 if eval_s else 0.0)
    out = {
        "metric": METRIC,
        "value": regions / dt,
        "unit": "regions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Genz parameters)",
        "config": {
            "workload": "PAGANI, 8D Genz Gaussian exp(-sum a_i^2 (x_i-u_i)^2), "
                        "a_i=10(1+0.05(2U-1)), u_i=0.2+0.6U, epsrel 1e-6, initial 4^8 regions",
            "seed": args.seed,
            "epsrel": args.epsrel,
            "precision": "fp64",
        },
        "integrals_per_s": args.steps / dt,
        "estimate": res.estimate,
        "errorest": res.errorest,
        "analytic": exact,
        "abs_err_vs_analytic": abs(res.estimate - exact),
        "status": res.status,
        "nregions": res.nregions,
        "iters": res.iters,
        "regions_evaluated_per_step": regions // args.steps,
        "max_regions": st["max_regions"],
        "phase_ms_per_step": rep["phase_ms_per_integral"],
        "classify_passes_per_step": rep["classify_passes_per_integral"],
        "noneval_gbs": rep["noneval_gbs"],
        "roofline": {
            "bound": "fp64-valu",
            "kernel": "pagani::eval_kernel<8, Gaussian<double,8>, double>",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "flops_per_region": EVAL_EXEC_FLOP_PER_REGION_8D,
            "flops_per_region_kind": "executed FP64 FLOPs (counter-measured)",
            "algorithmic_flops_per_region": W.eval_flops_per_region(fam, d),
            "algorithmic_tflops": rep["eval_tflops"],
            "valu_issue_frac": valu_frac,
            "eval_launches": host_launches,
            "avg_launch_ms": st["eval_ms"] / max(host_launches, 1),
            "avg_regions_per_launch": host_regions / max(host_launches, 1),
            "traffic": EVAL_HBM_BYTES_PER_REGION_8D * host_regions / max(host_launches, 1),
            "traffic_unit": "bytes per launch (counter-measured per-region HBM bytes x "
                            "average regions per launch)",
            "algorithmic_bytes_per_launch": (16 * d + 17) * host_regions / max(host_launches, 1),
        },
    }
    if not args.no_mixed:
        # mixed precision beside FP64 at one stated tolerance both can certify
        eps_m = args.mixed_epsrel
        r32, dt32, st32, tr32 = run_gpu(pg, fam, a, u, eps_m, args.steps, args.warmup, pg.FP32,
                                        init_k=4)
        r64, dt64, st64, tr64 = run_gpu(pg, fam, a, u, eps_m, args.steps, args.warmup, pg.FP64,
                                        init_k=4)
        out["mixed"] = {
            "epsrel": eps_m,
            "fp32": dict(precision="fp32 integrand + fp64 sums/errors",
                         regions_per_s=st32["regions_evaluated"] / dt32,
                         integrals_per_s=args.steps / dt32, status=r32.status,
                         abs_err_vs_analytic=abs(r32.estimate - exact), errorest=r32.errorest,
                         nregions=r32.nregions),
            "fp64": dict(regions_per_s=st64["regions_evaluated"] / dt64,
                         integrals_per_s=args.steps / dt64, status=r64.status,
                         abs_err_vs_analytic=abs(r64.estimate - exact), errorest=r64.errorest,
                         nregions=r64.nregions),
        }
    if not args.no_mcubes:
        out["mcubes"] = mcubes_leg(args, pg, W)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(fam, a, u, args.cpu_seconds)
    return out


def config1(args, pg, W):
    """5D oscillatory, epsrel 1e-3, epsabs 1e-20, max 1e6 evaluations."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as O
    fam, a, u = W.oscillatory_5d(args.seed)
    d = len(a)
    prm = O.params(a, u)
    exact = W.exact(fam, a, u)
    # CPU baseline: the oracle's sequential Cuhre-equivalent on one core
    O.LIB.orc_set_num_threads(1)
    t0 = time.perf_counter()
    reps = 0
    while True:
        rc = O.cuhre_seq(d, fam, prm, 1e-3, 1e-20, 1_000_000)
        reps += 1
        if time.perf_counter() - t0 > 3.0:
            break
    cdt = (time.perf_counter() - t0) / reps
    res, dt, st, tr = run_gpu(pg, fam, a, u, 1e-3, args.steps, args.warmup, pg.FP64,
                              epsabs=1e-20)
    rep = phase_report(W, fam, d, res, dt, st, tr, args.steps)
    return {
        "metric": "integrals/s, 5D Genz oscillatory to epsrel 1e-3 (CPU plain method vs GPU)",
        "config": {"workload": "cos(2 pi u1 + sum a_i x_i), a_i=1+0.5U, u1=U, 5D, epsrel 1e-3, "
                               "epsabs 1e-20, max 1e6 evals", "seed": args.seed},
        "analytic": exact,
        "gpu": dict(rep, abs_err_vs_analytic=abs(res.estimate - exact)),
        "cpu_cuhre_equivalent": {
            "cores": 1, "integrals_per_s": 1 / cdt, "ms_per_integral": cdt * 1e3,
            "estimate": rc.estimate, "errorest": rc.errorest, "status": rc.status,
            "neval": rc.neval, "nregions": rc.nregions,
            "abs_err_vs_analytic": abs(rc.estimate - exact),
        },
    }


def config3(args, pg, W):
    rows = []
    for family in ("corner_peak", "product_peak"):
        fam, a, u = W.mixed_6d(family, args.seed)
        d = len(a)
        exact = W.exact(fam, a, u)
        for eps in (1e-3, 1e-4, 1e-5, 1e-6, 1e-7):
            for prec, name in ((pg.FP64, "fp64"), (pg.FP32, "fp32")):
                res, dt, st, tr = run_gpu(pg, fam, a, u, eps, 1, 0, prec)
                rows.append({
                    "family": fam, "epsrel": eps, "precision": name, "status": res.status,
                    "estimate": res.estimate, "errorest": res.errorest,
                    "rel_err_vs_analytic": abs(res.estimate - exact) / abs(exact),
                    "regions_per_s": st["regions_evaluated"] / dt,
                    "ms_per_integral": dt * 1e3, "nregions": res.nregions,
                })
    return {"metric": "mixed precision (FP32 integrand) vs FP64, 6D corner/product peak",
            "config": {"workload": "a_i = 5 + 10U, u_i = U, 6D, epsrel 1e-3..1e-7",
                       "seed": args.seed}, "runs": rows}


def config5(args, pg, W):
    from gpuintegration_amd.workloads import MT19937_64
    g = MT19937_64(args.seed)
    d = 8
    cases = [("oscillatory", [float(i + 1) for i in range(d)], [0.0] * d)]
    a6 = [5.0 + 10.0 * g.uniform() for _ in range(d)]
    u6 = [0.3 + 0.4 * g.uniform() for _ in range(d)]
    cases.append(("discontinuous", a6, u6))
    out = []
    for fam, a, u in cases:
        exact = W.exact(fam, a, u)
        # after one untimed integral on the same workspace the pool is at its
        # full size: the timed integral allocates nothing (a first integral's
        # growth reallocates up to hundreds of GB, whose cost depends on how
        # recently the driver got that memory back)
        res, dt, st, tr = run_gpu(pg, fam, a, u, args.large_epsrel, 1, 1, pg.FP64,
                                  mem_model="actual")
        rep = phase_report(W, fam, d, res, dt, st, tr, 1)
        out.append(dict(rep, family=fam, abs_err_vs_analytic=abs(res.estimate - exact),
                        analytic=exact))
    return {"metric": "memory-bound phases: non-eval HBM GB/s, 8D, epsrel 1e-8",
            "config": {"workload": "8D cos(sum (i+1) x_i) and 8D discontinuous "
                                   "(a_i=5+10U, u_i=0.3+0.4U), epsrel %g, memory model: actual"
                                   % args.large_epsrel, "seed": args.seed},
            "runs": out}


def config4(args, pg, W):
    """512 independent integrals (family, d in {3..8}, epsrel log-uniform in
    [1e-6, 1e-3]) from one pagani_integrate_batch call."""
    import collections
    items = W.batch(args.batch_count, args.seed)
    b = pg.Batch(0, args.workers)
    b.set_timing(True)
    for _ in range(args.warmup):
        b.integrate(items[:16])
    _sync()
    b.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, codes = b.integrate(items)
    _sync()
    dt = time.perf_counter() - t0
    st = b.stats()
    b.close()
    conv = sum(1 for r in res if r.status == 0)
    within = 0
    fam_rows = collections.defaultdict(lambda: {"count": 0, "converged": 0,
                                                "rel_err_within_epsrel": 0})
    for (fam, a, u, eps), r in zip(items, res):
        exact = W.exact(fam, a, u)
        ok = abs(r.estimate - exact) <= eps * abs(exact)
        within += ok
        row = fam_rows[fam]
        row["count"] += 1
        row["converged"] += r.status == 0
        row["rel_err_within_epsrel"] += ok
    n = len(items) * args.steps
    return {
        "metric": "integrals/s, 512-integral Genz batch (FP64)",
        "value": n / dt, "unit": "integrals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.batch_count} Genz integrals, family ~ U(6), d ~ U{{3..8}}, "
                               "parameters per family, epsrel = 10^(-3-3U), epsabs 1e-20",
                   "seed": args.seed, "workers": args.workers},
        "regions_per_s": st["regions_evaluated"] / dt,
        "converged": conv, "rel_err_within_epsrel": within,
        "codes_negative": sum(1 for c in codes if c < 0),
        "phase_ms_per_step": {k: st[k] / args.steps for k in
                              ("eval_ms", "refine_ms", "compact_ms", "classify_ms")},
        "by_family": fam_rows,
    }


def mcubes_leg(args, pg, W, runs=3):
    """m-Cubes on the same seeded 8D Gaussian: 1e9 calls per iteration,
    10 iterations (5 adjusting the grid, 2 skipped), default streams; epsrel
    1e-12 so every run does the same work.  Timed over `runs` runs after
    one warm-up run."""
    fam, a, u = W.gaussian_8d(args.seed)
    d = len(a)
    exact = W.gaussian_exact(a, u)
    kw = dict(epsrel=1e-12, epsabs=1e-30, ncall=args.mc_ncall, total_iters=10, adjust_iters=5,
              skip_iters=2, generator="default", seed=args.seed)
    with pg.MCubes(d) as mc:
        mc.integrate_genz(fam, a, u, **kw)
        mc.set_timing(True)
        mc.reset_stats()
        _sync()
        t0 = time.perf_counter()
        for _ in range(runs):
            r = mc.integrate_genz(fam, a, u, **kw)
        dt = time.perf_counter() - t0
        st = mc.stats()
    fl = W.mc_flops_per_sample(fam, d)
    kern_ms = st["sample_ms"] / max(st["launches"], 1)
    per_launch = st["samples"] / max(st["launches"], 1)
    out = {
        "workload": "m-Cubes VEGAS, same 8D Gaussian, %.0e calls/iteration, 10 iterations "
                    "(5 grid-adjusting, 2 skipped), default streams" % args.mc_ncall,
        "samples_per_s": st["samples"] / dt,
        "ms_per_run": dt / runs * 1e3,
        "ms_per_iteration": dt / runs / 10 * 1e3,
        "sample_kernel_avg_ms": kern_ms,
        "samples_per_launch": per_launch,
        "flops_per_sample": fl,
        "fp64_tflops_algorithmic": fl * per_launch / (kern_ms * 1e-3) / 1e12 if kern_ms else None,
        "fp64_frac_algorithmic": (fl * per_launch / (kern_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS
                                  if kern_ms else None),
        "estimate": r["estimate"], "errorest": r["errorest"], "chi_sq": r["chi_sq"],
        "analytic": exact, "deviation_in_errorest": (r["estimate"] - exact) / r["errorest"],
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_api as O
        prm = O.params(a, u)
        n = 2e6
        t0 = time.perf_counter()
        O.mc_integrate(fam, d, prm, 1e-12, 1e-30, n, total_iters=2, adjust_iters=1,
                       skip_iters=0, generator="default", seed=args.seed)
        cdt = time.perf_counter() - t0
        s = O.mc_setup(d, n)
        out["cpu_baseline"] = {"value": 2 * s.calls / cdt, "unit": "samples/s", "cores": 1,
                               "kind": "port",
                               "sample": "oracle m-Cubes (one core), 2 iterations of %d calls "
                                         "in %.1f s" % (s.calls, cdt)}
    return out


def eval_only(args, pg, W):
    """The evaluation kernel in isolation, for hardware-counter runs: one
    device-to-device copy of known size (the reference point that scales
    FETCH_SIZE / WRITE_SIZE to bytes), then `--eval-launches` evaluations of
    the 8D Gaussian over a uniform k^8 split of the unit cube."""
    import torch
    fam, a, u = W.gaussian_8d(args.seed)
    d = len(a)
    k = args.eval_k
    n = k ** d
    dev = torch.device("cuda:0")
    L = torch.empty(n * d, dtype=torch.float64, device=dev)
    H = torch.empty(n * d, dtype=torch.float64, device=dev)
    est = torch.empty(n, dtype=torch.float64, device=dev)
    err = torch.empty(n, dtype=torch.float64, device=dev)
    dim = torch.empty(n, dtype=torch.uint8, device=dev)
    ws = pg.Workspace(d, 0)
    ws.uniform_split(k, L, H)
    nbytes = args.copy_mib << 20
    src = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.eval_launches):
        ws.eval_genz(fam, a, u, n, L, H, est, err, dim)
    dt = time.perf_counter() - t0
    ws.close()
    return {
        "mode": "eval-only", "kernel": "pagani::eval_kernel<8, Gaussian<double,8>, double>",
        "regions_per_launch": n, "launches": args.eval_launches,
        "algorithmic_bytes_per_launch": n * (16 * d + 17),
        "flops_per_launch": n * W.eval_flops_per_region(fam, d),
        "copy_bytes": nbytes,
        "avg_launch_ms_host": dt / args.eval_launches * 1e3,
        "tflops_host_timed": n * W.eval_flops_per_region(fam, d) * args.eval_launches / dt / 1e12,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--eval-only", action="store_true",
                    help="time the evaluation kernel alone (for counter runs)")
    ap.add_argument("--eval-k", type=int, default=6)
    ap.add_argument("--eval-launches", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--workers", type=int, default=1)
    ap.add_argument("--epsrel", type=float, default=1e-6)
    ap.add_argument("--mixed-epsrel", type=float, default=1e-5)
    ap.add_argument("--large-epsrel", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-mcubes", action="store_true")
    ap.add_argument("--mcubes-only", action="store_true",
                    help="the m-Cubes leg alone (for profiles)")
    ap.add_argument("--mc-ncall", type=float, default=1e9)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(each integral runs on one device); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import gpuintegration_amd as pg
    from gpuintegration_amd import workloads as W
    fn = {1: config1, 2: config2, 3: config3, 4: config4, 5: config5}[args.config]
    if args.eval_only:
        fn = eval_only
    if args.mcubes_only:
        fn = mcubes_leg
    print(json.dumps(fn(args, pg, W)), flush=True)


if __name__ == "__main__":
    main()
