"""Benchmark: PAGANI on one MI355X, 8D Genz Gaussian to epsrel 1e-6.

One step = one full integration to convergence of the seeded 8D Gaussian
f(x) = exp(-sum a_i^2 (x_i - u_i)^2), a_i = 10 (1 + 0.05 (2U - 1)),
u_i = 0.2 + 0.6 U (std::mt19937_64, U = (x >> 11) 2^-53), epsrel 1e-6,
initial uniform split 4 per axis (65,536 regions), FP64.  Inputs are
synthetic and resident on the device; the whole loop (evaluation,
refinement, classification, filtering, splitting) is inside the timed
region.

value = regions evaluated per second (all iterations, all steps).
Also reported: integrals/s, the FP32-integrand (mixed precision) run of the
same workload, the evaluation kernel's roofline (algorithmic FP64 FLOPs /
its HIP-event time vs 78.6 TFLOP/s) and the CPU oracle on the host cores.
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
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def run_gpu(pg, W, fam, a, u, epsrel, steps, warmup, precision, init_k=4):
    d = len(a)
    ws = pg.Workspace(d, 0)
    ws.set_timing(True)
    for _ in range(warmup):
        ws.integrate_genz(fam, a, u, epsrel=epsrel, epsabs=1e-20, precision=precision,
                          init_k=init_k)
    _sync()
    ws.reset_stats()
    t0 = time.perf_counter()
    res = None
    for _ in range(steps):
        res = ws.integrate_genz(fam, a, u, epsrel=epsrel, epsabs=1e-20, precision=precision,
                                init_k=init_k)
    _sync()
    dt = time.perf_counter() - t0
    st = ws.stats()
    tr = ws.trace()
    ws.close()
    return res, dt, st, tr


def cpu_baseline(fam, a, u, budget_s):
    """The oracle's region evaluation on the host cores: repeated passes over
    the workload's first iteration (the 4^8 uniform split) for ~budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--epsrel", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(each integral runs on one device); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import gpuintegration_amd as pg
    from gpuintegration_amd import workloads as W

    fam, a, u = W.gaussian_8d(args.seed)
    d = len(a)
    exact = W.gaussian_exact(a, u)
    res, dt, st, tr = run_gpu(pg, W, fam, a, u, args.epsrel, args.steps, args.warmup, pg.FP64)
    regions = st["regions_evaluated"]
    value = regions / dt
    flops_region = W.eval_flops_per_region(fam, d)
    eval_s = st["eval_ms"] / 1e3
    achieved = flops_region * regions / eval_s / 1e12 if eval_s > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": value,
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
        "phase_ms_per_step": {k: st[k] / args.steps for k in
                              ("eval_ms", "refine_ms", "compact_ms", "classify_ms")},
        "roofline": {
            "bound": "fp64-valu",
            "kernel": "pagani::eval_kernel<8, Gaussian<double,8>, double>",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "flops_per_region": flops_region,
            "eval_launches": st["eval_launches"],
            "avg_launch_ms": st["eval_ms"] / max(st["eval_launches"], 1),
            "traffic": None,
        },
    }
    if not args.no_mixed:
        r32, dt32, st32, _ = run_gpu(pg, W, fam, a, u, args.epsrel, args.steps, args.warmup,
                                     pg.FP32)
        e32 = st32["eval_ms"] / 1e3
        out["mixed"] = {
            "precision": "fp32 integrand + fp64 sums/errors",
            "value": st32["regions_evaluated"] / dt32,
            "unit": "regions/s",
            "integrals_per_s": args.steps / dt32,
            "estimate": r32.estimate,
            "errorest": r32.errorest,
            "abs_err_vs_analytic": abs(r32.estimate - exact),
            "status": r32.status,
            "nregions": r32.nregions,
            "eval_regions_per_s": st32["regions_evaluated"] / e32 if e32 > 0 else None,
        }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(fam, a, u, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
