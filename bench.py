"""Benchmark: PAGANI on one MI355X.

Default (--config 2, the headline line): the seeded 8D Genz Gaussian
f(x) = exp(-sum a_i^2 (x_i - u_i)^2), a_i = 10 (1 + 0.05 (2U - 1)),
u_i = 0.2 + 0.6 U (std::mt19937_64, U = (x >> 11) 2^-53), integrated to
epsrel 1e-6 from an initial uniform split of 4 per axis (65,536 regions),
FP64.  One step = one full integration to convergence; inputs are synthetic
and resident on the device; the whole loop (evaluation, refinement,
classification, filtering, splitting) is inside the timed region.
value = regions evaluated per second.

The headline line also carries, each timed in the same run:
  generic  the same integral through the C++ drop-in header with the
           integrand written as a plain functor (only operator(), the
           reference's formula): the kernel's point-by-point path
  mixed    FP32 integrand + FP64 sums beside FP64 at epsrel 1e-5
  mixed_6d the 6D corner-peak and product-peak sweep (epsrel 1e-3..1e-7),
           FP32 integrand beside FP64 (also --config 3)
  mcubes   the m-Cubes VEGAS integrator on the same integrand
  batch    the 512-integral Genz batch (family and d in {3..8} drawn,
           epsrel log-uniform in [1e-6, 1e-3]) from one call, the short
           integrals in the persistent kernel (--segments regions each)
  persistent_batch  the batch's persistent kernel (whole integrals, one
           workgroup each) off and on, on 512 short integrals and on the
           512-integral batch
  large    8D cos(sum (i+1) x_i) and 8D discontinuous at epsrel 1e-8 with
           the library's own memory model (pools of 2e8-5e8 regions)
  plain_cpu_vs_gpu  5D oscillatory at epsrel 1e-3: the sequential
           Cuhre-equivalent on one host core beside the GPU
  cpu_baseline  the oracle's whole PAGANI loop (OpenMP, all host threads)
           on the headline integral to a stated epsrel cap
(--quick leaves out batch, large, plain_cpu_vs_gpu, mcubes and the CPU
baselines.)  Other configurations print their own line: --config 1, 3, 4, 5.

The roofline's per-region figures (executed FP64 FLOPs, VALU instructions,
HBM bytes) are counter measurements read from
profiles/eval_gaussian8d_counters.json; the launch durations are measured
here with HIP events on the workspace's stream.  A fraction above 1 means
those figures no longer describe the kernel, and the run fails.
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
VALU_ISSUE_PEAK = 256 * 2.4e9  # wave-instructions per second (one per CU per clock)
COUNTERS_PATH = os.path.join(ROOT, "profiles", "eval_gaussian8d_counters.json")


def load_counters():
    with open(COUNTERS_PATH) as f:
        return json.load(f)


def check_fraction(name, frac):
    """A roofline fraction above one cannot be measured: the per-region
    counter figures are stale for the kernel that ran."""
    if frac is not None and frac > 1.0:
        raise SystemExit(f"bench.py: {name} = {frac:.3f} > 1: the counter figures in "
                         f"{COUNTERS_PATH} do not describe this kernel; measure them again")
    return frac


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


_T0 = time.perf_counter()


def _note(msg):
    """A progress line on stderr (the JSON line stays alone on stdout)."""
    print(f"bench: {msg} ({time.perf_counter() - _T0:.1f} s)", file=sys.stderr, flush=True)


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def noneval_bytes(trace, d):
    """Algorithmic HBM bytes of this library's refinement and filter + split
    over a run, per iteration of n regions keeping m: the refinement reads
    each region's estimate and error (16), its share of the parents'
    estimates (4), writes the refined error (8) and the flag (1); the
    compaction reads every flag (1) and, per kept region, its parent's row
    (8d lower bounds + 2d width levels), the parent's halved axis, its own
    axis (1 + 1) and estimate (8), and writes it as a parent of the next
    list (10d row + 1 axis + 8 estimate): 30 n + (20d + 19) m."""
    return sum(int(n) * 30 + int(m) * (20 * d + 19) for n, m in zip(trace["n"], trace["m"]))


def reference_noneval_bytes(trace, d):
    """The bytes the reference's phases move for the same trajectory
    (n (16d + 33) + m (32d + 16): both children written out in full)."""
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
    ref_bytes = reference_noneval_bytes(tr, d) * steps
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
        "noneval_frac_hbm": check_fraction("non-eval HBM fraction",
                                           ne_bytes / ne_s / 1e9 / HBM_PEAK_GBS)
        if ne_s > 0 else None,
        "noneval_reference_bytes_per_s": ref_bytes / ne_s if ne_s > 0 else None,
        "phase_ms_per_integral": {k: st[k] / steps for k in
                                  ("eval_ms", "refine_ms", "compact_ms", "classify_ms")},
        "classify_passes_per_integral": st["classify_passes"] / steps,
        "device_loop": {"ms": st["device_ms"] / steps,
                        "iterations": st["device_iterations"] / steps,
                        "rounds": st["device_rounds"] / steps},
        "alloc_ms_per_integral": st["alloc_ms"] / steps,
        "allocations_per_integral": st["allocations"] / steps,
    }


def cpu_baseline(fam, a, u, epsrel):
    """The oracle's whole PAGANI loop (evaluation OpenMP-parallel over the
    host threads, the list phases and the classifier sequential) on the
    headline integral from the same 4^8 start, to `epsrel` (the headline's
    1e-6 would take the host about 15 minutes).  Regions evaluated per
    second, like the headline value."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_api as O
    d = len(a)
    prm = O.params(a, u)
    t0 = time.perf_counter()
    r, tr = O.integrate(d, fam, prm, epsrel, 1e-20, init_k=4, trace_cap=800)
    dt = time.perf_counter() - t0
    regions = int(sum(int(v) for v in tr["n"]))
    return {
        "value": regions / dt,
        "unit": "regions/s",
        "cores": int(O.LIB.orc_num_threads()),
        "kind": "port",
        "cpu": cpu_model(),
        "sample": f"oracle PAGANI loop ({O.LIB.orc_num_threads()} OpenMP threads) on the same "
                  f"8D Gaussian from the 4^8 split to epsrel {epsrel:g}: {regions} regions in "
                  f"{len(tr['n'])} iterations, {dt:.1f} s (status {r.status})",
        "estimate": r.estimate, "errorest": r.errorest,
    }


def generic_leg(args, W, counters, steps=2, warmup=1):
    """The headline integral through the C++ header with a plain functor
    (libpagani_header_bench.so, compiled as a user translation unit)."""
    import ctypes
    import gpuintegration_amd as pg
    pg.load()
    lib = ctypes.CDLL(os.path.join(ROOT, "gpuintegration_amd", "libpagani_header_bench.so"))
    fam, a, u = W.gaussian_8d(args.seed)
    d = len(a)
    A = (ctypes.c_double * 8)(*a)
    U = (ctypes.c_double * 8)(*u)
    res = pg._Result()
    st = pg.Stats()
    secs = ctypes.c_double(0)
    rc = lib.pagani_header_bench_gaussian8(A, U, ctypes.c_double(args.epsrel),
                                           ctypes.c_double(1e-20), steps, warmup,
                                           ctypes.byref(res), ctypes.byref(st), ctypes.byref(secs))
    if rc != 0:
        raise SystemExit(f"bench.py: header bench failed with {rc}")
    st = st.as_dict()
    dt = secs.value
    host_regions = st["regions_evaluated"] - st["device_regions"]
    host_launches = max(st["eval_launches"] - st["device_iterations"], 1)
    eval_s = st["eval_ms"] / 1e3
    alg = W.eval_flops_per_region(fam, d)
    out = {
        "workload": "same integral; Workspace<double, 8>::integrate through the C++ header, "
                    "integrand a plain functor exp(-sum a_i^2 (x_i - u_i)^2) (operator() only)",
        "regions_per_s": st["regions_evaluated"] / dt,
        "integrals_per_s": steps / dt,
        "ms_per_integral": dt / steps * 1e3,
        "estimate": res.estimate, "errorest": res.errorest, "status": res.status,
        "nregions": res.nregions,
        "eval_ms_per_integral": st["eval_ms"] / steps,
        "avg_eval_launch_ms": st["eval_ms"] / host_launches,
        "eval_regions_per_s": host_regions / eval_s if eval_s else None,
        "algorithmic_flops_per_region": alg,
        "algorithmic_tflops": alg * host_regions / eval_s / 1e12 if eval_s else None,
    }
    out["algorithmic_frac_fp64_peak"] = check_fraction(
        "generic algorithmic FP64 fraction",
        out["algorithmic_tflops"] / FP64_PEAK_TFLOPS if eval_s else None)
    pc = counters.get("plain", {})
    if pc.get("exec_fp64_flop_per_region") and eval_s:
        ex = pc["exec_fp64_flop_per_region"] * host_regions / eval_s / 1e12
        out["executed_tflops"] = ex
        out["executed_frac_fp64_peak"] = check_fraction("generic executed FP64 fraction",
                                                        ex / FP64_PEAK_TFLOPS)
        out["valu_issue_frac"] = check_fraction(
            "generic VALU issue fraction",
            pc["valu_insts_per_region"] / 64 * host_regions / eval_s / VALU_ISSUE_PEAK)
    return out


def calibration_copy(mib):
    """One device-to-device copy of `mib` MiB: the reference point that scales
    the FETCH_SIZE / WRITE_SIZE counters of a run to bytes."""
    import torch
    nbytes = mib << 20
    src = torch.zeros(nbytes, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()


def config2(args, pg, W):
    if args.calibrate_copy:
        calibration_copy(args.copy_mib)
    fam, a, u = W.gaussian_8d(args.seed)
    d = len(a)
    exact = W.gaussian_exact(a, u)
    counters = load_counters()
    fc = counters["factored"]
    res, dt, st, tr = run_gpu(pg, fam, a, u, args.epsrel, args.steps, args.warmup, pg.FP64,
                              init_k=4)
    rep = phase_report(W, fam, d, res, dt, st, tr, args.steps)
    regions = st["regions_evaluated"]
    # the roofline covers the host-driven evaluation launches (timed one by
    # one with HIP events on the workspace's stream); the device loop's small
    # iterations are reported beside it
    host_regions = regions - st["device_regions"]
    host_launches = st["eval_launches"] - st["device_iterations"]
    eval_s = st["eval_ms"] / 1e3
    achieved = fc["exec_fp64_flop_per_region"] * host_regions / eval_s / 1e12 if eval_s else 0.0
    valu_frac = (fc["valu_insts_per_region"] / 64 * host_regions / eval_s / VALU_ISSUE_PEAK
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
        "noneval_gbs_algorithmic": rep["noneval_gbs"],
        "roofline": {
            "bound": "fp64-valu",
            "kernel": fc["kernel"],
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": check_fraction("roofline.frac", achieved / FP64_PEAK_TFLOPS),
            "flops_per_region": fc["exec_fp64_flop_per_region"],
            "flops_per_region_kind": "executed FP64 FLOPs (counter-measured, "
                                     "profiles/eval_gaussian8d_counters.json)",
            "algorithmic_flops_per_region": W.eval_flops_per_region(fam, d),
            "algorithmic_tflops": rep["eval_tflops"],
            "valu_issue_frac": check_fraction("roofline.valu_issue_frac", valu_frac),
            "eval_launches": host_launches,
            "avg_launch_ms": st["eval_ms"] / max(host_launches, 1),
            "avg_regions_per_launch": host_regions / max(host_launches, 1),
            # the host-driven launches evaluate split-form lists with widths
            # as levels: a parent's row (8d lower bounds + 2d levels) serves
            # its two children, plus the halved axis (1 per parent), and each
            # child writes estimate, error and axis (17)
            "traffic": fc["split_hbm_bytes_per_region"] * host_regions / max(host_launches, 1),
            "traffic_unit": "bytes per launch (counter-measured per-region HBM bytes of the "
                            "headline's split-form launches x average regions per launch)",
            "algorithmic_bytes_per_launch": (5 * d + 17.5) * host_regions / max(host_launches, 1),
            "traffic_vs_algorithmic": fc["split_hbm_bytes_per_region"] / (5 * d + 17.5),
        },
    }
    _note("headline done")
    if not args.no_generic:
        out["generic"] = generic_leg(args, W, counters)
        _note("generic leg done")
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
        _note("mixed leg done")
    if args.quick:
        return out
    out["mixed_6d"] = config3(args, pg, W)
    _note("6D mixed-precision sweep done")
    if not args.no_mcubes:
        out["mcubes"] = mcubes_leg(args, pg, W)
        _note("mcubes leg done")
    out["batch"] = config4(args, pg, W, leg=True)
    _note("batch leg done")
    out["persistent_batch"] = persistent_leg(args, pg, W)
    _note("persistent batch leg done")
    # the same batch with the fixed launch-shape rule instead of the
    # autotuner's table
    rule = config4(args, pg, W, leg=True, tuned=False)
    out["launch_shape"] = {"batch_integrals_per_s_tuned": out["batch"]["value"],
                           "batch_integrals_per_s_rule": rule["value"],
                           "batch_eval_ms_tuned": out["batch"]["phase_ms_per_step"]["eval_ms"],
                           "batch_eval_ms_rule": rule["phase_ms_per_step"]["eval_ms"]}
    _note("launch-shape leg done")
    out["large"] = config5(args, pg, W)
    _note("large leg done")
    out["plain_cpu_vs_gpu"] = config1(args, pg, W)
    _note("plain CPU/GPU leg done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(fam, a, u, args.cpu_epsrel)
        _note("CPU baseline done")
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
    threads = O.LIB.orc_num_threads()
    O.LIB.orc_set_num_threads(1)
    try:
        t0 = time.perf_counter()
        reps = 0
        while True:
            rc = O.cuhre_seq(d, fam, prm, 1e-3, 1e-20, 1_000_000)
            reps += 1
            if time.perf_counter() - t0 > 3.0:
                break
        cdt = (time.perf_counter() - t0) / reps
    finally:
        O.LIB.orc_set_num_threads(threads)
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


def config4(args, pg, W, leg=False, tuned=True, segments=None):
    """512 independent integrals (family, d in {3..8}, epsrel log-uniform in
    [1e-6, 1e-3]) from one pagani_integrate_batch call.  As a leg of the
    headline line: one timed call after a 16-item warm-up.  segments: the
    persistent kernel's segment size (pagani_batch_set_segments; None:
    --segments, 0: off)."""
    import collections
    items = W.batch(args.batch_count, args.seed)
    steps = 1 if leg else args.steps
    warmup = 1 if leg else args.warmup
    b = pg.Batch(0, args.workers)
    b.set_timing(True)
    b.set_tuned_slices(tuned)
    seg = args.segments if segments is None else segments
    b.set_segments(seg)
    for _ in range(warmup):
        b.integrate(items[:16])
    _sync()
    b.reset_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
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
    n = len(items) * steps
    return {
        "metric": "integrals/s, 512-integral Genz batch (FP64)",
        "value": n / dt, "unit": "integrals/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": dt / steps * 1e3,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.batch_count} Genz integrals, family ~ U(6), d ~ U{{3..8}}, "
                               "parameters per family, epsrel = 10^(-3-3U), epsabs 1e-20",
                   "seed": args.seed, "workers": args.workers, "segment_regions": seg},
        "regions_per_s": st["regions_evaluated"] / dt,
        "segment_integrals": st["segment_integrals"] // steps,
        "converged": conv, "rel_err_within_epsrel": within,
        "codes_negative": sum(1 for c in codes if c < 0),
        "phase_ms_per_step": {k: st[k] / steps for k in
                              ("eval_ms", "refine_ms", "compact_ms", "classify_ms",
                               "device_ms", "alloc_ms")},
        "by_family": fam_rows,
    }


def persistent_leg(args, pg, W, reps=3):
    """The batch's persistent kernel (pagani_batch_set_segments) on many
    short integrals (workloads.small_batch: 512 items, d <= 6, epsrel >=
    1e-4) and on the 512-integral batch, off and on: median of `reps` timed
    calls after one untimed call, one worker."""
    import statistics
    seg = args.persistent_segments

    def timed(items, regions, n):
        b = pg.Batch(0, 1)
        b.set_segments(regions)
        b.integrate(items)
        _sync()
        ts = []
        for _ in range(n):
            b.reset_stats()
            t0 = time.perf_counter()
            res, codes = b.integrate(items)
            _sync()
            ts.append(time.perf_counter() - t0)
        st = b.stats()
        b.close()
        return statistics.median(ts), st, res

    out = {"segment_regions": seg, "workers": 1}
    for name, items, n in (("small_batch", W.small_batch(512, 7), reps),
                           ("batch", W.batch(args.batch_count, args.seed), 1)):
        t_off, _, r_off = timed(items, 0, n)
        t_on, st, r_on = timed(items, seg, n)
        same = sum(1 for x, y in zip(r_on, r_off)
                   if (x.nregions, x.iters, x.status) == (y.nregions, y.iters, y.status))
        out[name] = {
            "items": len(items), "timed_calls": n,
            "integrals_per_s_off": len(items) / t_off, "integrals_per_s_on": len(items) / t_on,
            "speedup": t_off / t_on,
            "segment_integrals": st["segment_integrals"], "segment_ms": st["segment_ms"],
            "same_iterations": same,
            "max_rel_diff_estimate": max(abs(x.estimate - y.estimate) / max(abs(y.estimate), 1e-300)
                                         for x, y in zip(r_on, r_off)),
        }
    return out


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
    if args.plain:
        # the plain functor's kernel, compiled in libpagani_header_bench.so
        import ctypes
        lib = ctypes.CDLL(os.path.join(ROOT, "gpuintegration_amd", "libpagani_header_bench.so"))
        lib.pagani_header_bench_eval_gaussian8.argtypes = [ctypes.c_void_p] * 3 + [
            ctypes.c_size_t] + [ctypes.c_void_p] * 5 + [ctypes.c_int]
        A = (ctypes.c_double * d)(*a)
        U = (ctypes.c_double * d)(*u)
        ptrs = [t.data_ptr() for t in (L, H, est, err, dim)]
        kernel = "pagani::eval_kernel<8, PlainGaussian8, double>"
        t0 = time.perf_counter()
        rc = lib.pagani_header_bench_eval_gaussian8(ws._ws, A, U, n, *ptrs, args.eval_launches)
        dt = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"pagani_header_bench_eval_gaussian8 returned {rc}")
    else:
        kernel = "pagani::eval_kernel<8, Gaussian<double,8>, double>"
        t0 = time.perf_counter()
        for _ in range(args.eval_launches):
            ws.eval_genz(fam, a, u, n, L, H, est, err, dim)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    ws.close()
    return {
        "mode": "eval-only", "kernel": kernel,
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
    ap.add_argument("--segments", type=int, default=65536,
                    help="the batch's persistent kernel: regions per segment (0: off)")
    ap.add_argument("--persistent-segments", type=int, default=65536,
                    help="segment size of the persistent-kernel leg")
    ap.add_argument("--eval-only", action="store_true",
                    help="time the evaluation kernel alone (for counter runs)")
    ap.add_argument("--plain", action="store_true",
                    help="with --eval-only: the plain functor's kernel (operator() only, "
                         "compiled through the C++ header) instead of the Genz descriptor's")
    ap.add_argument("--eval-k", type=int, default=6)
    ap.add_argument("--eval-launches", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="start the headline with a device copy of --copy-mib MiB (counter runs)")
    ap.add_argument("--workers", type=int, default=1)
    ap.add_argument("--epsrel", type=float, default=1e-6)
    ap.add_argument("--mixed-epsrel", type=float, default=1e-5)
    ap.add_argument("--large-epsrel", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-epsrel", type=float, default=1e-3,
                    help="epsrel cap of the oracle's loop in the CPU baseline")
    ap.add_argument("--quick", action="store_true",
                    help="headline, generic and mixed legs only")
    ap.add_argument("--no-generic", action="store_true")
    ap.add_argument("--generic-only", action="store_true",
                    help="the generic leg alone (for counter runs)")
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
    if args.generic_only:
        def fn(args, pg, W):
            return generic_leg(args, W, load_counters(), steps=1, warmup=0)
    print(json.dumps(fn(args, pg, W)), flush=True)


if __name__ == "__main__":
    main()
