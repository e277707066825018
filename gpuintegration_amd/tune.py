"""Autotuner for the evaluation kernel's launch shape.

The reference launches one 64-thread block per region (PaganiUtils.cuh:
313-331).  Here a wave holds 64 regions, one per lane, and for short lists
`slices` waves share a batch of 64 regions and split its points, so that
the device has enough waves in flight (include/pagani/eval_kernel.hpp).  How
many waves pay off depends on the integrand's cost per point, the dimension
(points per region), the precision and the list's size.  This script
measures it: for every Genz family, dimension and precision, and for lists
in each power-of-two size class, it times the evaluation kernel at 1, 2, 4
and 8 slices (HIP events on the workspace's stream, the median of a few
launches) and records the fastest.  The table goes to
gpuintegration_amd/tuning/eval_slices.json; the build compiles it into the
library, whose automatic launch shape then takes the table's entry for the
running integrand and list size (pagani_ws_set_tuned_slices turns it off).

    python -m gpuintegration_amd.tune                # every family, d 2..10, FP64 and FP32
    python -m gpuintegration_amd.tune --dims 8 --families gaussian

Runs on the GPU; takes about a minute for the whole table.
"""
import argparse
import json
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tuning", "eval_slices.json")
FAMILIES = ["oscillatory", "product_peak", "corner_peak", "gaussian", "c0", "discontinuous"]
SLICES = (1, 2, 4, 8)


def family_params(family, d, seed):
    """Parameters of the batch workload's distributions (workloads.batch)."""
    from gpuintegration_amd.workloads import MT19937_64
    g = MT19937_64(seed)
    if family == "oscillatory":
        a = [1.0 + 0.5 * g.uniform() for _ in range(d)]
        u = [g.uniform() for _ in range(d)]
    elif family == "gaussian":
        a = [10.0 * (1.0 + 0.05 * (2.0 * g.uniform() - 1.0)) for _ in range(d)]
        u = [0.2 + 0.6 * g.uniform() for _ in range(d)]
    elif family == "discontinuous":
        a = [5.0 + 10.0 * g.uniform() for _ in range(d)]
        u = [0.3 + 0.4 * g.uniform() for _ in range(d)]
    else:
        a = [5.0 + 10.0 * g.uniform() for _ in range(d)]
        u = [g.uniform() for _ in range(d)]
    return a, u


def regions(d, n, seed, device):
    """n random sub-boxes of the unit cube (SoA), widths 10^U(-3, -0.3)."""
    import torch
    gen = torch.Generator().manual_seed(seed)
    H = 10.0 ** (-3.0 + 2.7 * torch.rand((d, n), generator=gen, dtype=torch.float64))
    L = torch.rand((d, n), generator=gen, dtype=torch.float64) * (1.0 - H)
    return L.reshape(-1).contiguous().to(device), H.reshape(-1).contiguous().to(device)


def time_shape(ws, family, a, u, n, L, H, est, err, dim, precision, slices, launches):
    ws.set_slices(slices)
    ws.eval_genz(family, a, u, n, L, H, est, err, dim, precision=precision)  # warm-up
    times = []
    for _ in range(launches):
        ws.reset_stats()
        ws.eval_genz(family, a, u, n, L, H, est, err, dim, precision=precision)
        times.append(ws.stats()["eval_ms"])
    return statistics.median(times)


def tune(dims, families, precisions, buckets, launches, log=print):
    import torch
    import gpuintegration_amd as pg
    device = torch.device("cuda:0")
    tables = []
    for d in dims:
        with pg.Workspace(d, 0) as ws:
            ws.set_timing(True)
            for b in buckets:
                n = 3 << (b - 1)  # the middle of the size class [2^b, 2^(b+1))
                L, H = regions(d, n, seed=1000 * d + b, device=device)
                est = torch.empty(n, dtype=torch.float64, device=device)
                err = torch.empty_like(est)
                dim = torch.empty(n, dtype=torch.uint8, device=device)
                for fam in families:
                    a, u = family_params(fam, d, seed=77 + d)
                    for prec in precisions:
                        t = {s: time_shape(ws, fam, a, u, n, L, H, est, err, dim, prec, s,
                                           launches) for s in SLICES}
                        best = min(SLICES, key=lambda s: (t[s], s))
                        row = next((r for r in tables if r["family"] == fam and r["ndim"] == d
                                    and r["precision"] == prec), None)
                        if row is None:
                            row = {"family": fam, "ndim": d, "precision": prec, "slices": {},
                                   "times_ms": {}}
                            tables.append(row)
                        row["slices"][str(b)] = best
                        row["times_ms"][str(b)] = {str(s): round(t[s], 5) for s in SLICES}
                log(f"d={d} n={n}: " + ", ".join(
                    f"{r['family']}/{r['precision']}:{r['slices'][str(b)]}" for r in tables
                    if r["ndim"] == d))
            ws.set_slices(0)
    return tables


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--dims", default="2-10", help="e.g. 8 or 2-10")
    ap.add_argument("--families", default=",".join(FAMILIES))
    ap.add_argument("--precisions", default="64,32")
    ap.add_argument("--buckets", default="6-19",
                    help="size classes 2^b <= n < 2^(b+1); above 2^19 one wave per batch "
                         "already fills the device")
    ap.add_argument("--launches", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args(argv)

    def span(text):
        if "-" in text:
            lo, hi = text.split("-")
            return list(range(int(lo), int(hi) + 1))
        return [int(x) for x in text.split(",")]

    t0 = time.perf_counter()
    tables = tune(span(args.dims), args.families.split(","),
                  [int(p) for p in args.precisions.split(",")], span(args.buckets),
                  args.launches, log=lambda m: print(m, flush=True))
    import torch
    out = {
        "about": "fastest slices (waves per batch of 64 regions) of the evaluation kernel per "
                 "Genz family, dimension, precision and list size class (key b: 2^b <= n < "
                 "2^(b+1)), measured by gpuintegration_amd/tune.py",
        "device": torch.cuda.get_device_name(0),
        "launches_per_shape": args.launches,
        "seconds": round(time.perf_counter() - t0, 1),
        "tables": tables,
    }
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {args.out} ({len(tables)} tables, {out['seconds']} s)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
