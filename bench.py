"""Benchmark: simulated days per day (SDPD) of the dynamical-core step on one MI355X.

Workload: glevel 7 (163,842 cells, 40 regions of 66 x 66 points), 40 layers
up to 30 km (100 m at the surface), dt = 300 s, RK3 with 2/3/6 HEVI small
steps, DIRECT hyperdiffusion and 3-D divergence damping; dry balanced zonal
flow with seeded potential-temperature bumps (seed 1), flat topography; plain
FP64 HIP form.  A step is one large step of the whole globe with the state
resident in HBM.  SDPD = dt / wall time per step.

Prints one JSON line (rank 0).  The dominant kernel's achieved HBM bandwidth
is its algorithmic bytes per launch divided by its average launch time,
measured with events on the launch stream during the timed steps.  The CPU
baseline times the numpy FP64 oracle on one gl05 z40 step on the host and
scales it to gl07 by the cell ratio (the step's work is proportional to the
number of cells).
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "SDPD at gl07 z40, FP32+FP64-solve vs plain FP64; adjoint/forward step cost"
HBM_PEAK_GBS = 8000.0


def pkg(name):
    return importlib.import_module(f"pynicam-dc_amd.{name}")


def kernel_bytes(t) -> dict:
    """Algorithmic HBM bytes per launch of each kernel (FP64).

    Each distinct (i, j, k, l) field read or written once counts F bytes; each
    per-column coefficient component counts C = F / kall; k-only profiles are
    ignored.  Field counts follow the arrays each kernel touches.
    """
    ncell = t.lall * t.gall_1d ** 2
    F = 8.0 * (ncell * t.kall + 12 * t.kall)
    Cc = 8.0 * ncell
    return {
        "prepost": 14 * F,                      # prog 6 + qv in; diag 6 + eth out
        "adv_mom": 13 * F + 24 * Cc,            # prog 5 + diag 4 in; 4 out; div + xyz coefs
        "hdiff_pass1": 12 * F + 7 * Cc,         # 6 in, 6 out; Laplacian coefs
        "hdiff_pass2": 13 * F + 43 * Cc,        # 6 + rhog in, 6 out; lap + intp + diff coefs
        "divdamp_3d": 7 * F + 36 * Cc,          # 4 in, 3 out; intp + diff coefs
        "divdamp_post": 6 * F + 39 * Cc,        # 3 in, 3 out; intp + diff + xyz
        "vi_path0": 34 * F + 45 * Cc,           # 23 in, 11 out; div + grad + xyz
        "ethtot_mean": 12 * F,                  # prog 5 + eth in; ethtot + 5 means out
        "split_init": 15 * F,                   # prog0 6 + prog 6 in (2 of 3 stages), 6 out
        "vi_path1": 20 * F + 24 * Cc,           # 14 in, 6 out; grad + xyz
        "vi_main": 30 * F + 21 * Cc,            # 25 in, 5 out; div coefs
        "prog_update": 18 * F + 3 * Cc,         # 12 in, 6 out
    }


def cpu_baseline(seconds_cap: float = 60.0) -> dict:
    """One gl05 z40 large step of the numpy oracle on the host, scaled to gl07 SDPD."""
    from oracle.dynamics import OracleContext, OracleDynamics
    case = pkg("setup").build_case(glevel=5, vlayer=40, dtl=1200.0, seed=1)
    dyn = OracleDynamics(OracleContext(case.geo, case.const, None), case.state)
    t0 = time.perf_counter()
    dyn.step()
    t_gl05 = time.perf_counter() - t0
    cells5 = 10 * 4 ** 5 + 2
    cells7 = 10 * 4 ** 7 + 2
    t_gl07 = t_gl05 * cells7 / cells5
    threads = os.environ.get("OMP_NUM_THREADS", "unset")
    return {
        "value": 300.0 / t_gl07,
        "unit": "SDPD",
        "cores": 1,
        "kind": "port",
        "sample": (f"numpy FP64 oracle, one gl05 z40 large step (dt 1200 s) = {t_gl05:.2f} s on 1 host core "
                   f"(OMP_NUM_THREADS={threads}, os.cpu_count()={os.cpu_count()}), scaled by the cell ratio "
                   f"{cells7}/{cells5} to a gl07 step of {t_gl07:.1f} s at dt 300 s"),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--glevel", type=int, default=7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the run, copy 512 MiB device-to-device once (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU path is a design, not a run", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    gl = args.glevel
    dt = {5: 1200.0, 6: 600.0, 7: 300.0, 8: 150.0}[gl]
    print(f"bench: setting up gl{gl:02d} z40 dt {dt:g} s", file=sys.stderr, flush=True)
    case = pkg("setup").build_case(glevel=gl, vlayer=40, dtl=dt, seed=1)
    hipdyn = pkg("hipdyn")
    ctx = hipdyn.HipContext(case.geo, case.const, device="cuda:0")
    dyn = hipdyn.HipDynamics(ctx, case.state)
    for w in range(args.warmup):
        dyn.dynamics_step()
    ctx.sync()
    print(f"bench: {args.warmup} warmup steps done", file=sys.stderr, flush=True)
    ctx.prof_report()                      # clear
    ctx.prof_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dyn.dynamics_step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.prof_enable(False)
    prof = ctx.prof_report()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    sdpd = dt / (ms_per_step / 1000.0)
    st = dyn.state()
    import numpy as np
    g = case.topo.gall_1d
    if not np.all(np.isfinite(st["PROG"][1:g - 1, 1:g - 1, 1:-1])):
        print("bench: state is not finite", file=sys.stderr)
        sys.exit(1)
    kb = kernel_bytes(case.topo)
    dom = max((k for k in prof if k in kb), key=lambda k: prof[k][0])
    tot, n = prof[dom]
    avg_s = tot / n / 1000.0
    achieved = kb[dom] / avg_s / 1e9
    kernels = {k: {"ms_per_step": v[0] / args.steps, "launches_per_step": v[1] / args.steps,
                   "avg_us": 1000.0 * v[0] / v[1],
                   "GBps": (kb[k] / (v[0] / v[1] / 1000.0) / 1e9) if k in kb else None}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}
    print("bench: kernels " + json.dumps(kernels), file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": sdpd,
        "unit": "SDPD",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"gl{gl:02d} z40 dynamics step, dt {dt:g} s, plain FP64 HIP form, "
                               "dry balanced zonal flow + seeded theta bumps (seed 1), flat topography",
                   "glevel": gl, "vlayer": 40, "dt": dt, "cells": 10 * 4 ** gl + 2, "form": "fp64",
                   "parallelism": "single GPU, 40 regions"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": kb[dom], "avg_launch_us": avg_s * 1e6},
    }
    if args.copy_probe:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print("bench: copy probe of 536870912 bytes done", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print("bench: timing the CPU oracle", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
