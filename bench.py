"""Benchmark: simulated days per day (SDPD) of the dynamical-core step on one MI355X.

Workload: glevel 7 (163,842 cells, 40 regions of 66 x 66 points), 40 layers
up to 30 km (100 m at the surface), dt = 300 s, RK3 with 2/3/6 HEVI small
steps, DIRECT hyperdiffusion and 3-D divergence damping, tracer transport of
qv; dry balanced zonal flow with seeded potential-temperature bumps (seed 1),
flat topography.  A step is one large step of the whole globe with the state
resident in HBM.  SDPD = dt / wall time per step.  Both forms are timed on the
same state: the mixed form (FP32 fields and stencils, FP64 tridiagonal solve,
the value) and the plain FP64 form (the drop-in, the comparison).

Prints one JSON line (rank 0).  The dominant kernel's achieved HBM bandwidth
is its algorithmic bytes per launch divided by its average launch time,
measured with events on the launch stream during the timed steps; its HBM
traffic per launch comes from two one-step rocprofv3 counter passes
(FETCH_SIZE, WRITE_SIZE) run as child processes first (--no-traffic skips them).  The CPU
baseline times the numpy FP64 oracle on the host: one warm-up and three
steady gl06 z40 steps, scaled to gl07 by the cell ratio (the step's work is
proportional to the number of cells).  The accuracy leg runs both forms for
one simulated day from the same state and reports the mixed form's
surface-pressure difference from the FP64 form and both forms' mass and
energy drift.
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


def kernel_bytes(t, nq: int = 1, elem: int = 8) -> dict:
    """Algorithmic HBM bytes per launch of each kernel.

    Each distinct (i, j, k, l) field read or written once counts F bytes
    (elem bytes per point: 8 in the FP64 form, 4 in the mixed form), each of
    the FP64 fields of the tridiagonal solve (matrix, forward-sweep scratch)
    D = 8 bytes per point; each per-column coefficient component counts
    C = elem bytes per column; k-only profiles are ignored.  Field counts
    follow the arrays each kernel touches.
    """
    ncell = t.lall * t.gall_1d ** 2
    npts = ncell * t.kall + 12 * t.kall
    F = float(elem) * npts
    D = 8.0 * npts
    Cc = float(elem) * ncell
    return {
        "prepost": 14 * F,                      # prog 6 + qv in; diag 6 + eth out
        "adv_mom": 13 * F + 24 * Cc,            # prog 5 + diag 4 in; 4 out; div + xyz coefs
        "hdiff_pass1": 12 * F + 7 * Cc,         # 6 in, 6 out; Laplacian coefs
        "hdiff_pass2_lap": 8 * F + 7 * Cc,      # fields 0..3 in, 4 out; Laplacian coefs
        "hdiff_pass2_kh": 5 * F + 36 * Cc,      # fields 4, 5 + rhog in, 2 out; intp + diff coefs
        "divdamp_3d": 7 * F + 36 * Cc,          # 4 in, 3 out; intp + diff coefs
        "divdamp_post": 6 * F + 39 * Cc,        # 3 in, 3 out; intp + diff + xyz
        "vi_path0": 31 * F + 3 * D + 45 * Cc,   # 23 in, 8 out + FP64 matrix 3 out; div + grad + xyz
        "ethtot_mean": 12 * F,                  # prog 5 + eth in; ethtot + 5 means out
        "split_init": 15 * F,                   # prog0 6 + prog 6 in (2 of 3 stages), 6 out
        "vi_path1": 20 * F + 24 * Cc,           # 14 in, 6 out; grad + xyz
        "vi_main_h": 12 * F + 21 * Cc,          # dv 3, prog 3, eth, ethtot, gT_e in; 3 out; div coefs
        "vi_main_v": 8 * F + 5 * D,             # split rho/w/e, gT rho/w, divergences 2 in, rhogw out;
                                                # FP64 matrix 3 in, sweep scratch u, gamma
        "vi_main_post": 26 * F,                 # state 5, split 6, dv 3, ethtot, gT, getot, div 2, rhogw, means 2 in; 4 out
        "prog_update": 18 * F + 3 * Cc,         # 12 in, 6 out
        "trc_vert1": (5 + 2 * nq) * F,          # w mean, rhog_in, frhog + tracers in; flux, rhog1 + tracers out
        "trc_hflux": 16 * F + 24 * Cc,          # 4 mean fluxes in; 3 arc fluxes + 9 centroids out
        "trc_rhog2": 6 * F + 1 * Cc,
        "trc_gradq": 5 * F + 21 * Cc,
        "trc_remap": 20 * F + 4 * Cc,
        "trc_qout": 8 * F + 1 * Cc,
        "trc_apply": 11 * F + 1 * Cc,
        "trc_vert2": 5 * F,
    }


# rocprof kernel name -> library kernel name (the functor type inside k_run / k_pts)
KERNEL_TYPES = {
    "PrePostP": "prepost", "AdvMomP": "adv_mom", "HDiffPassP<0>": "hdiff_pass1", "HDiffPassP<1>": "hdiff_pass2",
    "DivDamp1P": "divdamp_3d", "DivDamp2P": "divdamp_post", "ViPath0P": "vi_path0", "EthTotPMP": "ethtot_mean",
    "SplitInit": "split_init", "Path1P": "vi_path1", "ViMainH": "vi_main_h", "ViMainV<39>": "vi_main_v",
    "ViMainV<0>": "vi_main_v",
    "ViMainPost": "vi_main_post",
    "ProgUpdateP": "prog_update", "HaloGather": "halo_gather", "TrcVert1": "trc_vert1", "TrcVert2": "trc_vert2",
    "TrcHFlux": "trc_hflux", "TrcRhog2": "trc_rhog2", "TrcGradQ": "trc_gradq", "TrcRemap": "trc_remap",
    "TrcQout": "trc_qout", "TrcApply": "trc_apply", "DivDampK<1>": "divdamp_3d", "DivDampK<2>": "divdamp_post",
    "ViMainHC": "vi_main_h", "DiffKhK": "hdiff_pass2_kh", "AdvMomC": "adv_mom",
}


def _kernel_key(name: str):
    import re
    lap = re.search(r"ColStencil<nh::f(?:64|32)::LapC<(\d+),", name)
    if lap:
        return "hdiff_pass1" if lap.group(1) == "6" else "hdiff_pass2_lap"
    m = re.search(r"k_(?:pts|run|tiles|cols)<nh::f(?:64|32)::(?:(?:StagedPointT|ColStencil)<nh::f(?:64|32)::)?"
                  r"(\w+(?:<\d+>)?)\s*>", name)
    return KERNEL_TYPES.get(m.group(1)) if m else None


def hbm_traffic(glevel: int, precision: str, timeout: float = 300.0):
    """HBM bytes per dispatch of each kernel from two rocprofv3 counter passes.

    Runs this benchmark for one step under ``rocprofv3 --pmc FETCH_SIZE`` and
    ``--pmc WRITE_SIZE`` (one counter per pass) as child processes, before this
    process touches the GPU.  Each counter is scaled by what it reads for the
    512 MiB device-to-device copy of the copy probe in the same run.  Returns
    ({kernel: bytes}, {counter: factor}) or (None, reason).
    """
    import csv
    import glob
    import subprocess
    import tempfile
    per = {}
    factors = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(td, ctr)
            cmd = ["rocprofv3", "--pmc", ctr, "--output-format", "csv", "-d", d, "-o", "pmc", "--",
                   sys.executable, os.path.abspath(__file__), "--steps", "1", "--warmup", "0",
                   "--glevel", str(glevel), "--precision", precision, "--no-cpu-baseline", "--no-traffic",
                   "--no-adjoint", "--no-accuracy", "--copy-probe"]
            print(f"bench: counter pass {ctr}", file=sys.stderr, flush=True)
            try:
                r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), capture_output=True,
                                   text=True, timeout=timeout)
            except (OSError, subprocess.TimeoutExpired) as e:
                return None, f"{ctr} pass failed: {e}"
            files = glob.glob(os.path.join(d, "**", "pmc_counter_collection.csv"), recursive=True)
            if r.returncode != 0 or not files:
                return None, f"{ctr} pass failed (rc {r.returncode})"
            rows = list(csv.DictReader(open(files[0])))
            copies = [row for row in rows if "copyBuffer" in row["Kernel_Name"]]
            if not copies:
                return None, "copy probe not found"
            probe = max(copies, key=lambda row: int(row["Dispatch_Id"]))
            factors[ctr] = 524288.0 / float(probe["Counter_Value"])
            acc = {}
            for row in rows:
                key = _kernel_key(row["Kernel_Name"])
                if key:
                    acc.setdefault(key, []).append(float(row["Counter_Value"]))
                    RESOURCES[key] = {"vgpr": int(row["VGPR_Count"]), "scratch": int(row["Scratch_Size"]),
                                      "lds": int(row["LDS_Block_Size"])}
            per[ctr] = {k: sum(v) / len(v) for k, v in acc.items()}
    traffic = {k: 1024.0 * (per["FETCH_SIZE"][k] * factors["FETCH_SIZE"] + per["WRITE_SIZE"].get(k, 0.0)
                            * factors["WRITE_SIZE"]) for k in per["FETCH_SIZE"]}
    return traffic, factors


RESOURCES: dict = {}


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(glevel: int = 6, steady: int = 3) -> dict:
    """The numpy FP64 oracle on one host core: one warm-up large step, then
    ``steady`` timed steps at gl06 z40 (the step's work is proportional to the
    number of cells, so the gl07 rate is the gl06 rate over the cell ratio)."""
    from oracle.dynamics import OracleContext, OracleDynamics
    dts = {5: 1200.0, 6: 600.0}
    case = pkg("setup").build_case(glevel=glevel, vlayer=40, dtl=dts[glevel], seed=1)
    dyn = OracleDynamics(OracleContext(case.geo, case.const, None), case.state)
    dyn.step()
    t0 = time.perf_counter()
    for _ in range(steady):
        dyn.step()
    t_step = (time.perf_counter() - t0) / steady
    cells = 10 * 4 ** glevel + 2
    cells7 = 10 * 4 ** 7 + 2
    t_gl07 = t_step * cells7 / cells
    env = {k: os.environ.get(k, "unset") for k in ("OMP_NUM_THREADS", "MKL_NUM_THREADS", "OPENBLAS_NUM_THREADS")}
    return {
        "value": 300.0 / t_gl07,
        "unit": "SDPD",
        "cores": 1,
        "kind": "port",
        "sample": (f"numpy FP64 oracle, gl{glevel:02d} z40: one warm-up large step, then {steady} steady steps "
                   f"of {t_step:.2f} s each on one host core ({_cpu_model()}, os.cpu_count()={os.cpu_count()}, "
                   f"{', '.join(f'{k}={v}' for k, v in env.items())}; numpy's elementwise kernels run on one "
                   f"thread); gl07 step = {t_gl07:.1f} s by the cell ratio {cells7}/{cells}, SDPD at dt 300 s"),
        "gl_measured": glevel,
        "s_per_step_measured": t_step,
    }


def sizing_leg(hipdyn, glevel: int, steps: int = 10, warmup: int = 2) -> dict:
    """The mixed form at a larger grid: ms per step and the device memory it holds."""
    import torch
    dt = 300.0 * 2.0 ** (7 - glevel)
    print(f"bench: sizing: setting up gl{glevel:02d} z40", file=sys.stderr, flush=True)
    case = pkg("setup").build_case(glevel=glevel, vlayer=40, dtl=dt, seed=1)
    torch.cuda.empty_cache()
    free0, total = torch.cuda.mem_get_info()
    ctx = hipdyn.HipContext(case.geo, case.const, device="cuda:0", precision="mixed")
    try:
        dyn = hipdyn.HipDynamics(ctx, case.state)
        for _ in range(warmup):
            dyn.dynamics_step()
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        t0 = time.perf_counter()
        for _ in range(steps):
            dyn.dynamics_step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        b = dyn.budget()
        ok = all(map(lambda v: v == v, b.values()))
    finally:
        ctx.close()
    res = {"workload": f"gl{glevel:02d} z40 ({10 * 4 ** glevel + 2} cells), dt {dt:g} s, mixed form",
           "ms_per_step": ms, "sdpd": dt / (ms / 1000.0), "device_mem_GB": (free0 - free1) / 1e9,
           "device_total_GB": total / 1e9, "budget_finite": ok}
    print("bench: sizing " + json.dumps(res), file=sys.stderr, flush=True)
    return res


def accuracy_leg(hipdyn, case, nsteps: int) -> dict:
    """Both forms over ``nsteps`` steps from the same state: surface-pressure
    difference of the mixed form from the FP64 form (BNDCND_pre_sfc,
    mod_bndcnd.py:521, over every regional interior column and the poles), the
    largest and RMS difference of each prognostic relative to its largest
    value, and each form's drift of dry mass and total energy (compensated
    global sums, Embudget._store_var, mod_embudget.py:82-139, which leave the
    pole cells out) and of the dry mass over every cell, poles included."""
    import numpy as np
    t = case.topo
    g = t.gall_1d
    geo = case.geo
    ks = slice(t.kmin, t.kmax + 1)

    def dry_mass_all_cells(s):
        """Dry mass over every prognostic cell: the regional interior and the two
        pole cells (the budget's _gsum leaves the pole cells out, so its dry mass
        moves with the exchange across their faces)."""
        P, Pp, Q, Qp = s["PROG"], s["PROG_pl"], s["PROGq"], s["PROGq_pl"]
        reg = P[1:g - 1, 1:g - 1, ks, :, 0] - Q[1:g - 1, 1:g - 1, ks, :, 0]
        pl = Pp[0, ks, :, 0] - Qp[0, ks, :, 0]
        return float((reg.astype(np.float64) * geo.VOLUME[1:g - 1, 1:g - 1, ks]).sum()
                     + (pl.astype(np.float64) * geo.VOLUME_pl[0, ks]).sum())

    out, ps, st = {}, {}, {}
    for f in ("float64", "mixed"):
        ctx = hipdyn.HipContext(case.geo, case.const, device="cuda:0", precision=f)
        try:
            dyn = hipdyn.HipDynamics(ctx, case.state)
            m0 = dry_mass_all_cells(dyn.state())
            b0 = dyn.budget()
            for n in range(nsteps):
                dyn.dynamics_step()
                if (n + 1) % 96 == 0:
                    ctx.sync()
                    print(f"bench: accuracy {f}: step {n + 1}/{nsteps}", file=sys.stderr, flush=True)
            b1 = dyn.budget()
            ps[f] = dyn.sfc_pressure().cpu().numpy()
            st[f] = dyn.state()
        finally:
            ctx.close()
        out[FORMS[f][0]] = {"dry_mass_drift_rel_all_cells": (dry_mass_all_cells(st[f]) - m0) / m0,
                            "dry_mass_drift_rel": (b1["dry"] - b0["dry"]) / b0["dry"],
                            "total_energy_drift_rel": (b1["tot"] - b0["tot"]) / b0["tot"],
                            "budget_start": b0, "budget_end": b1}
    nreg = t.lall * g * g
    inner = np.zeros((t.lall, g, g), dtype=bool)
    inner[:, 1:g - 1, 1:g - 1] = True
    cols = np.concatenate([inner.reshape(-1), np.array([True] + [False] * 5 + [True] + [False] * 5)])
    assert cols.size == nreg + 12
    d = (ps["mixed"] - ps["float64"])[cols]
    out["ps_rms_diff_hPa"] = float(np.sqrt(np.mean(d * d)) / 100.0)
    out["ps_max_abs_diff_hPa"] = float(np.abs(d).max() / 100.0)
    out["ps_fp64_range_hPa"] = [float(ps["float64"][cols].min() / 100.0), float(ps["float64"][cols].max() / 100.0)]
    a = st["float64"]["PROG"][1:g - 1, 1:g - 1, t.kmin:t.kmax + 1].reshape(-1, 6)
    b = st["mixed"]["PROG"][1:g - 1, 1:g - 1, t.kmin:t.kmax + 1].reshape(-1, 6).astype(np.float64)
    scale = np.abs(a).max(axis=0)
    names = ("rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge")
    out["prog_max_rel_diff"] = dict(zip(names, [float(e) for e in np.abs(a - b).max(axis=0) / scale]))
    out["prog_rms_rel_diff"] = dict(zip(names, [float(e) for e in np.sqrt(np.mean((a - b) ** 2, axis=0)) / scale]))
    out["workload"] = (f"gl{t.glevel:02d} z{t.kmax - t.kmin + 1}, dt {case.const.DTL:g} s, {nsteps} steps "
                       f"({nsteps * case.const.DTL / 86400.0:g} simulated days), mixed form against the FP64 form")
    return out


FORMS = {"float64": ("fp64", "f64", 8), "mixed": ("mixed", "f32+f64-solve", 4)}


def run_form(hipdyn, case, precision: str, steps: int, warmup: int, candidates, prof_steps: int = 3):
    """Warm up; run `prof_steps` steps with every kernel timed (the per-kernel
    table; an event pair around every launch adds to the wall time, so this run
    is not the timed one); pick the dominant kernel among `candidates` (most
    time per step); then time `steps` steps with the timers on that kernel
    only, so its average launch time comes from events over the timed region.
    Returns (ms per step, {kernel: (ms per step, launches per step)}, dominant
    kernel, its average launch ms in the timed steps, state, budget)."""
    import numpy as np
    import torch
    ctx = hipdyn.HipContext(case.geo, case.const, device="cuda:0", precision=precision)
    dyn = hipdyn.HipDynamics(ctx, case.state)
    for _ in range(warmup):
        dyn.dynamics_step()
    ctx.sync()
    print(f"bench: {precision}: {warmup} warmup steps done", file=sys.stderr, flush=True)
    ctx.prof_report()                      # clear
    ctx.prof_enable(True)
    for _ in range(prof_steps):
        dyn.dynamics_step()
    ctx.sync()
    ctx.prof_enable(False)
    prof = {k: (v[0] / prof_steps, v[1] / prof_steps) for k, v in ctx.prof_report().items()}
    dom = max((k for k in prof if k in candidates), key=lambda k: prof[k][0])
    ctx.prof_only(dom)
    ctx.prof_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dyn.dynamics_step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.prof_enable(False)
    tot, n = ctx.prof_report()[dom]
    ctx.prof_only(None)
    st = dyn.state()
    budget = dyn.budget()
    ctx.close()
    g = case.topo.gall_1d
    if not np.all(np.isfinite(st["PROG"][1:g - 1, 1:g - 1, 1:-1])):
        print(f"bench: {precision}: state is not finite", file=sys.stderr)
        sys.exit(1)
    ms = (t1 - t0) * 1000.0 / steps
    print(f"bench: {precision}: {ms:.2f} ms per step", file=sys.stderr, flush=True)
    return ms, prof, dom, tot / n, st, budget


def adjoint_leg(hipdyn, precision: str, nsteps: int = 18, reps: int = 2, ref=None, store: str = "tape",
                same=None) -> dict:
    """Initial-state sensitivity at gl06 z40 (dt 600 s): J = sum over the 30-50N,
    0-20E box of surface pressure squared after ``nsteps`` steps, its gradient
    through the adjoint window.  ``store="tape"``: per-step tapes of the forward
    steps, the adjoint steps recompute only the diagnostics; ``"stages"``: only
    the three RK-stage inputs of every step are kept and the adjoint recomputes
    each stage.  Times the plain forward steps, the checkpointed forward
    window and the backward pass (best of ``reps``), checks the gradient of the
    FP64 form by a central finite difference along one seeded direction (the
    mixed form's against the FP64 gradient ``ref``; a "stages" run's against the
    same form's "tape" gradient ``same``), and reports the device memory in use
    with the window's checkpoints alive (hipMemGetInfo)."""
    import numpy as np
    import torch
    adj = pkg("adjoint")
    case = pkg("setup").build_case(glevel=6, vlayer=40, dtl=600.0, seed=0)
    ctx = hipdyn.HipContext(case.geo, case.const, device="cuda:0", precision=precision)
    try:
        dyn = hipdyn.HipDynamics(ctx, case.state)
        m = adj.box_mask(dyn)
        x0 = dyn.prog.clone()
        q0 = dyn.progq.clone()

        def plain():
            dyn.prog.copy_(x0)
            dyn.progq.copy_(q0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(nsteps):
                dyn.dynamics_step()
            torch.cuda.synchronize()
            return time.perf_counter() - t0

        mem = {}

        def window():
            x = x0.clone().requires_grad_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ps = adj.sfc_pressure(dyn, adj.forecast(dyn, x, nsteps, store=store))
            J = (m * ps * ps).sum()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            if "peak" not in mem:
                # device memory in use with every step's tape alive (cached free blocks released)
                torch.cuda.empty_cache()
                free, total = torch.cuda.mem_get_info()
                mem["peak"] = (total - free) / 1e9
            J.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            return float(J.detach()), x.grad.detach(), t1 - t0, t2 - t1

        plain()
        window()
        t_plain = min(plain() for _ in range(reps))
        runs = [window() for _ in range(reps)]
        Jv, grad = runs[-1][0], runs[-1][1]
        t_fwd = min(r[2] for r in runs)
        t_bwd = min(r[3] for r in runs)
        tape_gb = dyn.tape_bytes() / 1e9
        # central difference along a seeded direction scaled per variable
        rng = np.random.default_rng(7)
        sc = x0.abs().amax(dim=1, keepdim=True).clone()
        sc[3] = sc[1]
        sc[4] = 0.01 * sc[1]
        v = torch.from_numpy(rng.standard_normal(tuple(x0.shape))).to(x0.device, x0.dtype) * sc
        ad = float((grad.double() * v.double()).sum())
        check = {}
        if same is not None:
            gs = same["grad"]
            per = (grad.double() - gs).abs().amax(dim=1) / gs.abs().amax(dim=1)
            check = {"vs_tape_max_rel_err": float(per.max())}
        elif ref is None:
            # the step size balances truncation (eps^2) against round-off
            eps = 1e-6

            def J_of(x):
                dyn.prog.copy_(x)
                dyn.progq.copy_(q0)
                for _ in range(nsteps):
                    dyn.dynamics_step()
                ps = dyn.sfc_pressure()
                return float((m * ps * ps).sum())

            fd = (J_of(x0 + eps * v) - J_of(x0 - eps * v)) / (2 * eps)
            check = {"fd_rel_err": abs(fd - ad) / abs(fd), "fd_eps": eps}
        else:
            g64 = ref["grad"]
            a64 = float((g64 * v.double()).sum())
            gd = grad.double()
            per = (gd - g64).abs().amax(dim=1) / g64.abs().amax(dim=1)
            check = {"vs_fp64_dir_rel_err": abs(ad - a64) / abs(a64),
                     "vs_fp64_max_rel_err": dict(zip(("rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge"),
                                                     [float(e) for e in per]))}
    finally:
        ctx.close()
    res = {"workload": f"gl06 z40, dt 600 s, {nsteps} steps, J = sum over 30-50N 0-20E of ps^2",
           "store": store,
           "forward_ms_per_step": 1000.0 * t_plain / nsteps,
           "window_forward_ms_per_step": 1000.0 * t_fwd / nsteps,
           "adjoint_ms_per_step": 1000.0 * t_bwd / nsteps,
           "adjoint_over_forward": t_bwd / t_plain,
           "J": Jv, **check,
           "tape_GB_per_step": tape_gb if store == "tape" else None, "peak_device_mem_GB": mem["peak"]}
    print(f"bench: adjoint {precision}: " + json.dumps(res), file=sys.stderr, flush=True)
    res["grad"] = grad.double()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--glevel", type=int, default=7)
    ap.add_argument("--precision", choices=("both", "mixed", "float64"), default="both",
                    help="forms to time; the value is the mixed form's SDPD when it is timed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure HBM traffic per kernel")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint window at gl06 z40")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint window at gl06 z40 (both forms) and print its JSON")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the one-simulated-day run of both forms (surface-pressure difference, drift)")
    ap.add_argument("--accuracy-steps", type=int, default=None,
                    help="steps of the accuracy leg (default: one simulated day)")
    ap.add_argument("--sizing-glevel", type=int, default=None,
                    help="also time the mixed form at this glevel (e.g. 8) and report its device memory")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the run, copy 512 MiB device-to-device once (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU path is a design, not a run", file=sys.stderr)
        sys.exit(2)
    gl = args.glevel
    forms = ["float64", "mixed"] if args.precision == "both" else [args.precision]
    if args.adjoint_only:
        import torch
        if not torch.cuda.is_available():
            print("bench.py: no GPU visible", file=sys.stderr)
            sys.exit(1)
        hipdyn = pkg("hipdyn")
        adj, ref = {}, None
        for f in forms:
            r = adjoint_leg(hipdyn, f, ref=ref)
            if f == "float64":
                ref = r
            adj[FORMS[f][0]] = r
            adj[FORMS[f][0] + "_stages"] = adjoint_leg(hipdyn, f, store="stages", same=r)
        for r in adj.values():
            r.pop("grad", None)
        print(json.dumps({"metric": "adjoint/forward step cost at gl06 z40 over 18 steps", "adjoint": adj}),
              flush=True)
        return
    main_form = forms[-1]
    traffic, factors = (None, "not measured")
    if not args.no_traffic:
        # before this process initialises the GPU
        traffic, factors = hbm_traffic(gl, main_form)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    dt = {5: 1200.0, 6: 600.0, 7: 300.0, 8: 150.0}[gl]
    print(f"bench: setting up gl{gl:02d} z40 dt {dt:g} s", file=sys.stderr, flush=True)
    case = pkg("setup").build_case(glevel=gl, vlayer=40, dtl=dt, seed=1)
    hipdyn = pkg("hipdyn")
    res = {f: run_form(hipdyn, case, f, args.steps, args.warmup, kernel_bytes(case.topo, elem=FORMS[f][2]))
           for f in forms}
    ms_per_step, prof, dom, dom_ms, st, budget = res[main_form]
    sdpd = dt / (ms_per_step / 1000.0)
    tag, dtype, elem = FORMS[main_form]
    kb = kernel_bytes(case.topo, elem=elem)
    avg_s = dom_ms / 1000.0
    achieved = kb[dom] / avg_s / 1e9
    kernels = {k: {"ms_per_step": v[0], "launches_per_step": v[1], "avg_us": 1000.0 * v[0] / v[1],
                   "GBps": (kb[k] / (v[0] / v[1] / 1000.0) / 1e9) if k in kb else None}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])}
    if isinstance(traffic, dict):
        for k, v in kernels.items():
            if k in traffic:
                v["traffic_GB"] = traffic[k] / 1e9
            if k in RESOURCES:
                v.update(RESOURCES[k])
    print(f"bench: {main_form} kernels " + json.dumps(kernels), file=sys.stderr, flush=True)
    form_info = {}
    for f, (ms, _, _, _, _, b) in res.items():
        form_info[FORMS[f][0]] = {"ms_per_step": ms, "sdpd": dt / (ms / 1000.0), "budget": b}
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
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": f"gl{gl:02d} z40 dynamics step, dt {dt:g} s, {tag} HIP form, "
                               "dry balanced zonal flow + seeded theta bumps (seed 1), flat topography",
                   "glevel": gl, "vlayer": 40, "dt": dt, "cells": 10 * 4 ** gl + 2, "form": tag,
                   "parallelism": "single GPU, 40 regions"},
        "forms": form_info,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (traffic.get(dom) if isinstance(traffic, dict) else None),
                     "traffic_note": (f"rocprofv3 FETCH_SIZE x {factors['FETCH_SIZE']:.3f} + WRITE_SIZE x "
                                      f"{factors['WRITE_SIZE']:.3f} per dispatch, factors from a 512 MiB copy"
                                      if isinstance(traffic, dict) else factors),
                     "algorithmic_bytes_per_launch": kb[dom], "avg_launch_us": avg_s * 1e6},
    }
    if "fp64" in form_info and "mixed" in form_info:
        out["speedup_mixed_vs_fp64"] = form_info["mixed"]["sdpd"] / form_info["fp64"]["sdpd"]
        import numpy as np
        t = case.topo
        a = res["float64"][4]["PROG"][1:-1, 1:-1, t.kmin:t.kmax + 1]
        b = res["mixed"][4]["PROG"][1:-1, 1:-1, t.kmin:t.kmax + 1].astype(np.float64)
        scale = np.abs(a).reshape(-1, 6).max(axis=0)
        err = np.abs(a - b).reshape(-1, 6).max(axis=0) / scale
        out["mixed_vs_fp64_max_rel_err"] = dict(zip(("rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge"),
                                                    [float(e) for e in err]))
    if "fp64" in form_info and "mixed" in form_info and not args.no_accuracy:
        nacc = args.accuracy_steps or int(round(86400.0 / dt))
        out["accuracy"] = accuracy_leg(hipdyn, case, nacc)
        print("bench: accuracy " + json.dumps(out["accuracy"]), file=sys.stderr, flush=True)
        # the same day with the FP64 hydrostatic basic state (ref_type 'INPUT': the
        # resting atmosphere), which the mixed form subtracts in FP64
        case_bs = pkg("setup").build_case(glevel=gl, vlayer=40, dtl=dt, seed=1, ref_type="INPUT")
        out["accuracy_basic_state"] = accuracy_leg(hipdyn, case_bs, nacc)
        print("bench: accuracy with the basic state " + json.dumps(out["accuracy_basic_state"]), file=sys.stderr,
              flush=True)
    if args.sizing_glevel:
        out["sizing"] = sizing_leg(hipdyn, args.sizing_glevel)
    if not args.no_adjoint:
        adj = {}
        ref = None
        for f in forms:
            r = adjoint_leg(hipdyn, f, ref=ref)
            if f == "float64":
                ref = r
            adj[FORMS[f][0]] = r
            # per-step checkpoints of the stage inputs only, each stage recomputed
            adj[FORMS[f][0] + "_stages"] = adjoint_leg(hipdyn, f, store="stages", same=r)
        for r in adj.values():
            r.pop("grad", None)
        out["adjoint"] = adj
        out["adjoint_over_forward"] = adj[FORMS[main_form][0]]["adjoint_over_forward"]
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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
