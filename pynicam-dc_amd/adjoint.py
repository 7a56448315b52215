"""Initial-state sensitivity of a forecast functional: the adjoint of the
dynamics step behind ``torch.autograd.Function``.

The reference has no adjoint; the forward operator differentiated here is
``Dyn.dynamics_step`` (nhm/dynamics/mod_dynamics.py:1725), run through
``HipDynamics.dynamics_step``.  ``forecast(dyn, prog, nsteps)`` advances a
differentiable copy of the state, keeping per step either its tape (every
intermediate field the transposes read: ``nh_dyn_step_tape``; the default)
or only its RK-stage inputs (``nh_dyn_step_record``, less memory); its
backward pass calls the library's transposed step once per step, newest
first (``nh_dyn_step_adj_tape`` or ``nh_dyn_step_adj_rec``, which then
recomputes the stages and small steps inside).
``sfc_pressure(dyn, prog)`` is the surface-pressure diagnostic
(Bndc.BNDCND_pre_sfc, nhm/share/mod_bndcnd.py:521) with its transpose, and
``box_mask`` selects the cells of a latitude/longitude box, so that

    J = (box_mask(...) * sfc_pressure(dyn, forecast(dyn, prog0, 18)) ** 2).sum()
    J.backward()

gives dJ/dprog0.  Tracer masses are carried along the trajectory but held
fixed in the transpose: the gradient is exact when they are zero (dry runs).

``tracer_forecast(run, progq, nsteps)`` is the same for the offline tracer
transport (TIME_integ_type 'TRCADV', mod_dynamics.py:2030): the tracers after
``nsteps`` steps of prescribed winds, differentiable with respect to the
initial tracers (``nh_tracer_step_adj`` per step).
"""
from __future__ import annotations

import numpy as np

from . import layout


def _torch():
    import torch
    return torch


def _check_state(dyn, prog):
    ctx = dyn.ctx
    if prog.dtype != ctx.dtype or prog.device != ctx.device or tuple(prog.shape) != (6, ctx.NF):
        raise ValueError(f"expected a ({6}, {ctx.NF}) {ctx.dtype} tensor on {ctx.device}")


def _make_functions():
    torch = _torch()

    class Forecast(torch.autograd.Function):
        """prog -> state after ``nsteps`` dynamics steps (tracers advance alongside)."""

        @staticmethod
        def forward(fctx, prog, dyn, nsteps, store):
            _check_state(dyn, prog)
            saved_prog, saved_q = dyn.prog, dyn.progq
            progq = dyn.progq.clone()
            checkpoints = []
            work = prog.detach().clone().contiguous()
            dyn.prog, dyn.progq = work, progq
            try:
                for _ in range(nsteps):
                    q_in = progq.clone()
                    if store == "tape":
                        rec = dyn.new_tape()
                        dyn.dynamics_step_tape(rec)
                    else:
                        rec = dyn.ctx.empty(18)
                        dyn.dynamics_step_record(rec)
                    checkpoints.append((rec, q_in))
            finally:
                dyn.prog, dyn.progq = saved_prog, saved_q
            fctx.dyn = dyn
            fctx.store = store
            fctx.checkpoints = checkpoints
            fctx.progq_out = progq
            return work

        @staticmethod
        def backward(fctx, grad):
            dyn = fctx.dyn
            gb = grad.detach().to(dtype=dyn.ctx.dtype).contiguous().clone()
            for rec, q in reversed(fctx.checkpoints):
                if fctx.store == "tape":
                    dyn.dynamics_step_adj_tape(rec, q, gb)
                else:
                    dyn.dynamics_step_adj_rec(rec, q, gb)
            fctx.checkpoints = None
            return gb, None, None, None

    class SfcPressure(torch.autograd.Function):
        """prog -> surface pressure per column (FP64, nreg + 12)."""

        @staticmethod
        def forward(fctx, prog, dyn, progq):
            _check_state(dyn, prog)
            prog = prog.detach().contiguous()
            fctx.dyn = dyn
            fctx.save_for_backward(prog, progq)
            return dyn.sfc_pressure(prog, progq)

        @staticmethod
        def backward(fctx, grad):
            prog, progq = fctx.saved_tensors
            gb = torch.zeros_like(prog)
            fctx.dyn.sfc_pressure_adj(prog, progq, grad.detach(), gb)
            return gb, None, None

    class TracerForecast(torch.autograd.Function):
        """progq -> tracers after ``nsteps`` offline tracer steps (winds prescribed)."""

        @staticmethod
        def forward(fctx, progq, run, nsteps):
            ctx = run.ctx
            if progq.dtype != ctx.dtype or progq.device != ctx.device or tuple(progq.shape) != tuple(run.progq.shape):
                raise ValueError(f"expected a {tuple(run.progq.shape)} {ctx.dtype} tensor on {ctx.device}")
            saved_q, saved_prog, saved_time = run.progq, run.prog.clone(), run.time
            work = progq.detach().clone().contiguous()
            checkpoints = []
            run.progq = work
            try:
                for _ in range(nsteps):
                    checkpoints.append((run.prog.clone(), work.clone()))
                    run.dynamics_step()
            finally:
                run.progq = saved_q
                run.prog.copy_(saved_prog)
                run.time = saved_time
            fctx.run = run
            fctx.checkpoints = checkpoints
            return work

        @staticmethod
        def backward(fctx, grad):
            run = fctx.run
            gb = grad.detach().to(dtype=run.ctx.dtype).contiguous().clone()
            for prog, q in reversed(fctx.checkpoints):
                run.tracer_step_adj(prog, q, gb)
            fctx.checkpoints = None
            return gb, None, None

    return Forecast, SfcPressure, TracerForecast


_FN = {}


def _fn(name):
    if not _FN:
        _FN["forecast"], _FN["sfc"], _FN["tracer"] = _make_functions()
    return _FN[name]


def forecast(dyn, prog, nsteps: int, store: str = "tape"):
    """State after ``nsteps`` steps of ``dyn`` from ``prog`` (``dyn.progq`` as the initial tracers).

    Differentiable with respect to ``prog``; ``dyn``'s own state is left unchanged.
    ``store``: "tape" keeps every step's full tape (``nh_tape_bytes`` per step;
    the backward pass recomputes only diagnostics), "stages" only the RK-stage
    inputs (18 fields per step; the backward pass recomputes each step's stages).
    """
    if store not in ("tape", "stages"):
        raise ValueError("store must be 'tape' or 'stages'")
    return _fn("forecast").apply(prog, dyn, int(nsteps), store)


def tracer_forecast(run, progq, nsteps: int):
    """Tracers after ``nsteps`` steps of the offline tracer run ``run`` (a
    ``HipOfflineTracer``) from ``progq``, differentiable with respect to ``progq``.

    The winds are prescribed (they do not depend on the tracers); ``run``'s own
    tracers, winds and time are left unchanged.  Each step keeps its winds and
    input tracers; the backward pass calls ``nh_tracer_step_adj`` per step,
    newest first, with the limiters' selections of the forward pass.
    """
    return _fn("tracer").apply(progq, run, int(nsteps))


def sfc_pressure(dyn, prog, progq=None):
    """Surface pressure per column of ``prog`` (differentiable with respect to ``prog``)."""
    return _fn("sfc").apply(prog, dyn, dyn.progq if progq is None else progq)


def box_mask(dyn, lat_deg=(30.0, 50.0), lon_deg=(0.0, 20.0)):
    """1.0 on the cells (interior regional points and pole centres) inside a
    latitude/longitude box, 0.0 elsewhere; FP64 per column of ``sfc_pressure``."""
    torch = _torch()
    ctx = dyn.ctx
    geo = ctx.geo
    t = geo.topo
    g = t.gall_1d
    lat = np.degrees(layout._soa(geo.lat)[0])
    lon = np.degrees(layout._soa(geo.lon)[0])
    lon = np.where(lon < -180.0, lon + 360.0, np.where(lon >= 180.0, lon - 360.0, lon))
    idx = np.arange(lat.size)
    i = idx % g
    j = (idx // g) % g
    inner = (i >= 1) & (i <= g - 2) & (j >= 1) & (j <= g - 2)
    inbox = (lat >= lat_deg[0]) & (lat <= lat_deg[1]) & (lon >= lon_deg[0]) & (lon <= lon_deg[1])
    m = np.zeros(ctx.ncol)
    m[:lat.size] = (inner & inbox).astype(np.float64)
    lat_pl = np.degrees(np.asarray(geo.lat_pl).reshape(6, 2))
    for lp in range(2):
        if lat_deg[0] <= lat_pl[0, lp] <= lat_deg[1]:
            m[lat.size + lp * 6] = 1.0
    return torch.from_numpy(m).to(ctx.device)
