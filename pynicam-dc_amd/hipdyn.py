"""Device-resident dynamical core behind the nicamhip C-ABI.

``HipContext`` owns one library context (geometry, halo map, step workspace)
on one device, in one of two forms: ``precision="float64"`` (the drop-in) or
``"mixed"`` (float fields and stencils; FP64 tridiagonal solve, global sums and
time steps).  ``HipDynamics`` owns the prognostic state as torch tensors in
the device layout and advances it with ``dynamics_step()``, the counterpart of
``Dyn.dynamics_step(msc)`` (nhm/dynamics/mod_dynamics.py:1725-2741).
Operators (``OPRT_laplacian``, ``OPRT_gradient``, divergence, halo fill) are
exposed with reference-style argument order for tests and callers.

torch provides device memory and the stream only; every computation is a
nicamhip kernel.  Host (numpy) arrays are accepted at the boundary in the
reference layout and converted once.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import layout, native


def _torch():
    import torch
    return torch


class HipContext:
    """Library context for one grid on one device.

    ``device`` is a torch device string.  The default is the GPU; tests of the
    kernel source on the host pass ``device="cpu"`` together with the host
    test build of the library (``lib_path``).
    """

    PRECISIONS = {"float64": native.PREC_FP64, "mixed": native.PREC_MIXED}

    def __init__(self, geo, const, device: str = "cuda:0", lib_path: str | None = None,
                 precision: str = "float64"):
        torch = _torch()
        if precision not in self.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(self.PRECISIONS)}, not {precision!r}")
        self.precision = precision
        self.mixed = precision == "mixed"
        self.dtype = torch.float32 if self.mixed else torch.float64
        self.geo = geo
        self.const = const
        self.topo = t = geo.topo
        self.device = torch.device(device)
        if self.device.type == "cuda" and not torch.cuda.is_available():
            raise RuntimeError("backend='hip' needs a GPU (torch.cuda.is_available() is False)")
        self.lib = native.load(lib_path)
        lib = self.lib
        self.dims = native.NhDims(t.gall_1d, t.kall, t.kmin, t.kmax, t.lall)
        self.NF = int(lib.nh_field_size(C.byref(self.dims)))
        self.ncol = t.lall * t.gall_1d * t.gall_1d + 12   # regional columns, then the pole columns
        if self.NF != layout.field_size(t.gall_1d, t.kall, t.lall):
            raise RuntimeError("field size mismatch between library and host layout")
        c = const
        self.consts = native.NhConsts(c.RADIUS, c.OHM, c.GRAV, c.Rdry, c.CVdry, c.Rvap, float(c.CVW[c.I_QV]),
                                      c.ALPHA, c.GAMMA_H, c.HDIFF_FACT_RHO, c.ALPHA_D,
                                      1 if c.THUBURN_LIM else 0)
        if c.LAP_ORDER_HDIFF != 2 or c.LAP_ORDER_DIVDAMP != 2:
            raise NotImplementedError("backend='hip' implements lap_order 2 for hyperdiffusion and divergence damping")
        dev_index = self.device.index if self.device.type == "cuda" and self.device.index is not None else 0
        if self.device.type == "cuda":
            torch.cuda.set_device(dev_index)
        h = C.c_void_p()
        native.check(lib, lib.nh_create_ex(C.byref(self.dims), C.byref(self.consts), dev_index,
                                           self.PRECISIONS[precision], C.byref(h)), "nh_create")
        self.handle = h
        # geometry (host arrays kept alive until the synchronous upload returns)
        arrs = layout.geometry_arrays(geo)
        g = native.NhGeometry()
        dp = C.POINTER(C.c_double)
        for name, _ in native.NhGeometry._fields_:
            a = arrs[name]
            ptr = a.ctypes.data_as(C.POINTER(C.c_int32) if name == "sgp" else dp)
            setattr(g, name, ptr)
        native.check(lib, lib.nh_upload_geometry(self.handle, C.byref(g)), "nh_upload_geometry")
        dst = np.ascontiguousarray(t.halo_dst, dtype=np.int32)
        src = np.ascontiguousarray(t.halo_src, dtype=np.int32)
        ip = C.POINTER(C.c_int32)
        native.check(lib, lib.nh_upload_halo_map(self.handle, dst.ctypes.data_as(ip), src.ctypes.data_as(ip),
                                                 dst.size), "nh_upload_halo_map")

    # ------------------------------------------------------------------
    def fn(self, name: str):
        """The library entry point for this context's field precision (``name`` or ``name_f32``)."""
        return getattr(self.lib, name + "_f32" if self.mixed else name)

    def stream(self):
        if self.device.type == "cuda":
            return C.c_void_p(_torch().cuda.current_stream(self.device).cuda_stream)
        return C.c_void_p(None)

    def empty(self, nv: int):
        return _torch().empty((nv, self.NF), dtype=self.dtype, device=self.device)

    def to_device(self, var: np.ndarray, var_pl: np.ndarray):
        """Reference layout (g,g,kall,l,nv)+(6,kall,2,nv) -> device tensor (nv, NF)."""
        return _torch().from_numpy(layout.pack(var, var_pl)).to(device=self.device, dtype=self.dtype)

    def to_host(self, flat):
        t = self.topo
        return layout.unpack(flat.detach().cpu().numpy(), t.gall_1d, t.kall, t.lall)

    def sync(self):
        if self.device.type == "cuda":
            _torch().cuda.synchronize(self.device)

    def prof_enable(self, on: bool = True):
        """Start/stop per-launch kernel timers (events on the launch stream)."""
        native.check(self.lib, self.lib.nh_prof_enable(self.handle, 1 if on else 0), "nh_prof_enable")

    def prof_only(self, kernel: str | None):
        """Time only the launches of ``kernel`` (its report name); None = every kernel."""
        native.check(self.lib, self.lib.nh_prof_only(self.handle, (kernel or "").encode()), "nh_prof_only")

    def prof_report(self, max_kernels: int = 64) -> dict:
        """{kernel name: (total ms, launches)} of the launches timed since the last report."""
        names = C.create_string_buffer(32 * max_kernels)
        ms = (C.c_double * max_kernels)()
        cnt = (C.c_int64 * max_kernels)()
        n = C.c_int32()
        native.check(self.lib, self.lib.nh_prof_report(self.handle, max_kernels, names, ms, cnt, C.byref(n)),
                     "nh_prof_report")
        raw = names.raw
        out = {}
        for q in range(n.value):
            nm = raw[32 * q:32 * (q + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (ms[q], int(cnt[q]))
        return out

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.nh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    # operators on device tensors (device layout)
    def COMM_data_transfer(self, fields):
        """In-place halo/pole fill of a (nv, NF) tensor (share/mod_comm.py:1066)."""
        self._check_tensor(fields)
        nv = fields.shape[0]
        ptrs = (C.c_void_p * nv)(*[fields[q].data_ptr() for q in range(nv)])
        native.check(self.lib, self.fn("nh_halo_fill")(self.handle, ptrs, nv, self.stream()), "nh_halo_fill")

    def OPRT_laplacian(self, scl, out=None):
        """7-point Laplacian (share/mod_oprt.py:1429)."""
        self._check_tensor(scl)
        out = _torch().zeros_like(scl) if out is None else out
        native.check(self.lib, self.fn("nh_laplacian")(self.handle, C.c_void_p(scl.data_ptr()),
                                                     C.c_void_p(out.data_ptr()), self.stream()), "nh_laplacian")
        return out

    def divergence(self, vx, vy, vz, out=None):
        """7-point divergence (kernels/fluxconv.py:238)."""
        for a in (vx, vy, vz):
            self._check_tensor(a)
        out = _torch().zeros_like(vx) if out is None else out
        native.check(self.lib, self.fn("nh_divergence")(self.handle, *(C.c_void_p(a.data_ptr()) for a in
                                                                     (vx, vy, vz, out)), self.stream()),
                     "nh_divergence")
        return out

    def OPRT_gradient(self, scl, out=None):
        """7-point gradient (share/mod_oprt.py:1387); out is (3, NF)."""
        self._check_tensor(scl)
        out = _torch().zeros((3, self.NF), dtype=scl.dtype, device=scl.device) if out is None else out
        native.check(self.lib, self.fn("nh_gradient")(self.handle, C.c_void_p(scl.data_ptr()),
                                                    *(C.c_void_p(out[q].data_ptr()) for q in range(3)),
                                                    self.stream()), "nh_gradient")
        return out

    def _check_tensor(self, a):
        torch = _torch()
        if a.dtype != self.dtype or a.device != self.device or a.shape[-1] != self.NF or not a.is_contiguous():
            raise ValueError(f"expected a contiguous {self.dtype} tensor (..., {self.NF}) on {self.device}")


class HipDynamics:
    """Prognostic state on the device and the large step (mod_dynamics.py:1725)."""

    def __init__(self, ctx: HipContext, state: dict):
        self.ctx = ctx
        c = ctx.const
        self.prog = ctx.to_device(state["PROG"], state["PROG_pl"]).contiguous()
        self.progq = ctx.to_device(state["PROGq"], state["PROGq_pl"]).contiguous()
        if self.prog.shape[0] != 6 or self.progq.shape[0] < 1:
            raise ValueError("PROG needs 6 variables and PROGq at least qv")
        self.dtl = float(c.DTL)
        self.sstep_max = int(c.SSTEP_MAX)

    @staticmethod
    def _check_state_pair(progq, *progs):
        """PROG-shaped arguments hold the 6 prognostics; PROGq at least qv."""
        for p in progs:
            if p.dim() != 2 or p.shape[0] != 6:
                raise ValueError("a PROG argument needs 6 fields")
        if progq.dim() != 2 or progq.shape[0] < 1:
            raise ValueError("PROGq needs at least one tracer field (qv)")

    def dynamics_step(self):
        lib = self.ctx.lib
        native.check(lib, self.ctx.fn("nh_dyn_step")(self.ctx.handle, C.c_void_p(self.prog.data_ptr()),
                                          C.c_void_p(self.progq.data_ptr()), self.progq.shape[0],
                                          self.dtl, self.sstep_max, self.ctx.stream()), "nh_dyn_step")

    def dynamics_step_record(self, stage_in):
        """``dynamics_step`` that also writes the input of each RK stage to ``stage_in`` (18 fields)."""
        ctx = self.ctx
        ctx._check_tensor(stage_in)
        if stage_in.shape[0] != 18:
            raise ValueError("stage_in needs 3 x 6 fields")
        native.check(ctx.lib, ctx.fn("nh_dyn_step_record")(
            ctx.handle, C.c_void_p(self.prog.data_ptr()), C.c_void_p(self.progq.data_ptr()), self.progq.shape[0],
            self.dtl, self.sstep_max, C.c_void_p(stage_in.data_ptr()), ctx.stream()), "nh_dyn_step_record")

    def dynamics_step_adj_rec(self, stage_in, progq, prog_b):
        """``dynamics_step_adj`` from the stage inputs ``dynamics_step_record`` wrote."""
        ctx = self.ctx
        for a in (stage_in, progq, prog_b):
            ctx._check_tensor(a)
        self._check_state_pair(progq, prog_b)
        if stage_in.shape[0] != 18:
            raise ValueError("stage_in needs 3 x 6 fields")
        native.check(ctx.lib, ctx.fn("nh_dyn_step_adj_rec")(
            ctx.handle, C.c_void_p(stage_in.data_ptr()), C.c_void_p(progq.data_ptr()), progq.shape[0], self.dtl,
            self.sstep_max, C.c_void_p(prog_b.data_ptr()), ctx.stream()), "nh_dyn_step_adj_rec")

    def tape_bytes(self) -> int:
        """Bytes of one step's tape (``dynamics_step_tape``)."""
        n = int(self.ctx.lib.nh_tape_bytes(self.ctx.handle, self.sstep_max))
        if n < 0:
            native.check(self.ctx.lib, n, "nh_tape_bytes")
        return n

    def new_tape(self):
        """A device buffer for one step's tape."""
        return _torch().empty(self.tape_bytes(), dtype=_torch().uint8, device=self.ctx.device)

    def dynamics_step_tape(self, tape):
        """``dynamics_step`` writing its intermediate fields to ``tape`` (for ``dynamics_step_adj_tape``)."""
        ctx = self.ctx
        if tape.dtype != _torch().uint8 or tape.numel() < self.tape_bytes() or tape.device != ctx.device \
                or not tape.is_contiguous():
            raise ValueError(f"tape must be a contiguous uint8 tensor of at least {self.tape_bytes()} bytes "
                             f"on {ctx.device}")
        native.check(ctx.lib, ctx.fn("nh_dyn_step_tape")(
            ctx.handle, C.c_void_p(self.prog.data_ptr()), C.c_void_p(self.progq.data_ptr()), self.progq.shape[0],
            self.dtl, self.sstep_max, C.c_void_p(tape.data_ptr()), ctx.stream()), "nh_dyn_step_tape")

    def dynamics_step_adj_tape(self, tape, progq, prog_b):
        """``dynamics_step_adj`` from the tape of the step (``progq``: the step's input tracers)."""
        ctx = self.ctx
        for a in (progq, prog_b):
            ctx._check_tensor(a)
        self._check_state_pair(progq, prog_b)
        if tape.dtype != _torch().uint8 or tape.numel() < self.tape_bytes() or tape.device != ctx.device \
                or not tape.is_contiguous():
            raise ValueError(f"tape must be a contiguous uint8 tensor of at least {self.tape_bytes()} bytes "
                             f"on {ctx.device}")
        native.check(ctx.lib, ctx.fn("nh_dyn_step_adj_tape")(
            ctx.handle, C.c_void_p(tape.data_ptr()), C.c_void_p(progq.data_ptr()), progq.shape[0], self.dtl,
            self.sstep_max, C.c_void_p(prog_b.data_ptr()), ctx.stream()), "nh_dyn_step_adj_tape")

    def dynamics_step_adj(self, prog, progq, prog_b):
        """Transpose of one ``dynamics_step`` started from (prog, progq), applied in place to ``prog_b``.

        On entry ``prog_b`` is the gradient of a scalar with respect to the state
        after the step; on exit it is the gradient with respect to ``prog``.  The
        tracer masses are held at ``progq`` (exact for zero tracers, the dry runs).
        """
        ctx = self.ctx
        for a in (prog, progq, prog_b):
            ctx._check_tensor(a)
        self._check_state_pair(progq, prog_b, prog)
        native.check(ctx.lib, ctx.fn("nh_dyn_step_adj")(
            ctx.handle, C.c_void_p(prog.data_ptr()), C.c_void_p(progq.data_ptr()), progq.shape[0], self.dtl,
            self.sstep_max, C.c_void_p(prog_b.data_ptr()), ctx.stream()), "nh_dyn_step_adj")

    def sfc_pressure(self, prog=None, progq=None):
        """Surface pressure per column (BNDCND_pre_sfc, mod_bndcnd.py:521), FP64, shape (nreg + 12,)."""
        ctx = self.ctx
        prog = self.prog if prog is None else prog
        progq = self.progq if progq is None else progq
        for a in (prog, progq):
            ctx._check_tensor(a)
        self._check_state_pair(progq, prog)
        ps = _torch().empty(ctx.ncol, dtype=_torch().float64, device=ctx.device)
        native.check(ctx.lib, ctx.fn("nh_sfc_pressure")(ctx.handle, C.c_void_p(prog.data_ptr()),
                                                        C.c_void_p(progq.data_ptr()), C.c_void_p(ps.data_ptr()),
                                                        ctx.stream()), "nh_sfc_pressure")
        return ps

    def sfc_pressure_adj(self, prog, progq, ps_b, prog_b):
        """Add the transpose of ``sfc_pressure`` at (prog, progq) applied to ``ps_b`` into ``prog_b``."""
        ctx = self.ctx
        for a in (prog, progq, prog_b):
            ctx._check_tensor(a)
        self._check_state_pair(progq, prog_b, prog)
        if ps_b.numel() != ctx.ncol or ps_b.device != ctx.device:
            raise ValueError(f"ps_b needs {ctx.ncol} values (one per column) on {ctx.device}")
        ps_b = ps_b.to(dtype=_torch().float64).contiguous()
        native.check(ctx.lib, ctx.fn("nh_sfc_pressure_adj")(
            ctx.handle, C.c_void_p(prog.data_ptr()), C.c_void_p(progq.data_ptr()), C.c_void_p(ps_b.data_ptr()),
            C.c_void_p(prog_b.data_ptr()), ctx.stream()), "nh_sfc_pressure_adj")

    def forcing_step(self, af_type: str = "HELD-SUAREZ", moist_case: bool = False, dt: float | None = None):
        """The forcing after a dynamics step (``Dyn.forcing_step``, mod_dynamics.py:597-720), in place:
        Held-Suarez Rayleigh friction and temperature relaxation (AF_heldsuarez,
        mod_af_heldsuarez.py:43-95) over ``dt`` (default the large step), then the halo fill.
        The regional points are forced, the pole points are not, as in the reference."""
        if af_type != "HELD-SUAREZ":
            raise NotImplementedError(f"backend='hip' implements AF_TYPE 'HELD-SUAREZ', not {af_type!r}")
        ctx = self.ctx
        dt = self.dtl if dt is None else float(dt)
        native.check(ctx.lib, ctx.fn("nh_forcing_hs")(
            ctx.handle, C.c_void_p(self.prog.data_ptr()), C.c_void_p(self.progq.data_ptr()), dt,
            1 if moist_case else 0, ctx.stream()), "nh_forcing_hs")

    def budget(self, lhv: float | None = None) -> dict:
        """Global energies and masses of the current state (Embudget._store_var, mod_embudget.py:82-139).

        Keys as the reference's monitor: pot, int, kin, tot [J], dry, vap [kg].
        ``lhv`` defaults to the constant set's latent heat (CONST_LHV).
        """
        lib, c = self.ctx.lib, self.ctx.const
        out = (C.c_double * 6)()
        lhv = float(c.LHV if lhv is None else lhv)
        native.check(lib, self.ctx.fn("nh_budget")(self.ctx.handle, C.c_void_p(self.prog.data_ptr()),
                                        C.c_void_p(self.progq.data_ptr()), self.progq.shape[0], lhv, out,
                                        self.ctx.stream()), "nh_budget")
        return dict(zip(("pot", "int", "kin", "tot", "dry", "vap"), list(out)))

    def state(self) -> dict:
        """Copy the state back to the host in the reference layout."""
        self.ctx.sync()
        P, Pp = self.ctx.to_host(self.prog)
        Q, Qp = self.ctx.to_host(self.progq)
        return {"PROG": P, "PROG_pl": Pp, "PROGq": Q, "PROGq_pl": Qp}


WIND_KINDS = {"dcmip11": 0, "deform": 1}


class HipOfflineTracer(HipDynamics):
    """Offline tracer run: ``TIME_integ_type = 'TRCADV'`` of ``Dyn.dynamics_step``.

    Each step transports the tracers with the mass fluxes held in PROG
    (mod_dynamics.py:2030-2068) and then writes the prescribed winds of the
    new time into PROG (``Forcing.forcing_update``, mod_forcing.py:113-206).
    ``wind`` is ``"dcmip11"`` (Trcadv.test11_velocity, mod_af_trcadv.py:23-98)
    or ``"deform"`` with ``params = (period, umax, w0, ztop)``: the
    non-divergent deformational flow (Nair and Lauritzen 2010, case 1)
    u = umax sin^2(lon/2) sin(2 lat) cos(pi t/T), v = umax/2 sin(lon) cos(lat) cos(pi t/T),
    w = w0 sin(pi z/ztop) cos(2 pi t/T).
    The winds of ``time0`` are written at construction.
    """

    def __init__(self, ctx: HipContext, state: dict, wind: str = "dcmip11", params=None, time0: float = 0.0):
        if wind not in WIND_KINDS:
            raise ValueError(f"unknown wind {wind!r}; expected one of {sorted(WIND_KINDS)}")
        p = [0.0] * 4 if params is None else [float(v) for v in params]
        if len(p) != 4:
            raise ValueError("params = (period, umax, w0, ztop)")
        super().__init__(ctx, state)
        self.kind = WIND_KINDS[wind]
        self._params = (C.c_double * 4)(*p)
        self.time = float(time0)
        self.set_winds(self.time)

    def set_winds(self, time: float):
        lib = self.ctx.lib
        native.check(lib, self.ctx.fn("nh_trcadv_winds")(self.ctx.handle, C.c_void_p(self.prog.data_ptr()), float(time),
                                              self.kind, self._params, self.ctx.stream()), "nh_trcadv_winds")

    @staticmethod
    def _check_state_pair(progq, *progs):
        """PROG-shaped arguments hold the 6 prognostics; PROGq at least qv."""
        for p in progs:
            if p.dim() != 2 or p.shape[0] != 6:
                raise ValueError("a PROG argument needs 6 fields")
        if progq.dim() != 2 or progq.shape[0] < 1:
            raise ValueError("PROGq needs at least one tracer field (qv)")

    def dynamics_step(self):
        lib = self.ctx.lib
        native.check(lib, self.ctx.fn("nh_tracer_step")(self.ctx.handle, C.c_void_p(self.prog.data_ptr()),
                                             C.c_void_p(self.progq.data_ptr()), self.progq.shape[0], self.dtl,
                                             self.ctx.stream()), "nh_tracer_step")
        self.time += self.dtl
        self.set_winds(self.time)

    def tracer_step_adj(self, prog, progq, progq_b):
        """Transpose of one tracer step with respect to the tracers, in place on ``progq_b``.

        ``prog``: the winds (mass fluxes) the step ran with, ``progq``: the tracers
        it started from.  On entry ``progq_b`` is the gradient of a scalar with
        respect to the tracers after the step, on exit with respect to ``progq``.
        The fluxes are held fixed and the limiters' selections are those of the
        forward pass (the offline-tracer sensitivity).
        """
        ctx = self.ctx
        for a in (prog, progq, progq_b):
            ctx._check_tensor(a)
        if progq_b.shape != progq.shape:
            raise ValueError("progq_b must have the shape of progq")
        native.check(ctx.lib, ctx.fn("nh_tracer_step_adj")(
            ctx.handle, C.c_void_p(prog.data_ptr()), C.c_void_p(progq.data_ptr()), progq.shape[0], self.dtl,
            C.c_void_p(progq_b.data_ptr()), ctx.stream()), "nh_tracer_step_adj")
