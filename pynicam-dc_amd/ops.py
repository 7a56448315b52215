"""The reference's operator API on the HIP backend.

``HipOperators`` runs the single operators of the step on device tensors in
the device layout ((nv, NF) blocks, ``HipContext.empty``): the tridiagonal
matrix and column solve of the vertical implicit step, the diagnostics /
boundary conditions, the hyperdiffusion and divergence-damping tendencies and
the two Thuburn limiters.

Beside it, functions with the names and argument lists of the reference's
pure kernels (``nhm/dynamics/kernels/*.py``) take and return numpy arrays in
the reference layout, so that a caller written against those kernels can
switch to this backend: ``compute_rhow_matrix_reg/_pl`` (vimatrix.py:111,139),
``compute_rhow_solver_reg/_pl`` (virhowsolver.py:67,123),
``compute_vertical_limiter`` (verticallimiter.py:41).  Their metric arguments
are accepted for the signature and checked to be the flat-topography ones the
context was built with (every metric factor 1).  Operators whose reference
form spans a halo exchange (``numfilter_hdiffusion``, ``numfilter_divdamp``,
the two-stage horizontal limiter) take the regional and pole arrays together:
``hdiffusion``, ``divdamp``, ``horizontal_limiter``, ``prepost``.

Every call is one or a few library launches; host arrays are converted to the
device layout and back at the boundary (``layout.pack`` / ``unpack``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import layout, native


def _p(t):
    return C.c_void_p(t.data_ptr())


class HipOperators:
    """Operators of one ``HipContext`` on device-layout tensors."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.NF = ctx.NF

    # -- helpers -------------------------------------------------------------
    def _field(self, a, nv: int | None = None, dtype=None):
        torch = __import__("torch")
        ctx = self.ctx
        want = dtype or ctx.dtype
        if not isinstance(a, torch.Tensor) or a.dtype != want or a.device != ctx.device or not a.is_contiguous():
            raise TypeError(f"expected a contiguous {want} tensor on {ctx.device}")
        n = a.numel()
        if nv is not None and n != nv * self.NF:
            raise ValueError(f"expected {nv} field(s) of {self.NF} values, got {n} values")
        if nv is None and n % self.NF:
            raise ValueError(f"expected whole fields of {self.NF} values, got {n} values")
        return a

    def _call(self, name, *args):
        ctx = self.ctx
        native.check(ctx.lib, ctx.fn(name)(ctx.handle, *args, ctx.stream()), name)

    def _new(self, nv: int, dtype=None):
        torch = __import__("torch")
        return torch.zeros((nv, self.NF), dtype=dtype or self.ctx.dtype, device=self.ctx.device)

    # -- operators ------------------------------------------------------------
    def vi_matrix(self, eth_h, g_tilde, dt: float):
        """Mc, Mu, Ml (3 FP64 fields; rows kmin+1..kmax) of the rhogw solve
        (compute_rhow_matrix_reg/_pl, kernels/vimatrix.py:111-165)."""
        self._field(eth_h, 1)
        self._field(g_tilde, 1)
        torch = __import__("torch")
        M = self._new(3, torch.float64)
        self._call("nh_vi_matrix", _p(eth_h), _p(g_tilde), float(dt), _p(M[0]), _p(M[1]), _p(M[2]))
        return M

    def vi_solve(self, rhogw, rhogw0, preg0, rhog0, Srho, Sw, Spre, M, dt: float, out=None):
        """New rhogw (rows kmin..kmax+1 solved or set, the rest copied) from the
        split-state sources (compute_rhow_solver_reg/_pl, kernels/virhowsolver.py:67-172)."""
        torch = __import__("torch")
        for a in (rhogw, rhogw0, preg0, rhog0, Srho, Sw, Spre):
            self._field(a, 1)
        self._field(M, 3, torch.float64)
        out = self._new(1)[0] if out is None else self._field(out, 1)
        self._call("nh_vi_solve", _p(rhogw), _p(rhogw0), _p(preg0), _p(rhog0), _p(Srho), _p(Sw), _p(Spre),
                   _p(M[0]), _p(M[1]), _p(M[2]), float(dt), _p(out))
        return out

    def prepost(self, prog, progq):
        """(diag (6 fields: pre, tem, vx, vy, vz, w), eth); the ghost levels of prog
        are set in place (compute_diagnostics, BNDCND_all, THRMDYN_eth;
        mod_dynamics.py:965-1283)."""
        self._field(prog, 6)
        self._field(progq)
        diag, eth = self._new(6), self._new(1)
        self._call("nh_prepost", _p(prog), _p(progq), _p(diag), _p(eth))
        return diag, eth[0]

    def hdiffusion(self, prog, diag):
        """Hyperdiffusion tendencies, 6 fields in PROG order
        (Numf.numfilter_hdiffusion, mod_numfilter.py:845-1578)."""
        self._field(prog, 6)
        self._field(diag, 6)
        tend = self._new(6)
        self._call("nh_hdiffusion", _p(prog), _p(diag), _p(tend))
        return tend

    def divdamp(self, rhogvx, rhogvy, rhogvz, rhogw):
        """gdx, gdy, gdz (3 fields) of the 3-D divergence damping
        (Numf.numfilter_divdamp, mod_numfilter.py:2032-2268)."""
        for a in (rhogvx, rhogvy, rhogvz, rhogw):
            self._field(a, 1)
        gd = self._new(3)
        self._call("nh_divdamp3d", _p(rhogvx), _p(rhogvy), _p(rhogvz), _p(rhogw), _p(gd[0]), _p(gd[1]), _p(gd[2]))
        return gd

    def vertical_limiter(self, q_h, q, d, ck):
        """Limit q_h in place (compute_vertical_limiter, kernels/verticallimiter.py:41-92);
        ck holds the two Courant-number fields."""
        self._field(q_h, 1)
        self._field(q, 1)
        self._field(d, 1)
        self._field(ck, 2)
        self._call("nh_limiter_v", _p(q_h), _p(q), _p(d), _p(ck))
        return q_h

    def horizontal_limiter(self, q_a, q, d, ch, cmask):
        """Limit the six face fields q_a in place (compute_horizontal_limiter_qout /
        _apply with the Qout exchange, kernels/horizontallimiter.py:51-255)."""
        self._field(q_a, 6)
        self._field(q, 1)
        self._field(d, 1)
        self._field(ch, 6)
        self._field(cmask, 6)
        self._call("nh_limiter_h", _p(q_a), _p(q), _p(d), _p(ch), _p(cmask))
        return q_a


# ----------------------------------------------------------------------------
# reference-signature wrappers (numpy in, numpy out)
# ----------------------------------------------------------------------------
def _ops(ctx) -> HipOperators:
    ops = getattr(ctx, "_operators", None)
    if ops is None:
        ops = HipOperators(ctx)
        ctx._operators = ops
    return ops


def _dev(ctx, reg=None, pl=None, dtype=None):
    """One field from a regional (g, g, kall, l) and/or pole (6, kall, 2) array (the other part
    zero), in the context's field precision or ``dtype``."""
    torch = __import__("torch")
    t = ctx.topo
    if reg is None:
        reg = np.zeros((t.gall_1d, t.gall_1d, t.kall, t.lall))
    if pl is None:
        pl = np.zeros((6, t.kall, 2))
    flat = layout.pack(np.asarray(reg, dtype=np.float64)[..., None], np.asarray(pl, dtype=np.float64)[..., None])
    return torch.from_numpy(flat[0]).to(device=ctx.device, dtype=dtype or ctx.dtype)


def _host(ctx, flat, pole: bool):
    t = ctx.topo
    reg, pl = layout.unpack1(flat.detach().cpu().numpy().astype(np.float64), t.gall_1d, t.kall, t.lall)
    return pl if pole else reg


def _require_flat(ctx, **metrics):
    for name, a in metrics.items():
        if a is not None and not np.all(np.asarray(a) == 1.0):
            raise ValueError(f"{name}: the HIP backend's operators take flat-topography metrics (all 1)")


def _check_vgrid(ctx, **prof):
    vg = ctx.geo.vgrid
    for name, a in prof.items():
        if a is not None and not np.array_equal(np.asarray(a, dtype=np.float64), np.asarray(getattr(vg, name))):
            raise ValueError(f"{name} differs from the vertical grid the context was built with")


def _check_coef(ctx, coef, cfg, pole):
    """The pre-combined metric products (precombine_matrix_reg/_pl, kernels/vimatrix.py:67-108)
    must be those of flat topography on the context's vertical grid."""
    if coef is None:
        return
    vg, c = ctx.geo.vgrid, ctx.const
    ks, ksm = slice(cfg.kmin + 1, cfg.kmax + 1), slice(cfg.kmin, cfg.kmax)
    GCVovR = c.GRAV * c.CVdry / c.Rdry
    want = {"Ca": vg.rdgz[ks] + vg.rdgz[ksm], "Cb": vg.dfact[ks] - vg.cfact[ksm], "Cu1": vg.rdgz[ks],
            "Cl1": vg.rdgz[ks], "Cu2": np.full(cfg.kmax - cfg.kmin, GCVovR), "Cl2": np.full(cfg.kmax - cfg.kmin, GCVovR)}
    for name, w in want.items():
        a = np.asarray(coef[name])
        shape = (1, -1, 1) if pole else (1, 1, -1, 1)
        if not np.array_equal(a, np.broadcast_to(w.reshape(shape), a.shape)):
            raise ValueError(f"coef[{name!r}] is not the flat-topography product of the context's vertical grid")


def _matrix(ctx, eth, g_tilde, RGSQRTH, rdgzh, dfact, cfact, coef, dt, cfg, pole):
    _require_flat(ctx, RGSQRTH=RGSQRTH)
    _check_vgrid(ctx, rdgzh=rdgzh, dfact=dfact, cfact=cfact)
    _check_coef(ctx, coef, cfg, pole)
    e = _dev(ctx, pl=eth) if pole else _dev(ctx, reg=eth)
    gt = _dev(ctx, pl=g_tilde) if pole else _dev(ctx, reg=g_tilde)
    M = _ops(ctx).vi_matrix(e, gt, dt)
    ks = slice(cfg.kmin + 1, cfg.kmax + 1)
    out = []
    for q in range(3):
        a = _host(ctx, M[q], pole)
        out.append(np.ascontiguousarray(a[:, ks] if pole else a[:, :, ks]))
    return tuple(out)


def compute_rhow_matrix_reg(eth, g_tilde, RGSQRTH, rdgzh, dfact, cfact, coef, dt, cfg, xp=None, *, ctx):
    """Regional Mc, Mu, Ml interior slabs (i, j, kmax-kmin, l), FP64
    (kernels/vimatrix.py:111 compute_rhow_matrix_reg)."""
    return _matrix(ctx, eth, g_tilde, RGSQRTH, rdgzh, dfact, cfact, coef, dt, cfg, False)


def compute_rhow_matrix_pl(eth_pl, g_tilde_pl, RGSQRTH_pl, rdgzh, dfact, cfact, coef, dt, cfg, xp=None, *, ctx):
    """Pole Mc, Mu, Ml interior slabs (g, kmax-kmin, l), FP64 (kernels/vimatrix.py:139)."""
    return _matrix(ctx, eth_pl, g_tilde_pl, RGSQRTH_pl, rdgzh, dfact, cfact, coef, dt, cfg, True)


def _full(ctx, slab, cfg, pole):
    """An interior-row matrix slab (kmax - kmin rows) in a full-height array."""
    t = ctx.topo
    slab = np.asarray(slab, dtype=np.float64)
    if slab.shape[1 if pole else 2] == t.kall:
        return slab
    ks = slice(cfg.kmin + 1, cfg.kmax + 1)
    if pole:
        out = np.zeros((6, t.kall, 2))
        out[:, ks] = slab
    else:
        out = np.zeros((t.gall_1d, t.gall_1d, t.kall, t.lall))
        out[:, :, ks] = slab
    return out


def _solve(ctx, fields, M, metrics, rdgzh, afact, bfact, dt, cfg, pole):
    torch = __import__("torch")
    _require_flat(ctx, **metrics)
    _check_vgrid(ctx, rdgzh=rdgzh, afact=afact, bfact=bfact)
    dev = [(_dev(ctx, pl=f) if pole else _dev(ctx, reg=f)) for f in fields]
    Md = torch.stack([(_dev(ctx, pl=_full(ctx, m, cfg, True), dtype=torch.float64) if pole else
                       _dev(ctx, reg=_full(ctx, m, cfg, False), dtype=torch.float64)) for m in M]).contiguous()
    out = _ops(ctx).vi_solve(*dev, Md, dt)
    return _host(ctx, out, pole)


def compute_rhow_solver_reg(rhogw, rhogw0, preg0, rhog0, Srho, Sw, Spre, Mc, Mu, Ml, RGAMH, RGSGAM2, RGAM,
                            RGSGAM2H, GSGAM2H, rdgzh, afact, bfact, dt, cfg, xp=None, *, ctx):
    """New regional rhogw (i, j, kall, l) (kernels/virhowsolver.py:67 compute_rhow_solver_reg).
    Mc, Mu, Ml may be full-height arrays or the interior slabs compute_rhow_matrix_reg returns."""
    return _solve(ctx, (rhogw, rhogw0, preg0, rhog0, Srho, Sw, Spre), (Mc, Mu, Ml),
                  dict(RGAMH=RGAMH, RGSGAM2=RGSGAM2, RGAM=RGAM, RGSGAM2H=RGSGAM2H, GSGAM2H=GSGAM2H),
                  rdgzh, afact, bfact, dt, cfg, False)


def compute_rhow_solver_pl(rhogw_pl, rhogw0_pl, preg0_pl, rhog0_pl, Srho_pl, Sw_pl, Spre_pl, Mc_pl, Mu_pl, Ml_pl,
                           RGAMH_pl, RGSGAM2_pl, RGAM_pl, RGSGAM2H_pl, GSGAM2H_pl, rdgzh, afact, bfact, dt, cfg,
                           xp=None, *, ctx):
    """New pole rhogw (g, kall, l) (kernels/virhowsolver.py:123 compute_rhow_solver_pl)."""
    return _solve(ctx, (rhogw_pl, rhogw0_pl, preg0_pl, rhog0_pl, Srho_pl, Sw_pl, Spre_pl), (Mc_pl, Mu_pl, Ml_pl),
                  dict(RGAMH=RGAMH_pl, RGSGAM2=RGSGAM2_pl, RGAM=RGAM_pl, RGSGAM2H=RGSGAM2H_pl, GSGAM2H=GSGAM2H_pl),
                  rdgzh, afact, bfact, dt, cfg, True)


def compute_vertical_limiter(q_h, q, d, ck, cfg, xp=None, *, ctx):
    """Limited face values q_h (kernels/verticallimiter.py:41 compute_vertical_limiter) of a
    regional (i, j, kall, l) or pole (6, kall, 2) set of arrays; ck (..., 2)."""
    torch = __import__("torch")
    pole = np.ndim(q) == 3
    put = (lambda a: _dev(ctx, pl=a)) if pole else (lambda a: _dev(ctx, reg=a))
    qh = put(q_h)
    ckd = torch.stack([put(np.asarray(ck)[..., 0]), put(np.asarray(ck)[..., 1])]).contiguous()
    _ops(ctx).vertical_limiter(qh, put(q), put(d), ckd)
    return _host(ctx, qh, pole)


# ----------------------------------------------------------------------------
# operators whose reference form spans a halo exchange: regional and pole together
# ----------------------------------------------------------------------------
def prepost(ctx, PROG, PROG_pl, PROGq, PROGq_pl):
    """Diagnostics, boundary conditions and eth (mod_dynamics.py:965-1283):
    returns (PROG, PROG_pl with the ghost levels set, DIAG, DIAG_pl, eth, eth_pl)."""
    prog = ctx.to_device(PROG, PROG_pl)
    progq = ctx.to_device(PROGq, PROGq_pl)
    diag, eth = _ops(ctx).prepost(prog, progq)
    P, Pp = ctx.to_host(prog)
    D, Dp = ctx.to_host(diag)
    e, ep = ctx.to_host(eth[None])
    return P, Pp, D, Dp, e[..., 0], ep[..., 0]


def hdiffusion(ctx, PROG, PROG_pl, DIAG, DIAG_pl):
    """Hyperdiffusion tendency (…, 6) in PROG order, regional and pole
    (Numf.numfilter_hdiffusion, mod_numfilter.py:845-1578); PROG and DIAG as
    after the diagnostics and boundary conditions."""
    tend = _ops(ctx).hdiffusion(ctx.to_device(PROG, PROG_pl), ctx.to_device(DIAG, DIAG_pl))
    return ctx.to_host(tend)


def divdamp(ctx, rhogvx, rhogvx_pl, rhogvy, rhogvy_pl, rhogvz, rhogvz_pl, rhogw, rhogw_pl):
    """gdx, gdy, gdz (and their pole arrays) of the 3-D divergence damping
    (Numf.numfilter_divdamp, mod_numfilter.py:2032-2268)."""
    f = [_dev(ctx, reg=a, pl=b) for a, b in ((rhogvx, rhogvx_pl), (rhogvy, rhogvy_pl), (rhogvz, rhogvz_pl),
                                             (rhogw, rhogw_pl))]
    gd = _ops(ctx).divdamp(*f)
    return ctx.to_host(gd)


def horizontal_limiter(ctx, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl):
    """Limited face values (…, 6) and pole (6, kall, 2) (horizontal_limiter_thuburn,
    mod_src_tracer.py:2115-2662; kernels/horizontallimiter.py:51-255)."""
    torch = __import__("torch")

    def six(a, a_pl):
        a_pl = np.asarray(a_pl)
        full_pl = np.zeros(a_pl.shape[:3] + (6,))
        full_pl[..., 0] = a_pl
        return ctx.to_device(np.asarray(a, dtype=np.float64), full_pl)

    qa = six(q_a, q_a_pl)
    _ops(ctx).horizontal_limiter(qa, _dev(ctx, reg=q, pl=q_pl), _dev(ctx, reg=d, pl=d_pl), six(ch, ch_pl),
                                 six(cmask, cmask_pl))
    out, out_pl = ctx.to_host(qa)
    return out, out_pl[..., 0]
