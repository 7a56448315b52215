"""CPU oracle: numpy FP64 restatement of the pyNICAM-DC dynamical-core step.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP backend:
tests, the smoke check and the CPU-baseline leg of the benchmark may import it,
the product path never does.

Each function restates one reference routine (file:line cited in its
docstring), in the reference's array layout: regional fields (i, j, k, l),
pole fields (v, k, lpl).  Host setup (grid, operator coefficients, vertical
metrics) is taken from the product's setup modules, which are checked
separately against closed forms.
"""
from __future__ import annotations

import numpy as np

I_RHOG, I_RHOGVX, I_RHOGVY, I_RHOGVZ, I_RHOGW, I_RHOGE = range(6)
I_pre, I_tem, I_vx, I_vy, I_vz, I_w = range(6)
XDIR, YDIR, ZDIR = 0, 1, 2
TI, TJ = 0, 1


class OracleContext:
    """Geometry, constants and run parameters shared by the oracle routines."""

    def __init__(self, geo, const, run):
        self.geo = geo
        self.topo = geo.topo
        self.c = const
        self.run = run
        t = self.topo
        self.kmin, self.kmax, self.kall = t.kmin, t.kmax, t.kall
        self.g = t.gall_1d
        vg = geo.vgrid
        self.afact, self.bfact, self.cfact, self.dfact = vg.afact, vg.bfact, vg.cfact, vg.dfact
        self.rdgz, self.rdgzh = vg.rdgz, vg.rdgzh
        m = geo.vmtr
        self.m = m
        self.rscale = geo.radius
        # pntmask (share/mod_ppmask.py:20-31): (m0, m1) = (0, 1) on pentagon regions
        self.pnt_m0 = np.where(t.have_sgp, 0.0, 1.0)
        self.pnt_m1 = np.where(t.have_sgp, 1.0, 0.0)

    def bs(self, name, reg):
        """Basic-state array (Bsst.pre_bs / tem_bs / rho_bs, mod_bsstate.py:159-200)
        of the regional or pole layout; 0.0 for ref_type NOBASE."""
        b = getattr(self.geo, "bsstate", None)
        if b is None:
            return 0.0
        full = b.full(name)
        return full[0] if reg else full[1]

    # ------------------------------------------------------------------
    def halo(self, var, var_pl):
        """COMM_data_transfer (share/mod_comm.py:1563-1644) on one device."""
        self.topo.halo_fill_host(var, var_pl)


# ======================================================================
# horizontal operators (kernels/*.py of the reference)
# ======================================================================
def _st7(a):
    i, j = a.shape[0], a.shape[1]
    return [a[1:i - 1, 1:j - 1], a[2:i, 1:j - 1], a[2:i, 2:j], a[1:i - 1, 2:j],
            a[0:i - 2, 1:j - 1], a[0:i - 2, 0:j - 2], a[1:i - 1, 0:j - 2]]


def laplacian(ctx, scl, scl_pl):
    """7-point Laplacian (kernels/oprtlaplacian.py:41-85)."""
    g = ctx.g
    c = ctx.geo.coef_lap[1:g - 1, 1:g - 1][:, :, None]      # (.., 1, l, 7)
    s = _st7(scl)
    inner = (c[..., 0] * s[0] + c[..., 1] * s[1] + c[..., 2] * s[2] + c[..., 3] * s[3]
             + c[..., 4] * s[4] + c[..., 5] * s[5] + c[..., 6] * s[6])
    out = np.zeros_like(scl)
    out[1:g - 1, 1:g - 1] = inner
    out_pl = np.zeros_like(scl_pl)
    cp = ctx.geo.coef_lap_pl                                  # (6, 2)
    out_pl[0] = np.sum(cp[:, None, :] * scl_pl, axis=0)
    return out, out_pl


def divergence(ctx, vx, vy, vz, vx_pl, vy_pl, vz_pl):
    """7-point divergence (kernels/fluxconv.py:52-108)."""
    g = ctx.g
    c = ctx.geo.coef_div[1:g - 1, 1:g - 1][:, :, None]       # (.., 1, l, 3, 7)

    def st(v, d):
        s = _st7(v)
        return (c[..., d, 0] * s[0] + c[..., d, 1] * s[1] + c[..., d, 2] * s[2]
                + c[..., d, 3] * s[3] + c[..., d, 4] * s[4] + c[..., d, 5] * s[5]
                + c[..., d, 6] * s[6])

    out = np.zeros_like(vx)
    out[1:g - 1, 1:g - 1] = st(vx, XDIR) + st(vy, YDIR) + st(vz, ZDIR)
    cp = ctx.geo.coef_div_pl                                  # (6, 2, 3)
    per_v = (cp[:, None, :, XDIR] * vx_pl + cp[:, None, :, YDIR] * vy_pl
             + cp[:, None, :, ZDIR] * vz_pl)
    out_pl = np.zeros_like(vx_pl)
    out_pl[0] = per_v.sum(axis=0)
    return out, out_pl


def gradient(ctx, scl, scl_pl):
    """7-point gradient (kernels/oprtgradient.py:45-88), interior + pole centre."""
    g = ctx.g
    c = ctx.geo.coef_grad[1:g - 1, 1:g - 1][:, :, None]
    s = np.stack(_st7(scl), axis=-1)
    out = np.zeros(scl.shape + (3,))
    for d in range(3):
        out[1:g - 1, 1:g - 1, ..., d] = np.sum(c[..., d, :] * s, axis=-1)
    cp = ctx.geo.coef_grad_pl                                  # (6, 2, 3)
    out_pl = np.zeros(scl_pl.shape + (3,))
    out_pl[0] = np.sum(cp[:, None, :, :] * scl_pl[..., None], axis=0)
    return out, out_pl


def horizontalize_interior(ctx, vx, vy, vz):
    """Remove the radial part on interior points (kernels/horizontalizevec.py:34-55)."""
    g, R = ctx.g, ctx.rscale
    s = slice(1, g - 1)
    x = ctx.geo.GRD_x[s, s, 0]
    gx, gy, gz = x[:, :, None, :, XDIR], x[:, :, None, :, YDIR], x[:, :, None, :, ZDIR]
    prd = (vx[s, s] * gx + vy[s, s] * gy + vz[s, s] * gz) / R
    return vx[s, s] - prd * gx / R, vy[s, s] - prd * gy / R, vz[s, s] - prd * gz / R


def horizontalize_pole(ctx, vx, vy, vz):
    """Pole version (kernels/horizontalizevec.py:58-71)."""
    R = ctx.rscale
    xp = ctx.geo.GRD_x_pl[:, 0]
    gx, gy, gz = xp[:, None, :, XDIR], xp[:, None, :, YDIR], xp[:, None, :, ZDIR]
    prd = vx * gx / R + vy * gy / R + vz * gz / R
    return vx - prd * gx / R, vy - prd * gy / R, vz - prd * gz / R


def horizontalize_inplace(ctx, vx, vy, vz, vx_pl, vy_pl, vz_pl):
    """OPRT_horizontalize_vec (share/mod_oprt.py:1404-1427)."""
    g = ctx.g
    s = slice(1, g - 1)
    nx, ny, nz = horizontalize_interior(ctx, vx, vy, vz)
    vx[s, s], vy[s, s], vz[s, s] = nx, ny, nz
    px, py, pz = horizontalize_pole(ctx, vx_pl, vy_pl, vz_pl)
    vx_pl[...], vy_pl[...], vz_pl[...] = px, py, pz


def _vertex_values(ctx, sx, sy, sz, with_mask):
    """Triangle-mean gradient values for OPRT_divdamp-type operators.

    Returns sclt_TI, sclt_TJ on i, j = 0..gmax (kernels/oprtdivdamp.py:54-95).
    """
    gmax = ctx.topo.gmax
    G = gmax + 1
    R, Rp = slice(0, G), slice(1, G + 1)
    c = ctx.geo.coef_intp[R, R][:, :, None]                  # (G, G, 1, l, 3, 3, 2)
    ti = (c[..., XDIR, 0, TI] * sx[R, R] + c[..., XDIR, 1, TI] * sx[Rp, R] + c[..., XDIR, 2, TI] * sx[Rp, Rp]
          + c[..., YDIR, 0, TI] * sy[R, R] + c[..., YDIR, 1, TI] * sy[Rp, R] + c[..., YDIR, 2, TI] * sy[Rp, Rp]
          + c[..., ZDIR, 0, TI] * sz[R, R] + c[..., ZDIR, 1, TI] * sz[Rp, R] + c[..., ZDIR, 2, TI] * sz[Rp, Rp])
    tj = (c[..., XDIR, 0, TJ] * sx[R, R] + c[..., XDIR, 1, TJ] * sx[Rp, Rp] + c[..., XDIR, 2, TJ] * sx[R, Rp]
          + c[..., YDIR, 0, TJ] * sy[R, R] + c[..., YDIR, 1, TJ] * sy[Rp, Rp] + c[..., YDIR, 2, TJ] * sy[R, Rp]
          + c[..., ZDIR, 0, TJ] * sz[R, R] + c[..., ZDIR, 1, TJ] * sz[Rp, Rp] + c[..., ZDIR, 2, TJ] * sz[R, Rp])
    return ti, tj


def _hex_flux_sum(ctx, ti, tj, d):
    """Sum over the six hexagon arcs of coef_diff * (adjacent triangle values)."""
    gmax = ctx.topo.gmax
    G = gmax + 1
    II, IM = slice(1, G), slice(0, G - 1)
    cd = ctx.geo.coef_diff[II, II][:, :, None, :, d, :]      # (gmax, gmax, 1, l, 6)
    return (cd[..., 0] * (ti[II, II] + tj[II, II])
            + cd[..., 1] * (tj[II, II] + ti[IM, II])
            + cd[..., 2] * (ti[IM, II] + tj[IM, IM])
            + cd[..., 3] * (tj[IM, IM] + ti[IM, IM])
            + cd[..., 4] * (ti[IM, IM] + tj[II, IM])
            + cd[..., 5] * (tj[II, IM] + ti[II, II]))


def _pntmask_corner(ctx, ti, tj):
    """Pentagon fix: the (0,0) TI value is replaced by TJ(1,0) on pentagon regions."""
    m0 = ctx.pnt_m0
    m1 = ctx.pnt_m1
    ti[0, 0] = ti[0, 0] * m0 + tj[1, 0] * m1


def divdamp_op(ctx, vx, vy, vz, vx_pl, vy_pl, vz_pl):
    """OPRT_divdamp: gradient of the divergence (kernels/oprtdivdamp.py:54-173)."""
    g = ctx.g
    G = ctx.topo.gmax + 1
    ti, tj = _vertex_values(ctx, vx, vy, vz, False)
    outs = []
    for d in range(3):
        o = np.zeros_like(vx)
        o[1:G, 1:G] = _hex_flux_sum(ctx, ti, tj, d)
        outs.append(o)
    # pole
    cI = ctx.geo.coef_intp_pl[1:6]                            # (5, 2, 3, 3)
    cd = ctx.geo.coef_diff_pl[1:6]                            # (5, 2, 3)
    vxs, vys, vzs = vx_pl[0], vy_pl[0], vz_pl[0]
    vxi, vyi, vzi = vx_pl[1:6], vy_pl[1:6], vz_pl[1:6]
    vxp, vyp, vzp = np.roll(vxi, -1, 0), np.roll(vyi, -1, 0), np.roll(vzi, -1, 0)

    def b(a):
        return a[:, None, :]
    sclt = (b(cI[:, :, XDIR, 0]) * vxs + b(cI[:, :, XDIR, 1]) * vxi + b(cI[:, :, XDIR, 2]) * vxp
            + b(cI[:, :, YDIR, 0]) * vys + b(cI[:, :, YDIR, 1]) * vyi + b(cI[:, :, YDIR, 2]) * vyp
            + b(cI[:, :, ZDIR, 0]) * vzs + b(cI[:, :, ZDIR, 1]) * vzi + b(cI[:, :, ZDIR, 2]) * vzp)
    pair = np.roll(sclt, 1, 0) + sclt
    outs_pl = []
    for d in range(3):
        o = np.zeros_like(vx_pl)
        o[0] = (b(cd[:, :, d]) * pair).sum(axis=0)
        outs_pl.append(o)
    return outs, outs_pl


def diffusion(ctx, scl, scl_pl, kh, kh_pl):
    """Kh-weighted diffusion (kernels/oprtdiffusion.py:59-177)."""
    g = ctx.g
    gmin, gmax = ctx.topo.gmin, ctx.topo.gmax
    c = ctx.geo.coef_intp[0:g - 1, 0:g - 1][:, :, None]      # (g-1, g-1, 1, l, 3, 3, 2)
    s0 = scl[0:g - 1, 0:g - 1][..., None]
    sip = scl[1:g, 0:g - 1][..., None]
    sd = scl[1:g, 1:g][..., None]
    sjp = scl[0:g - 1, 1:g][..., None]
    c1i, c2i, c3i = c[..., 0, TI], c[..., 1, TI], c[..., 2, TI]
    c1j, c2j, c3j = c[..., 0, TJ], c[..., 1, TJ], c[..., 2, TJ]
    ti = ((2.0 * c1i - c2i - c3i) * s0 + (-c1i + 2.0 * c2i - c3i) * sip
          + (-c1i - c2i + 2.0 * c3i) * sd) / 3.0
    tj = ((2.0 * c1j - c2j - c3j) * s0 + (-c1j + 2.0 * c2j - c3j) * sd
          + (-c1j - c2j + 2.0 * c3j) * sjp) / 3.0
    m0 = ctx.pnt_m0[None, None, None, :, None]
    m1 = ctx.pnt_m1[None, None, None, :, None]
    ti[0:1, 0:1] = ti[0:1, 0:1] * m0 + tj[1:2, 0:1] * m1
    sl, slp, slm = slice(gmin, gmax + 1), slice(gmin + 1, gmax + 2), slice(gmin - 1, gmax)
    k0 = kh[sl, sl]
    kf = [0.5 * (k0 + kh[slp, slp]), 0.5 * (k0 + kh[sl, slp]), 0.5 * (k0 + kh[slm, sl]),
          0.5 * (k0 + kh[slm, slm]), 0.5 * (k0 + kh[sl, slm]), 0.5 * (k0 + kh[slp, sl])]
    II, IM = slice(gmin, gmax + 1), slice(gmin - 1, gmax)
    a_ti, a_tj = ti[II, II], tj[II, II]
    b_ti = ti[IM, II]
    c_tj, c_ti = tj[IM, IM], ti[IM, IM]
    d_tj = tj[II, IM]
    inner = None
    for d in range(3):
        cd = ctx.geo.coef_diff[II, II][:, :, None, :, d, :]
        blk = (kf[0] * cd[..., 0] * (a_ti[..., d] + a_tj[..., d])
               + kf[1] * cd[..., 1] * (a_tj[..., d] + b_ti[..., d])
               + kf[2] * cd[..., 2] * (b_ti[..., d] + c_tj[..., d])
               + kf[3] * cd[..., 3] * (c_tj[..., d] + c_ti[..., d])
               + kf[4] * cd[..., 4] * (c_ti[..., d] + d_tj[..., d])
               + kf[5] * cd[..., 5] * (d_tj[..., d] + a_ti[..., d]))
        inner = blk if inner is None else inner + blk
    out = np.zeros_like(scl)
    out[gmin:gmax + 1, gmin:gmax + 1] = inner
    # pole
    cI = ctx.geo.coef_intp_pl[1:6][:, None]                   # (5, 1, 2, 3, 3)
    c0, c1, c2 = cI[..., 0], cI[..., 1], cI[..., 2]
    ss = scl_pl[0][None, :, :, None]
    si = scl_pl[1:6][..., None]
    sp = np.roll(scl_pl[1:6], -1, 0)[..., None]
    vt = ((2.0 * c0 - c1 - c2) * ss + (-c0 + 2.0 * c1 - c2) * si + (-c0 - c1 + 2.0 * c2) * sp) / 3.0
    vsum = np.roll(vt, 1, 0) + vt
    kavg = 0.5 * (kh_pl[0][None] + kh_pl[1:6])
    cdp = ctx.geo.coef_diff_pl[1:6][:, None]                  # (5, 1, 2, 3)
    contrib = kavg * (cdp * vsum).sum(axis=-1)
    out_pl = np.zeros_like(scl_pl)
    out_pl[0] = contrib.sum(axis=0)
    return out, out_pl


# ======================================================================
# vertical / column pieces
# ======================================================================
def flux_convergence(ctx, rvx, rvy, rvz, rw, rvx_pl, rvy_pl, rvz_pl, rw_pl, vertical):
    """-(horizontal divergence + vertical flux difference) (kernels/fluxconv.py:111-208)."""
    m = ctx.m
    kmin, kmax = ctx.kmin, ctx.kmax
    out = []
    for reg, (vx, vy, vz, w) in ((True, (rvx, rvy, rvz, rw)), (False, (rvx_pl, rvy_pl, rvz_pl, rw_pl))):
        tag = "" if reg else "_pl"
        kax = 2 if reg else 1
        RGAM, RGAMH, RGSQRTH, GZ = m["RGAM" + tag], m["RGAMH" + tag], m["RGSQRTH" + tag], m["C2WfactGz" + tag]
        vflag = 1.0 if vertical else 0.0
        K = _ks(kax, kmin + 1, kmax + 1)
        Km = _ks(kax, kmin, kmax)
        horiz = (GZ[K + (..., 0)] * vx[K] + GZ[K + (..., 1)] * vx[Km] + GZ[K + (..., 2)] * vy[K]
                 + GZ[K + (..., 3)] * vy[Km] + GZ[K + (..., 4)] * vz[K] + GZ[K + (..., 5)] * vz[Km])
        vmh = np.zeros_like(w)
        vmh[K] = horiz * RGAMH[K] + vflag * w[K] * RGSQRTH[K]
        out.append((vx * RGAM, vy * RGAM, vz * RGAM, vmh))
    (ax, ay, az, vmh), (axp, ayp, azp, vmhp) = out
    div, div_pl = divergence(ctx, ax, ay, az, axp, ayp, azp)
    g = np.zeros_like(rw)
    g_pl = np.zeros_like(rw_pl)
    rd = ctx.rdgz[kmin:kmax + 1]
    g[:, :, kmin:kmax + 1] = -div[:, :, kmin:kmax + 1] - (vmh[:, :, kmin + 1:kmax + 2] - vmh[:, :, kmin:kmax + 1]) * rd[None, None, :, None]
    g_pl[:, kmin:kmax + 1] = -div_pl[:, kmin:kmax + 1] - (vmhp[:, kmin + 1:kmax + 2] - vmhp[:, kmin:kmax + 1]) * rd[None, :, None]
    return g, g_pl


def _ks(kax, lo, hi):
    return (slice(None),) * kax + (slice(lo, hi),)


def advection_convergence(ctx, rvx, rvy, rvz, rw, scl, rvx_pl, rvy_pl, rvz_pl, rw_pl, scl_pl, vertical):
    """Flux convergence of scl-weighted mass fluxes (kernels/advconv.py:42-108 + fluxconv)."""
    kmin, kmax = ctx.kmin, ctx.kmax
    af, bf = ctx.afact, ctx.bfact

    def wscl(w, s, kax):
        out = np.zeros_like(w)
        shp = [1] * w.ndim
        shp[kax] = kmax + 2 - kmin
        a = af[kmin:kmax + 2].reshape(shp)
        b = bf[kmin:kmax + 2].reshape(shp)
        out[_ks(kax, kmin, kmax + 2)] = w[_ks(kax, kmin, kmax + 2)] * (
            a * s[_ks(kax, kmin, kmax + 2)] + b * s[_ks(kax, kmin - 1, kmax + 1)])
        return out

    if vertical:
        ws, ws_pl = wscl(rw, scl, 2), wscl(rw_pl, scl_pl, 1)
    else:
        ws, ws_pl = np.zeros_like(rw), np.zeros_like(rw_pl)
    return flux_convergence(ctx, rvx * scl, rvy * scl, rvz * scl, ws,
                            rvx_pl * scl_pl, rvy_pl * scl_pl, rvz_pl * scl_pl, ws_pl, vertical)


def pres_gradient(ctx, P, P_pl, with_vertical):
    """Pressure gradient force (kernels/presgrad.py:126-265), first-layer remedy on."""
    m = ctx.m
    kmin, kmax, kall = ctx.kmin, ctx.kmax, ctx.kall
    grad_h, grad_h_pl = gradient(ctx, P * m["RGAM"], P_pl * m["RGAM_pl"])
    outs = []
    for reg in (True, False):
        tag = "" if reg else "_pl"
        kax = 2 if reg else 1
        PP = P if reg else P_pl
        gh = grad_h if reg else grad_h_pl
        GZ, RGAMH = m["C2WfactGz" + tag], m["RGAMH" + tag]
        K, Km = _ks(kax, kmin, kmax + 2), _ks(kax, kmin - 1, kmax + 1)
        pvmh = np.stack([(GZ[K + (..., 0)] * PP[K] + GZ[K + (..., 1)] * PP[Km]) * RGAMH[K],
                         (GZ[K + (..., 2)] * PP[K] + GZ[K + (..., 3)] * PP[Km]) * RGAMH[K],
                         (GZ[K + (..., 4)] * PP[K] + GZ[K + (..., 5)] * PP[Km]) * RGAMH[K]], axis=-1)
        shp = [1] * pvmh.ndim
        shp[kax] = kmax + 1 - kmin
        rd = ctx.rdgz[kmin:kmax + 1].reshape(shp)
        nk = kmax + 2 - kmin
        vterm = (pvmh[_ks(kax, 1, nk)] - pvmh[_ks(kax, 0, nk - 1)]) * rd
        gf = gh[_ks(kax, kmin, kmax + 1)] + vterm
        Pg = np.zeros(PP.shape + (3,))
        Pg[_ks(kax, kmin + 1, kmax + 1)] = gf[_ks(kax, 1, kmax + 1 - kmin)]
        Pg[_ks(kax, kmin, kmin + 1)] = gf[_ks(kax, 1, 2)]          # first-layer remedy
        if reg:
            g = ctx.g
            s = slice(1, g - 1)
            nx, ny, nz = horizontalize_interior(ctx, Pg[..., 0], Pg[..., 1], Pg[..., 2])
            Pg[s, s, ..., 0], Pg[s, s, ..., 1], Pg[s, s, ..., 2] = nx, ny, nz
        else:
            Pg[..., 0], Pg[..., 1], Pg[..., 2] = horizontalize_pole(ctx, Pg[..., 0], Pg[..., 1], Pg[..., 2])
        Pw = np.zeros_like(PP)
        if with_vertical:
            GAM2H, RGSGAM2 = m["GAM2H" + tag], m["RGSGAM2" + tag]
            K1, K0 = _ks(kax, kmin + 1, kmax + 1), _ks(kax, kmin, kmax)
            shp2 = [1] * PP.ndim
            shp2[kax] = kmax - kmin
            rdh = ctx.rdgzh[kmin + 1:kmax + 1].reshape(shp2)
            Pw[K1] = GAM2H[K1] * (PP[K1] * RGSGAM2[K1] - PP[K0] * RGSGAM2[K0]) * rdh
        outs.append((Pg, Pw))
    (Pg, Pw), (Pg_pl, Pw_pl) = outs
    return Pg, Pw, Pg_pl, Pw_pl


def buoyancy(ctx, rhog, rhog_pl):
    """-g * rho interpolated to half levels (kernels/buoyancy.py:40-81)."""
    kmin, kmax, grav = ctx.kmin, ctx.kmax, ctx.c.GRAV
    out = []
    for r, C2W, kax in ((rhog, ctx.m["C2Wfact"], 2), (rhog_pl, ctx.m["C2Wfact_pl"], 1)):
        b = np.zeros_like(r)
        K, Km = _ks(kax, kmin + 1, kmax + 1), _ks(kax, kmin, kmax)
        b[K] = -grav * (C2W[K + (..., 0)] * r[K] + C2W[K + (..., 1)] * r[Km])
        out.append(b)
    return out[0], out[1]


def rhogkin(ctx, rhog, rvx, rvy, rvz, rw, tag, kax):
    """Kinetic energy density at full levels (kernels/rhogkin.py:38-79)."""
    kmin, kmax = ctx.kmin, ctx.kmax
    C2W, W2C = ctx.m["C2Wfact" + tag], ctx.m["W2Cfact" + tag]
    kc, kvi, kcm = _ks(kax, kmin, kmax + 1), _ks(kax, kmin + 1, kmax + 1), _ks(kax, kmin, kmax)
    kh = 0.5 * (rvx[kc] ** 2 + rvy[kc] ** 2 + rvz[kc] ** 2) / rhog[kc]
    denom = C2W[kvi + (..., 0)] * rhog[kvi] + C2W[kvi + (..., 1)] * rhog[kcm]
    vint = 0.5 * rw[kvi] ** 2 / denom
    vb = np.zeros_like(rhog)
    vb[kvi] = vint
    up = vb[_ks(kax, kmin + 1, kmax + 2)]
    dn = vb[_ks(kax, kmin, kmax + 1)]
    tot = kh + W2C[kc + (..., 0)] * up + W2C[kc + (..., 1)] * dn
    out = np.zeros_like(rhog)
    out[kc] = tot
    return out


def hdiff_tendency(vt, vtp, rhog, rhog_p, rhog_h, rhog_h_p, kh, kh_p, kh_h, kh_h_p):
    """Hyperdiffusion tendency from the iterated Laplacians vt (6 components) with the
    Laplacian-1 filter off (its terms are the literal 0.0): Numf._hdiff_tendency_host
    (mod_numfilter.py:1519-1578)."""
    out = []
    for v_, rg, rgh, k, kh_ in ((vt, rhog, rhog_h, kh, kh_h), (vtp, rhog_p, rhog_h_p, kh_p, kh_h_p)):
        t = np.zeros(rg.shape + (6,))
        t[..., I_RHOGVX] = -(v_[0] * k + 0.0) * rg
        t[..., I_RHOGVY] = -(v_[1] * k + 0.0) * rg
        t[..., I_RHOGVZ] = -(v_[2] * k + 0.0) * rg
        t[..., I_RHOGW] = -(v_[3] * kh_ + 0.0) * rgh
        t[..., I_RHOGE] = -(v_[4] + 0.0)
        t[..., I_RHOG] = -(v_[5] + 0.0)
        out.append(t)
    return out[0], out[1]


def thomas(S, Mc, Mu, Ml):
    """Tridiagonal solve along axis 0 (kernels/virhowsolver.py:42-64)."""
    N = len(S)
    beta = Mc[0]
    u = [S[0] / beta]
    gam = [None] * N
    for i in range(1, N):
        gm = Mu[i - 1] / beta
        beta = Mc[i] - Ml[i] * gm
        u.append((S[i] - Ml[i] * u[i - 1]) / beta)
        gam[i] = gm
    x = [None] * N
    x[N - 1] = u[N - 1]
    for i in range(N - 2, -1, -1):
        x[i] = u[i] - gam[i + 1] * x[i + 1]
    return x


# ======================================================================
# the dynamics step
# ======================================================================
def default_tracer(dyn, rhog_in, rhog_in_p, PM, PMp, f_rhog, f_rhog_p, dtl):
    """MIURA2004 transport of all tracers with the Thuburn limiters (mod_dynamics.py:2346-2461)."""
    from .tracer import tracer_advection
    tracer_advection(dyn.ctx, dyn.PROGq, dyn.PROGq_pl, rhog_in, rhog_in_p, PM, PMp, f_rhog, f_rhog_p, dtl,
                     limiter=bool(dyn.ctx.c.THUBURN_LIM))

class OracleDynamics:
    """One RK3/HEVI large step (mod_dynamics.py:1725-2741, numpy branches).

    State lives in the reference layout: PROG (g, g, kall, l, 6), PROG_pl
    (6, kall, 2, 6), PROGq (g, g, kall, l, nq), PROGq_pl (6, kall, 2, nq).
    """

    def __init__(self, ctx, state):
        self.ctx = ctx
        self.PROG = state["PROG"].copy()
        self.PROG_pl = state["PROG_pl"].copy()
        self.PROGq = state["PROGq"].copy()
        self.PROGq_pl = state["PROGq_pl"].copy()
        g, kall, L = ctx.g, ctx.kall, ctx.topo.lall
        self.DIAG = np.zeros((g, g, kall, L, 6))
        self.DIAG_pl = np.zeros((6, kall, 2, 6))
        self.tracer = default_tracer   # tracer transport of the last RK stage (None: skip)

    # ------------------------------------------------------------------
    def prepost(self):
        """Diagnostics + boundary conditions + eth (mod_dynamics.py:965-1283)."""
        ctx, c, m = self.ctx, self.ctx.c, self.ctx.m
        kmin, kmax = ctx.kmin, ctx.kmax
        out = {}
        for reg in (True, False):
            tag = "" if reg else "_pl"
            kax = 2 if reg else 1
            PROG = self.PROG if reg else self.PROG_pl
            PROGq = self.PROGq if reg else self.PROGq_pl
            DIAG = self.DIAG if reg else self.DIAG_pl
            GSGAM2, C2W, PHI, GZ = m["GSGAM2" + tag], m["C2Wfact" + tag], m["PHI" + tag], m["C2WfactGz" + tag]
            rhog = PROG[..., I_RHOG]
            rho = rhog / GSGAM2
            vx = PROG[..., I_RHOGVX] / rhog
            vy = PROG[..., I_RHOGVY] / rhog
            vz = PROG[..., I_RHOGVZ] / rhog
            ein = PROG[..., I_RHOGE] / rhog
            q = PROGq / rhog[..., None]
            qw = q[..., c.NQW_STR:c.NQW_END + 1]
            qd = 1.0 - np.sum(qw, axis=-1)
            cv = np.sum(qw * c.CVW[c.NQW_STR:c.NQW_END + 1], axis=-1) + qd * c.CVdry
            tem = ein / cv
            pre = rho * tem * (qd * c.Rdry + q[..., c.I_QV] * c.Rvap)
            w = DIAG[..., I_w].copy()
            K, Km = _ks(kax, kmin + 1, kmax + 1), _ks(kax, kmin, kmax)
            w[K] = PROG[K + (..., I_RHOGW)] / (C2W[K + (..., 0)] * rhog[K] + C2W[K + (..., 1)] * rhog[Km])
            DIAG[...] = np.stack([pre, tem, vx, vy, vz, w], axis=-1)
            self._bndcnd(DIAG, PROG, rho, ein, GSGAM2, PHI, C2W, GZ, kax)
            eth = ein + DIAG[..., I_pre] / rho
            pbs, rbs = ctx.bs("pre", reg), ctx.bs("rho", reg)
            pregd = (DIAG[..., I_pre] - pbs) * GSGAM2
            rhogd = (rho - rbs) * GSGAM2
            out["rho" + tag], out["ein" + tag], out["eth" + tag] = rho, ein, eth
            out["pregd" + tag], out["rhogd" + tag], out["q" + tag] = pregd, rhogd, q
        return out

    def _bndcnd(self, DIAG, PROG, rho, ein, GSGAM2, PHI, C2W, GZ, kax):
        """BNDCND_all: TEM/TEM thermo, FREE top / RIGID bottom (mod_bndcnd.py:153-238)."""
        c = self.ctx.c
        kmin, kmax = self.ctx.kmin, self.ctx.kmax
        kt, kb = kmax + 1, kmin - 1

        def K(k):
            return (slice(None),) * kax + (k,)
        tem, pre = DIAG[..., I_tem], DIAG[..., I_pre]
        tem_t, tem_b = tem[K(kmax)].copy(), tem[K(kmin)].copy()
        pre_t = pre[K(kmax - 1)] - rho[K(kmax)] * (PHI[K(kt)] - PHI[K(kmax - 1)])
        pre_b = pre[K(kmin + 1)] - rho[K(kmin)] * (PHI[K(kb)] - PHI[K(kmin + 1)])
        tem[K(kt)], tem[K(kb)] = tem_t, tem_b
        pre[K(kt)], pre[K(kb)] = pre_t, pre_b
        rho[K(kt)] = pre_t / (c.Rdry * tem_t)
        rho[K(kb)] = pre_b / (c.Rdry * tem_b)
        rhog = PROG[..., I_RHOG]
        rhog[K(kt)] = rho[K(kt)] * GSGAM2[K(kt)]
        rhog[K(kb)] = rho[K(kb)] * GSGAM2[K(kb)]
        ein[K(kt)] = c.CVdry * tem[K(kt)]
        ein[K(kb)] = c.CVdry * tem[K(kb)]
        PROG[K(kt) + (..., I_RHOGE)] = rhog[K(kt)] * ein[K(kt)]
        PROG[K(kb) + (..., I_RHOGE)] = rhog[K(kb)] * ein[K(kb)]
        for iv in (I_RHOGVX, I_RHOGVY, I_RHOGVZ):
            top = 1.0 * (PROG[K(kmax) + (..., iv)] / rhog[K(kmax)]) * rhog[K(kt)]
            btm = -1.0 * (PROG[K(kmin) + (..., iv)] / rhog[K(kmin)]) * rhog[K(kb)]
            PROG[K(kt) + (..., iv)] = top
            PROG[K(kb) + (..., iv)] = btm
        for kk in (kt, kb):
            DIAG[K(kk) + (..., I_vx)] = PROG[K(kk) + (..., I_RHOGVX)] / rhog[K(kk)]
            DIAG[K(kk) + (..., I_vy)] = PROG[K(kk) + (..., I_RHOGVY)] / rhog[K(kk)]
            DIAG[K(kk) + (..., I_vz)] = PROG[K(kk) + (..., I_RHOGVZ)] / rhog[K(kk)]
        rvx, rvy, rvz = PROG[..., I_RHOGVX], PROG[..., I_RHOGVY], PROG[..., I_RHOGVZ]
        rw_top = -(GZ[K(kt) + (..., 0)] * rvx[K(kt)] + GZ[K(kt) + (..., 1)] * rvx[K(kmax)]
                   + GZ[K(kt) + (..., 2)] * rvy[K(kt)] + GZ[K(kt) + (..., 3)] * rvy[K(kmax)]
                   + GZ[K(kt) + (..., 4)] * rvz[K(kt)] + GZ[K(kt) + (..., 5)] * rvz[K(kmax)])
        PROG[K(kt) + (..., I_RHOGW)] = rw_top
        PROG[K(kmin) + (..., I_RHOGW)] = 0.0
        PROG[K(kb) + (..., I_RHOGW)] = 0.0
        rw = PROG[..., I_RHOGW]
        DIAG[K(kt) + (..., I_w)] = rw[K(kt)] / (C2W[K(kt) + (..., 0)] * rhog[K(kt)] + C2W[K(kt) + (..., 1)] * rhog[K(kmax)])
        DIAG[K(kmin) + (..., I_w)] = rw[K(kmin)] / (C2W[K(kmin) + (..., 0)] * rhog[K(kmin)] + C2W[K(kmin) + (..., 1)] * rhog[K(kb)])
        DIAG[K(kb) + (..., I_w)] = 0.0

    # ------------------------------------------------------------------
    def advection_momentum(self):
        """Momentum advection + Coriolis (mod_src.py:96-497, kernels/advconvmom.py)."""
        ctx, c = self.ctx, self.ctx.c
        kmin, kmax, R = ctx.kmin, ctx.kmax, ctx.rscale
        P, Pp, D, Dp = self.PROG, self.PROG_pl, self.DIAG, self.DIAG_pl
        vv = []
        for D_, x_, kax in ((D, ctx.geo.GRD_x[:, :, 0], 2), (Dp, ctx.geo.GRD_x_pl[:, 0], 1)):
            shp = [1] * D_[..., 0].ndim
            shp[kax] = kmax + 1 - kmin
            cf = ctx.cfact[kmin:kmax + 1].reshape(shp)
            df = ctx.dfact[kmin:kmax + 1].reshape(shp)
            w = D_[..., I_w]
            wc = cf * w[_ks(kax, kmin + 1, kmax + 2)] + df * w[_ks(kax, kmin, kmax + 1)]
            if kax == 2:
                gxyz = [x_[:, :, None, :, d] for d in range(3)]
            else:
                gxyz = [x_[:, None, :, d] for d in range(3)]
            out = []
            for d, iv in enumerate((I_vx, I_vy, I_vz)):
                a = np.zeros_like(w)
                a[_ks(kax, kmin, kmax + 1)] = D_[_ks(kax, kmin, kmax + 1) + (..., iv)] + wc * gxyz[d] / R
                out.append(a)
            vv.append(out)
        (vvx, vvy, vvz), (vvx_p, vvy_p, vvz_p) = vv
        args = (P[..., I_RHOGVX], P[..., I_RHOGVY], P[..., I_RHOGVZ], P[..., I_RHOGW])
        args_p = (Pp[..., I_RHOGVX], Pp[..., I_RHOGVY], Pp[..., I_RHOGVZ], Pp[..., I_RHOGW])
        dvv = [advection_convergence(ctx, *args, s, *args_p, sp, True)
               for s, sp in ((vvx, vvx_p), (vvy, vvy_p), (vvz, vvz_p))]
        res = []
        for reg in (True, False):
            idx = 0 if reg else 1
            kax = 2 if reg else 1
            tag = "" if reg else "_pl"
            rh = (P if reg else Pp)[..., I_RHOG]
            vvx_, vvy_ = (vvx, vvy) if reg else (vvx_p, vvy_p)
            dx, dy, dz = dvv[0][idx], dvv[1][idx], dvv[2][idx]
            x_ = ctx.geo.GRD_x[:, :, 0] if reg else ctx.geo.GRD_x_pl[:, 0]
            if reg:
                gx, gy, gz = (x_[:, :, None, :, d] / R for d in range(3))
            else:
                gx, gy, gz = (x_[:, None, :, d] / R for d in range(3))
            kc = _ks(kax, kmin, kmax + 1)
            dxc = dx[kc] - (-2.0 * rh[kc] * (c.OHM * vvy_[kc]))
            dyc = dy[kc] - (2.0 * rh[kc] * (c.OHM * vvx_[kc]))
            dzc = dz[kc]
            prd = dxc * gx + dyc * gy + dzc * gz
            gv = [np.zeros_like(rh) for _ in range(4)]
            gv[0][kc], gv[1][kc], gv[2][kc] = dxc - prd * gx, dyc - prd * gy, dzc - prd * gz
            gwc = prd * c.ALPHA
            C2W = ctx.m["C2Wfact" + tag]
            K = _ks(kax, kmin + 1, kmax + 1)
            nk = kmax + 1 - kmin
            gv[3][K] = C2W[K + (..., 0)] * gwc[_ks(kax, 1, nk)] + C2W[K + (..., 1)] * gwc[_ks(kax, 0, nk - 1)]
            res.append(gv)
        return res[0], res[1]

    def hdiffusion(self, pp):
        """Fourth-order horizontal diffusion tendency (mod_numfilter.py:845-1469)."""
        ctx, c = self.ctx, self.ctx.c
        kmin, kmax = ctx.kmin, ctx.kmax
        D, Dp = self.DIAG, self.DIAG_pl
        rhog, rhog_p = self.PROG[..., I_RHOG], self.PROG_pl[..., I_RHOG]
        Kh, Kh_p = c.GAMMA_H, c.GAMMA_H
        vt = [D[..., I_vx].copy(), D[..., I_vy].copy(), D[..., I_vz].copy(), D[..., I_w].copy(),
              D[..., I_tem] - ctx.bs("tem", True), pp["rho"] - ctx.bs("rho", True)]
        vtp = [Dp[..., I_vx].copy(), Dp[..., I_vy].copy(), Dp[..., I_vz].copy(), Dp[..., I_w].copy(),
               Dp[..., I_tem] - ctx.bs("tem", False), pp["rho_pl"] - ctx.bs("rho", False)]
        khf = np.full_like(rhog, Kh)
        khf_p = np.full_like(rhog_p, Kh_p)
        nlap = c.LAP_ORDER_HDIFF
        for p in range(nlap):
            new, newp = [], []
            for v in range(6):
                if p == nlap - 1 and v == 4:
                    o, op = diffusion(ctx, vt[4], vtp[4], rhog * c.CVdry * khf, rhog_p * c.CVdry * khf_p)
                elif p == nlap - 1 and v == 5:
                    o, op = diffusion(ctx, vt[5], vtp[5], rhog * c.HDIFF_FACT_RHO * khf,
                                      rhog_p * c.HDIFF_FACT_RHO * khf_p)
                else:
                    o, op = laplacian(ctx, vt[v], vtp[v])
                new.append(-o)
                newp.append(-op)
            V = np.stack(new, axis=-1)
            Vp = np.stack(newp, axis=-1)
            ctx.halo(V, Vp)
            vt = [V[..., i] for i in range(6)]
            vtp = [Vp[..., i] for i in range(6)]
        rgh = []
        for reg in (True, False):
            kax = 2 if reg else 1
            rg = rhog if reg else rhog_p
            C2W = ctx.m["C2Wfact" + ("" if reg else "_pl")]
            r = np.zeros_like(rg)
            K, Km = _ks(kax, kmin, kmax + 2), _ks(kax, kmin - 1, kmax + 1)
            r[K] = C2W[K + (..., 0)] * rg[K] + C2W[K + (..., 1)] * rg[Km]
            r[_ks(kax, kmin - 1, kmin)] = 0.0
            rgh.append(r)
        # the linear filter: Kh at half levels is Kh itself (mod_numfilter.py:1208-1211)
        T, Tp = hdiff_tendency(vt, vtp, rhog, rhog_p, rgh[0], rgh[1], khf, khf_p, khf, khf_p)
        horizontalize_inplace(ctx, T[..., 1], T[..., 2], T[..., 3], Tp[..., 1], Tp[..., 2], Tp[..., 3])
        return T, Tp

    def divdamp(self, rvx, rvy, rvz, rw, rvx_p, rvy_p, rvz_p, rw_p):
        """3D divergence damping, lap_order 2 (mod_numfilter.py:2032-2268)."""
        ctx, c, m = self.ctx, self.ctx.c, self.ctx.m
        kmin, kmax, g = ctx.kmin, ctx.kmax, ctx.g
        G = ctx.topo.gmax + 1
        # OPRT3D_divdamp (kernels/oprt3ddivdamp.py:50-173)
        outs = []
        for reg in (True, False):
            tag = "" if reg else "_pl"
            kax = 2 if reg else 1
            vx, vy, vz, w = (rvx, rvy, rvz, rw) if reg else (rvx_p, rvy_p, rvz_p, rw_p)
            GZ, RGAMH, RGSQRTH, RGAM = m["C2WfactGz" + tag], m["RGAMH" + tag], m["RGSQRTH" + tag], m["RGAM" + tag]
            K, Km = _ks(kax, kmin + 1, kmax + 1), _ks(kax, kmin, kmax)
            horiz = (GZ[K + (..., 0)] * vx[K] + GZ[K + (..., 1)] * vx[Km] + GZ[K + (..., 2)] * vy[K]
                     + GZ[K + (..., 3)] * vy[Km] + GZ[K + (..., 4)] * vz[K] + GZ[K + (..., 5)] * vz[Km])
            rwv = np.zeros_like(w)
            rwv[K] = horiz * RGAMH[K] + w[K] * RGSQRTH[K]
            Kc = _ks(kax, kmin, kmax + 1)
            outs.append((vx[Kc] * RGAM[Kc], vy[Kc] * RGAM[Kc], vz[Kc] * RGAM[Kc], rwv))
        (vmx, vmy, vmz, rwv), (vmxp, vmyp, vmzp, rwvp) = outs
        Rr, Rp = slice(0, G), slice(1, G + 1)
        UP, DN = slice(kmin + 1, kmax + 2), slice(kmin, kmax + 1)
        rdz = ctx.rdgz[kmin:kmax + 1][None, None, :, None]
        rw_ti = ((rwv[Rr, Rr, UP] + rwv[Rp, Rr, UP] + rwv[Rp, Rp, UP])
                 - (rwv[Rr, Rr, DN] + rwv[Rp, Rr, DN] + rwv[Rp, Rp, DN])) / 3.0 * rdz
        rw_tj = ((rwv[Rr, Rr, UP] + rwv[Rp, Rp, UP] + rwv[Rr, Rp, UP])
                 - (rwv[Rr, Rr, DN] + rwv[Rp, Rp, DN] + rwv[Rr, Rp, DN])) / 3.0 * rdz
        ti, tj = _vertex_values(ctx, vmx, vmy, vmz, True)
        ti = ti + rw_ti
        tj = tj + rw_tj
        _pntmask_corner(ctx, ti, tj)
        vt2 = np.zeros(rvx.shape + (3,))
        for d in range(3):
            vt2[1:G, 1:G, kmin:kmax + 1, :, d] = _hex_flux_sum(ctx, ti, tj, d)
        # pole
        cI = ctx.geo.coef_intp_pl[1:6]
        cd = ctx.geo.coef_diff_pl[1:6]

        def b(a):
            return a[:, None, :]
        Vs = [vmxp[0], vmyp[0], vmzp[0]]
        Vi = [vmxp[1:6], vmyp[1:6], vmzp[1:6]]
        Vp = [np.roll(a, -1, 0) for a in Vi]
        rup, rdn = rwvp[:, UP], rwvp[:, DN]
        srw = ((rup[0] + rup[1:6] + np.roll(rup[1:6], -1, 0))
               - (rdn[0] + rdn[1:6] + np.roll(rdn[1:6], -1, 0))) / 3.0 * ctx.rdgz[kmin:kmax + 1][None, :, None]
        sclt = (b(cI[:, :, XDIR, 0]) * Vs[0] + b(cI[:, :, XDIR, 1]) * Vi[0] + b(cI[:, :, XDIR, 2]) * Vp[0]
                + b(cI[:, :, YDIR, 0]) * Vs[1] + b(cI[:, :, YDIR, 1]) * Vi[1] + b(cI[:, :, YDIR, 2]) * Vp[1]
                + b(cI[:, :, ZDIR, 0]) * Vs[2] + b(cI[:, :, ZDIR, 1]) * Vi[2] + b(cI[:, :, ZDIR, 2]) * Vp[2]
                + srw)
        pair = np.roll(sclt, 1, 0) + sclt
        vt2p = np.zeros(rvx_p.shape + (3,))
        for d in range(3):
            vt2p[0, kmin:kmax + 1, :, d] = (b(cd[:, :, d]) * pair).sum(axis=0)
        ctx.halo(vt2, vt2p)
        # post-exchange: second application, coefficient, horizontalize
        o, op = divdamp_op(ctx, -vt2[..., 0], -vt2[..., 1], -vt2[..., 2],
                           -vt2p[..., 0], -vt2p[..., 1], -vt2p[..., 2])
        coef = c.ALPHA_D
        gd = [coef * o[d] for d in range(3)]
        gdp = [coef * op[d] for d in range(3)]
        s = slice(1, g - 1)
        nx, ny, nz = horizontalize_interior(ctx, gd[0], gd[1], gd[2])
        gd[0][s, s], gd[1][s, s], gd[2][s, s] = nx, ny, nz
        gdp = list(horizontalize_pole(ctx, gdp[0], gdp[1], gdp[2]))
        return gd, gdp

    # ------------------------------------------------------------------
    def vi_small_step(self, pp, g_TEND, g_TEND_pl, PROG_split, PROG_split_pl, num_itr, dt):
        """HEVI small steps (mod_vi.py:104-1590, numpy path)."""
        ctx, c, m = self.ctx, self.ctx.c, self.ctx.m
        kmin, kmax, kall = ctx.kmin, ctx.kmax, ctx.kall
        P, Pp, D, Dp = self.PROG, self.PROG_pl, self.DIAG, self.DIAG_pl
        alpha, GRAV, RovCV = c.ALPHA, c.GRAV, c.Rdry / c.CVdry
        eth, eth_p = pp["eth"], pp["eth_pl"]
        # half-level interpolations
        half = {}
        for reg in (True, False):
            tag = "" if reg else "_pl"
            kax = 2 if reg else 1
            PR = P if reg else Pp
            et = eth if reg else eth_p
            C2W = m["C2Wfact" + tag]
            K, Km = _ks(kax, kmin, kmax + 2), _ks(kax, kmin - 1, kmax + 1)
            rh = np.zeros_like(et)
            rh[K] = C2W[K + (..., 0)] * PR[K + (..., I_RHOG)] + C2W[K + (..., 1)] * PR[Km + (..., I_RHOG)]
            rh[_ks(kax, kmin - 1, kmin)] = rh[_ks(kax, kmin, kmin + 1)]
            shp = [1] * et.ndim
            shp[kax] = kmax + 2 - kmin
            eh = np.zeros_like(et)
            eh[K] = ctx.afact[kmin:kmax + 2].reshape(shp) * et[K] + ctx.bfact[kmin:kmax + 2].reshape(shp) * et[Km]
            eh[_ks(kax, kmin - 1, kmin)] = eh[_ks(kax, kmin, kmin + 1)]
            half["rhog_h" + tag], half["eth_h" + tag] = rh, eh
        flux = (P[..., I_RHOGVX], P[..., I_RHOGVY], P[..., I_RHOGVZ], P[..., I_RHOGW])
        flux_p = (Pp[..., I_RHOGVX], Pp[..., I_RHOGVY], Pp[..., I_RHOGVZ], Pp[..., I_RHOGW])
        drhog, drhog_p = flux_convergence(ctx, *flux, *flux_p, True)
        dd, ddp = self.divdamp(*flux, *flux_p)
        dpg, dpgw, dpg_p, dpgw_p = pres_gradient(ctx, pp["pregd"], pp["pregd_pl"], True)
        dbuo, dbuo_p = buoyancy(ctx, pp["rhogd"], pp["rhogd_pl"])
        drhoge, drhoge_p = advection_convergence(ctx, *flux, eth, *flux_p, eth_p, True)
        gT = []
        gz_t = {}
        for reg in (True, False):
            tag = "" if reg else "_pl"
            kax = 2 if reg else 1
            PR, DG = (P, D) if reg else (Pp, Dp)
            g0 = g_TEND if reg else g_TEND_pl
            dpg_, dpgw_, dbuo_ = (dpg, dpgw, dbuo) if reg else (dpg_p, dpgw_p, dbuo_p)
            dr, de = (drhog, drhoge) if reg else (drhog_p, drhoge_p)
            dd_ = dd if reg else ddp
            gz = GRAV - (dpgw_ - dbuo_) / half["rhog_h" + tag]
            pwh = -gz * PR[..., I_RHOGW]
            W2C = m["W2Cfact" + tag]
            pw = np.zeros_like(gz)
            kc, kp = _ks(kax, kmin, kmax + 1), _ks(kax, kmin + 1, kmax + 2)
            pw[kc] = (DG[kc + (..., I_vx)] * dpg_[kc + (..., XDIR)] + DG[kc + (..., I_vy)] * dpg_[kc + (..., YDIR)]
                      + DG[kc + (..., I_vz)] * dpg_[kc + (..., ZDIR)]
                      + W2C[kc + (..., 0)] * pwh[kp] + W2C[kc + (..., 1)] * pwh[kc])
            t = np.zeros_like(g0)
            t[..., I_RHOG] = g0[..., I_RHOG] + dr
            t[..., I_RHOGVX] = g0[..., I_RHOGVX] - dpg_[..., XDIR] + dd_[0] + 0.0
            t[..., I_RHOGVY] = g0[..., I_RHOGVY] - dpg_[..., YDIR] + dd_[1] + 0.0
            t[..., I_RHOGVZ] = g0[..., I_RHOGVZ] - dpg_[..., ZDIR] + dd_[2] + 0.0
            t[..., I_RHOGW] = g0[..., I_RHOGW] + 0.0 * alpha - dpgw_ + dbuo_
            t[..., I_RHOGE] = g0[..., I_RHOGE] + de + pw
            gT.append(t)
            gz_t[tag] = gz
        gTr, gTp = gT
        grhogetot0, grhogetot0_p = g_TEND[..., I_RHOGE].copy(), g_TEND_pl[..., I_RHOGE].copy()
        PROG_mean = P[..., 0:5].copy()
        PROG_mean_p = Pp[..., 0:5].copy()
        Mat = {}
        for reg in (True, False):
            tag = "" if reg else "_pl"
            Mat[tag] = self._rhow_matrix(half["eth_h" + tag], gz_t[tag], dt, tag, 2 if reg else 1)
        rw_itr = 1.0 / float(num_itr)
        PS, PSp = PROG_split, PROG_split_pl
        PR0 = P.copy()
        PR0p = Pp.copy()
        for _ in range(int(num_itr)):
            preg = PS[..., I_RHOGE] * RovCV
            preg[:, :, kmin - 1] = preg[:, :, kmin]
            preg[:, :, kmax + 1] = preg[:, :, kmax]
            PS[:, :, kmin - 1, :, I_RHOGE] = PS[:, :, kmin, :, I_RHOGE]
            PS[:, :, kmax + 1, :, I_RHOGE] = PS[:, :, kmax, :, I_RHOGE]
            preg_p = PSp[..., I_RHOGE] * RovCV
            preg_p[:, kmin - 1] = preg_p[:, kmin]
            preg_p[:, kmax + 1] = preg_p[:, kmax]
            PSp[:, kmin - 1, :, I_RHOGE] = PSp[:, kmin, :, I_RHOGE]
            PSp[:, kmax + 1, :, I_RHOGE] = PSp[:, kmax, :, I_RHOGE]
            ddv, ddvp = self.divdamp(PS[..., I_RHOGVX], PS[..., I_RHOGVY], PS[..., I_RHOGVZ], PS[..., I_RHOGW],
                                     PSp[..., I_RHOGVX], PSp[..., I_RHOGVY], PSp[..., I_RHOGVZ], PSp[..., I_RHOGW])
            DV, DVp, drhogw, drhogw_p = self._vi_path1(preg, preg_p, gTr, gTp, ddv, ddvp, PS, PSp, PR0, PR0p, dt)
            ctx.halo(DV, DVp)
            we = self._vi_main(DV, DVp, PS, PSp, preg, preg_p, PR0, PR0p, eth, eth_p,
                               gTr, gTp, drhogw, drhogw_p, grhogetot0, grhogetot0_p, Mat, dt)
            WE = np.stack(we[0], axis=-1)
            WEp = np.stack(we[1], axis=-1)
            ctx.halo(WE, WEp)
            self._vi_path2(PS, DV, WE, PROG_mean, rw_itr)
            self._vi_path2(PSp, DVp, WEp, PROG_mean_p, rw_itr)
        P += PS
        Pp += PSp
        horizontalize_inplace(ctx, P[..., 1], P[..., 2], P[..., 3], Pp[..., 1], Pp[..., 2], Pp[..., 3])
        ctx.halo(PROG_mean, PROG_mean_p)
        return PROG_mean, PROG_mean_p

    def _vi_path1(self, preg, preg_p, gTr, gTp, ddv, ddvp, PS, PSp, PR0, PR0p, dt):
        """Horizontal explicit momentum update of one small step with its boundary rows
        (kernels/vipath1.py:68-160; the vertical and 2-D divergence damping are off)."""
        ctx = self.ctx
        kmin, kmax, alpha = ctx.kmin, ctx.kmax, ctx.c.ALPHA
        Pg, _, Pg_p, _ = pres_gradient(ctx, preg, preg_p, False)
        dvh, dvh_p = [], []
        for reg in (True, False):
            kax = 2 if reg else 1
            PS_ = PS if reg else PSp
            gT_ = gTr if reg else gTp
            Pg_ = Pg if reg else Pg_p
            dd_ = ddv if reg else ddvp
            rg = (PR0 if reg else PR0p)[..., I_RHOG]
            comps = []
            for d, iv in enumerate((I_RHOGVX, I_RHOGVY, I_RHOGVZ)):
                drv = gT_[..., iv] - Pg_[..., d] + dd_[d] + 0.0
                comps.append(PS_[..., iv] + drv * dt)
            for a in comps:
                kt = _ks(kax, kmax + 1, kmax + 2)
                kb = _ks(kax, kmin - 1, kmin)
                top = 1.0 * (a[_ks(kax, kmax, kmax + 1)] / rg[_ks(kax, kmax, kmax + 1)]) * rg[kt]
                btm = -1.0 * (a[_ks(kax, kmin, kmin + 1)] / rg[_ks(kax, kmin, kmin + 1)]) * rg[kb]
                a[kt] = top
                a[kb] = btm
            (dvh if reg else dvh_p).extend(comps)
        drhogw, drhogw_p = gTr[..., I_RHOGW] + 0.0 * alpha, gTp[..., I_RHOGW] + 0.0 * alpha
        return np.stack(dvh, axis=-1), np.stack(dvh_p, axis=-1), drhogw, drhogw_p

    @staticmethod
    def _vi_path2(PS, DV, WE, PROG_mean, rw_itr):
        """Split-state write-back and running mean of one small step (kernels/vipath2.py:60-87)."""
        PS[..., I_RHOGVX], PS[..., I_RHOGVY], PS[..., I_RHOGVZ] = DV[..., 0], DV[..., 1], DV[..., 2]
        PS[..., I_RHOG], PS[..., I_RHOGW], PS[..., I_RHOGE] = WE[..., 0], WE[..., 1], WE[..., 2]
        PROG_mean += PS[..., 0:5] * rw_itr

    def _rhow_matrix(self, eth_h, g_tilde, dt, tag, kax):
        """Tridiagonal coefficients (kernels/vimatrix.py:67-136)."""
        ctx, c, m = self.ctx, self.ctx.c, self.ctx.m
        kmin, kmax = ctx.kmin, ctx.kmax
        ks, ksp, ksm = _ks(kax, kmin + 1, kmax + 1), _ks(kax, kmin + 2, kmax + 2), _ks(kax, kmin, kmax)
        shp = [1] * eth_h.ndim
        shp[kax] = kmax - kmin
        GCVovR = c.GRAV * c.CVdry / c.Rdry
        rg, rgm = ctx.rdgz[kmin + 1:kmax + 1].reshape(shp), ctx.rdgz[kmin:kmax].reshape(shp)
        RGSGAM2, GAM2H, RGAMH = m["RGSGAM2" + tag], m["GAM2H" + tag], m["RGAMH" + tag]
        Ca = (RGSGAM2[ks] * rg + RGSGAM2[ksm] * rgm) * GAM2H[ks]
        Cb = (ctx.dfact[kmin + 1:kmax + 1] - ctx.cfact[kmin:kmax]).reshape(shp)
        Cu1 = RGSGAM2[ks] * rg * GAM2H[ksp]
        Cu2 = GAM2H[ksp] * RGAMH[ks] ** 2 * GCVovR
        Cl1 = RGSGAM2[ks] * rg * GAM2H[ksm]
        Cl2 = GAM2H[ksm] * RGAMH[ks] ** 2 * GCVovR
        ACVovRt2 = c.ALPHA * c.CVdry / c.Rdry / (dt * dt)
        rgzh = ctx.rdgzh[kmin + 1:kmax + 1].reshape(shp)
        cf = ctx.cfact[kmin + 1:kmax + 1].reshape(shp)
        dfm = ctx.dfact[kmin:kmax].reshape(shp)
        Mc = ACVovRt2 / m["RGSQRTH" + tag][ks] + rgzh * (Ca * eth_h[ks] - Cb * (g_tilde[ks] + GCVovR))
        Mu = -rgzh * (Cu1 * eth_h[ksp] + cf * (g_tilde[ksp] + Cu2))
        Ml = -rgzh * (Cl1 * eth_h[ksm] - dfm * (g_tilde[ksm] + Cl2))
        return Mc, Mu, Ml

    def _vi_main(self, DV, DVp, PS, PSp, preg, preg_p, PR0, PR0p, eth, eth_p,
                 gT, gTp, drhogw, drhogw_p, getot, getot_p, Mat, dt):
        """Vertical implicit part of one small step (kernels/vimain.py:127-262)."""
        ctx, c, m = self.ctx, self.ctx.c, self.ctx.m
        kmin, kmax = ctx.kmin, ctx.kmax
        s1 = (DV[..., 0], DV[..., 1], DV[..., 2])
        s1p = (DVp[..., 0], DVp[..., 1], DVp[..., 2])
        w0, w0p = PS[..., I_RHOGW], PSp[..., I_RHOGW]
        drhog, drhog_p = flux_convergence(ctx, *s1, w0, *s1p, w0p, False)
        drhoge, drhoge_p = advection_convergence(ctx, *s1, w0, eth, *s1p, w0p, eth_p, False)
        outs = []
        for reg in (True, False):
            tag = "" if reg else "_pl"
            kax = 2 if reg else 1
            v1 = s1 if reg else s1p
            PS_, PR = (PS, PR0) if reg else (PSp, PR0p)
            g_ = gT if reg else gTp
            grhog1 = g_[..., I_RHOG] + (drhog if reg else drhog_p)
            grhoge1 = g_[..., I_RHOGE] + (drhoge if reg else drhoge_p)
            gpre = grhoge1 * c.Rdry / c.CVdry
            GZ = m["C2WfactGz" + tag]
            kt = _ks(kax, kmax + 1, kmax + 2)
            km = _ks(kax, kmax, kmax + 1)
            rw_bc = np.zeros_like(w0 if reg else w0p)
            rw_bc[kt] = -(GZ[kt + (..., 0)] * v1[0][kt] + GZ[kt + (..., 1)] * v1[0][km]
                          + GZ[kt + (..., 2)] * v1[1][kt] + GZ[kt + (..., 3)] * v1[1][km]
                          + GZ[kt + (..., 4)] * v1[2][kt] + GZ[kt + (..., 5)] * v1[2][km])
            rw_bc[_ks(kax, kmin, kmin + 1)] = 0.0
            Mc, Mu, Ml = Mat[tag]
            outs.append(self._rhow_solve(rw_bc, PS_[..., I_RHOGW], (preg if reg else preg_p),
                                         PS_[..., I_RHOG], grhog1,
                                         (drhogw if reg else drhogw_p), gpre, Mc, Mu, Ml, dt, tag, kax))
        rw1, rw1p = outs
        drhog2, drhog2_p = flux_convergence(ctx, *s1, rw1, *s1p, rw1p, True)
        res = []
        for reg in (True, False):
            tag = "" if reg else "_pl"
            kax = 2 if reg else 1
            PS_, PR = (PS, PR0) if reg else (PSp, PR0p)
            v1 = s1 if reg else s1p
            rw1_ = rw1 if reg else rw1p
            g_ = gT if reg else gTp
            rhog_s1 = PS_[..., I_RHOG] + (g_[..., I_RHOG] + (drhog2 if reg else drhog2_p)) * dt
            r0 = PR[..., I_RHOG]
            kin0 = rhogkin(ctx, r0, PR[..., I_RHOGVX], PR[..., I_RHOGVY], PR[..., I_RHOGVZ], PR[..., I_RHOGW], tag, kax)
            kin10 = rhogkin(ctx, r0 + PS_[..., I_RHOG], PR[..., I_RHOGVX] + PS_[..., I_RHOGVX],
                            PR[..., I_RHOGVY] + PS_[..., I_RHOGVY], PR[..., I_RHOGVZ] + PS_[..., I_RHOGVZ],
                            PR[..., I_RHOGW] + PS_[..., I_RHOGW], tag, kax)
            b_ = (r0 + rhog_s1, PR[..., I_RHOGVX] + v1[0], PR[..., I_RHOGVY] + v1[1],
                  PR[..., I_RHOGVZ] + v1[2], PR[..., I_RHOGW] + rw1_)
            kin11 = rhogkin(ctx, *b_, tag, kax)
            res.append((rhog_s1, kin0, kin10, kin11, b_))
        (rs1, k0, k10, k11, b), (rs1p, k0p, k10p, k11p, bp) = res
        PHI, PHIp = m["PHI"], m["PHI_pl"]
        et0 = eth + k0 / PR0[..., I_RHOG] + PHI
        et0p = eth_p + k0p / PR0p[..., I_RHOG] + PHIp
        dtot, dtot_p = advection_convergence(ctx, b[1], b[2], b[3], b[4], et0, bp[1], bp[2], bp[3], bp[4], et0p, True)
        re1 = PS[..., I_RHOGE] + (getot + dtot) * dt + (k10 - k11) + (PS[..., I_RHOG] - rs1) * PHI
        re1p = PSp[..., I_RHOGE] + (getot_p + dtot_p) * dt + (k10p - k11p) + (PSp[..., I_RHOG] - rs1p) * PHIp
        return (rs1, rw1, re1), (rs1p, rw1p, re1p)

    def _rhow_solve(self, rhogw, rhogw0, preg0, rhog0, Srho, Sw, Spre, Mc, Mu, Ml, dt, tag, kax):
        """Source assembly + Thomas sweep per column (kernels/virhowsolver.py:67-120)."""
        ctx, c, m = self.ctx, self.ctx.c, self.ctx.m
        kmin, kmax = ctx.kmin, ctx.kmax
        ks, ksm = _ks(kax, kmin + 1, kmax + 1), _ks(kax, kmin, kmax)
        shp = [1] * rhogw.ndim
        shp[kax] = kmax - kmin
        CVovRt2 = c.CVdry / c.Rdry / (dt * dt)
        rdgzh = ctx.rdgzh[kmin + 1:kmax + 1].reshape(shp)
        af = ctx.afact[kmin + 1:kmax + 1].reshape(shp)
        bf = ctx.bfact[kmin + 1:kmax + 1].reshape(shp)
        RGAMH, RGSGAM2, RGAM = m["RGAMH" + tag], m["RGSGAM2" + tag], m["RGAM" + tag]
        RGSGAM2H, GSGAM2H = m["RGSGAM2H" + tag], m["GSGAM2H" + tag]
        Sall = ((rhogw0[ks] * c.ALPHA + dt * Sw[ks]) * RGAMH[ks] ** 2
                - ((preg0[ks] + dt * Spre[ks]) * RGSGAM2[ks]
                   - (preg0[ksm] + dt * Spre[ksm]) * RGSGAM2[ksm]) * dt * rdgzh
                - ((rhog0[ks] + dt * Srho[ks]) * RGAM[ks] ** 2 * af
                   + (rhog0[ksm] + dt * Srho[ksm]) * RGAM[ksm] ** 2 * bf) * dt * c.GRAV) * CVovRt2

        def K(k):
            return (slice(None),) * kax + (k,)
        rw_kmin = rhogw[K(kmin)] * RGSGAM2H[K(kmin)]
        rw_top = rhogw[K(kmax + 1)] * RGSGAM2H[K(kmax + 1)]
        N = kmax - kmin
        S = [Sall[K(i)] for i in range(N)]
        Mcl = [Mc[K(i)] for i in range(N)]
        Mul = [Mu[K(i)] for i in range(N)]
        Mll = [Ml[K(i)] for i in range(N)]
        S[0] = S[0] - Mll[0] * rw_kmin
        S[N - 1] = S[N - 1] - Mul[N - 1] * rw_top
        x = thomas(S, Mcl, Mul, Mll)
        out = rhogw.copy()
        for i in range(N):
            out[K(kmin + 1 + i)] = x[i] * GSGAM2H[K(kmin + 1 + i)]
        out[K(kmin)] = rw_kmin * GSGAM2H[K(kmin)]
        out[K(kmax + 1)] = rw_top * GSGAM2H[K(kmax + 1)]
        return out

    # ------------------------------------------------------------------
    def step(self):
        """One large step: RK3 stages with 2/3/6 small steps (mod_dynamics.py:1914-2727)."""
        ctx, c = self.ctx, self.ctx.c
        dtl = c.DTL
        nsub = [c.SSTEP_MAX / 3, c.SSTEP_MAX / 2, c.SSTEP_MAX]
        dts = dtl / c.SSTEP_MAX
        PROG0 = self.PROG.copy()
        PROG0_p = self.PROG_pl.copy()
        rhog_in, rhog_in_p = self.PROG[..., I_RHOG].copy(), self.PROG_pl[..., I_RHOG].copy()
        f_rhog = f_rhog_p = None
        PM = PMp = None
        for nl in range(3):
            pp = self.prepost()
            adv, advp = self.advection_momentum()
            g_TEND = np.zeros(self.PROG.shape)
            g_TEND_p = np.zeros(self.PROG_pl.shape)
            for iv, a in zip((I_RHOGVX, I_RHOGVY, I_RHOGVZ, I_RHOGW), adv):
                g_TEND[..., iv] = a
            for iv, a in zip((I_RHOGVX, I_RHOGVY, I_RHOGVZ, I_RHOGW), advp):
                g_TEND_p[..., iv] = a
            fT, fTp = self.hdiffusion(pp)
            g_TEND += fT
            g_TEND_p += fTp
            if nl == 0:
                PS = np.zeros(self.PROG.shape)
                PSp = np.zeros(self.PROG_pl.shape)
            else:
                PS = PROG0 - self.PROG
                PSp = PROG0_p - self.PROG_pl
            PM, PMp = self.vi_small_step(pp, g_TEND, g_TEND_p, PS, PSp, nsub[nl], dts)
            if nl < 2:
                ctx.halo(self.PROG, self.PROG_pl)
            f_rhog, f_rhog_p = fT[..., I_RHOG], fTp[..., I_RHOG]
        if self.tracer is not None:
            self.tracer(self, rhog_in, rhog_in_p, PM, PMp, f_rhog, f_rhog_p, dtl)
        # f_TENDq is zero for MIURA2004 transport: PROGq += dtl * 0
        PRG = np.concatenate([self.PROG, self.PROGq], axis=-1)
        PRGp = np.concatenate([self.PROG_pl, self.PROGq_pl], axis=-1)
        ctx.halo(PRG, PRGp)
        self.PROG[...] = PRG[..., :6]
        self.PROGq[...] = PRG[..., 6:]
        self.PROG_pl[...] = PRGp[..., :6]
        self.PROGq_pl[...] = PRGp[..., 6:]
