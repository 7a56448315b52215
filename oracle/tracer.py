"""CPU oracle of the tracer advection (MIURA2004 with Thuburn limiters).

TEST INFRASTRUCTURE ONLY (imported by tests, the smoke check and the
benchmark's CPU baseline).  Restates the numpy path of
``Srctr.src_tracer_advection`` (nhm/dynamics/mod_src_tracer.py:83-1228):

  vertical half step    kernels/tracervertflux.py:62-167, mod_src_tracer.py:420-533,
                        vertical_limiter_thuburn (mod_src_tracer.py:1849-2113)
  horizontal step       horizontal_flux (:1230-1533), horizontal_remap (:1535-1847),
                        horizontal_limiter_thuburn (:2115-2662), update (:838-929)
  vertical half step    mod_src_tracer.py:947-1176, small-negative clip (:1190-1204)

with b1 = 0, b2 = 1, b3 = 0 (the hyperdiffusion mass tendency enters the
horizontal step only).  Regional arrays (i, j, k, l), pole arrays (v, k, lpl).
"""
from __future__ import annotations

import numpy as np

from .dynamics import gradient

TI, TJ = 0, 1
AI, AIJ, AJ = 0, 1, 2
XDIR, YDIR, ZDIR = 0, 1, 2
EPS = np.finfo(np.float64).eps
BIG = np.finfo(np.float64).max
B1, B2, B3 = 0.0, 1.0, 0.0


def _k(kax, lo, hi):
    return (slice(None),) * kax + (slice(lo, hi),)


# ----------------------------------------------------------------------
# vertical pieces
# ----------------------------------------------------------------------
def vert_flux(ctx, pm, rhog_in, frhog, dt, kax, GZ, RGAMH):
    """Vertical mass flux, Courant numbers, d and the half-step density (tracervertflux.py:44-122).

    pm: mean mass fluxes (..., 5) = rhog, rhogvx, rhogvy, rhogvz, rhogw; GZ the
    terrain-gradient metric C2WfactGz (..., 6) and RGAMH; RGSQRTH = 1 (flat terrain).
    """
    kmin, kmax = ctx.kmin, ctx.kmax
    shp = [1] * rhog_in.ndim
    shp[kax] = kmax + 1 - kmin
    rdgz = ctx.rdgz[kmin:kmax + 1].reshape(shp)
    flx = np.zeros_like(rhog_in)
    K, Km = _k(kax, kmin + 1, kmax + 1), _k(kax, kmin, kmax)
    vx, vy, vz, rw = pm[..., 1], pm[..., 2], pm[..., 3], pm[..., 4]
    horiz = (GZ[K + (..., 0)] * vx[K] + GZ[K + (..., 1)] * vx[Km] + GZ[K + (..., 2)] * vy[K]
             + GZ[K + (..., 3)] * vy[Km] + GZ[K + (..., 4)] * vz[K] + GZ[K + (..., 5)] * vz[Km]) * RGAMH[K]
    flx[K] = (horiz + rw[K] * 1.0) * 0.5 * dt
    kc, kcp = _k(kax, kmin, kmax + 1), _k(kax, kmin + 1, kmax + 2)
    ck = np.zeros(rhog_in.shape + (2,))
    ck[kc + (..., 0)] = -flx[kc] / rhog_in[kc] * rdgz
    ck[kc + (..., 1)] = flx[kcp] / rhog_in[kc] * rdgz
    d = B1 * frhog / rhog_in * dt
    rhog = np.zeros_like(rhog_in)
    rhog[kc] = rhog_in[kc] - (flx[kcp] - flx[kc]) * rdgz + B1 * frhog[kc] * dt
    rhog[_k(kax, kmin - 1, kmin)] = rhog_in[_k(kax, kmin, kmin + 1)]
    rhog[_k(kax, kmax + 1, kmax + 2)] = rhog_in[_k(kax, kmax, kmax + 1)]
    return flx, ck, d, rhog


def courant_v(ctx, flx, rhog, kax):
    """ck from the flux and a density (mod_src_tracer.py:951-973)."""
    kmin, kmax = ctx.kmin, ctx.kmax
    shp = [1] * rhog.ndim
    shp[kax] = kmax + 1 - kmin
    rdgz = ctx.rdgz[kmin:kmax + 1].reshape(shp)
    kc, kcp = _k(kax, kmin, kmax + 1), _k(kax, kmin + 1, kmax + 2)
    ck = np.zeros(rhog.shape + (2,))
    ck[kc + (..., 0)] = -flx[kc] / rhog[kc] * rdgz
    ck[kc + (..., 1)] = flx[kcp] / rhog[kc] * rdgz
    return ck


def vertical_limiter(ctx, q_h, q, d, ck, kax):
    """Thuburn (1996) limiter on face values (mod_src_tracer.py:1909-2074), in place on q_h."""
    kmin, kmax = ctx.kmin, ctx.kmax
    kc, kcm, kcp = _k(kax, kmin, kmax + 1), _k(kax, kmin - 1, kmax), _k(kax, kmin + 1, kmax + 2)
    inL = 0.5 - np.copysign(0.5, ck[kc + (..., 0)])
    inU = 0.5 + np.copysign(0.5, ck[kcp + (..., 0)])
    qc, qb, qa = q[kc], q[kcm], q[kcp]
    ck1, ck2 = ck[kc + (..., 0)], ck[kc + (..., 1)]
    minL = np.where(inL == 1.0, np.minimum(qc, qb), BIG)
    minU = np.where(inU == 1.0, np.minimum(qc, qa), BIG)
    maxL = np.where(inL == 1.0, np.maximum(qc, qb), -BIG)
    maxU = np.where(inU == 1.0, np.maximum(qc, qa), -BIG)
    nmin = np.minimum(np.minimum(minL, minU), qc)
    nmax = np.maximum(np.maximum(maxL, maxU), qc)
    Cin = inL * ck1 + inU * ck2
    Cout = (1.0 - inL) * ck1 + (1.0 - inU) * ck2
    CQmin = inL * ck1 * minL + inU * ck2 * minU
    CQmax = inL * ck1 * maxL + inU * ck2 * maxU
    zsw = 0.5 - np.copysign(0.5, np.abs(Cout) - EPS)
    dc = d[kc]
    Omin = ((qc - nmax) + nmax * (Cin + Cout - dc) - CQmax) / (Cout + zsw) * (1.0 - zsw) + qc * zsw
    Omax = ((qc - nmin) + nmin * (Cin + Cout - dc) - CQmin) / (Cout + zsw) * (1.0 - zsw) + qc * zsw
    n = kmax + 1 - kmin
    lo, hi = _k(kax, 0, n - 1), _k(kax, 1, n)
    ka = _k(kax, kmin + 1, kmax + 1)
    qh = q_h[ka]
    iL = inL[hi]
    c_km1 = np.minimum(np.maximum(qh, Omin[lo]), Omax[lo])
    c_k = np.minimum(np.maximum(qh, Omin[hi]), Omax[hi])
    q_h[ka] = iL * c_km1 + (1.0 - iL) * c_k


def vertical_step(ctx, rhogq_iq, rhog_den, flx, ck, d, kax, limiter):
    """q, face values, limiter and flux-form update of one tracer (mod_src_tracer.py:420-533)."""
    kmin, kmax = ctx.kmin, ctx.kmax
    q = rhogq_iq / rhog_den
    shp = [1] * q.ndim
    shp[kax] = kmax + 2 - kmin
    a = ctx.afact[kmin:kmax + 2].reshape(shp)
    b = ctx.bfact[kmin:kmax + 2].reshape(shp)
    q_h = np.zeros_like(q)
    q_h[_k(kax, kmin, kmax + 2)] = a * q[_k(kax, kmin, kmax + 2)] + b * q[_k(kax, kmin - 1, kmax + 1)]
    if limiter:
        vertical_limiter(ctx, q_h, q, d, ck, kax)
    q_h[_k(kax, kmin, kmin + 1)] = 0.0
    q_h[_k(kax, kmax + 1, kmax + 2)] = 0.0
    shp2 = [1] * q.ndim
    shp2[kax] = kmax + 1 - kmin
    rdgz = ctx.rdgz[kmin:kmax + 1].reshape(shp2)
    kc, kcp = _k(kax, kmin, kmax + 1), _k(kax, kmin + 1, kmax + 2)
    out = rhogq_iq.copy()
    out[kc] = rhogq_iq[kc] - (flx[kcp] * q_h[kcp] - flx[kc] * q_h[kc]) * rdgz
    out[_k(kax, kmin - 1, kmin)] = 0.0
    out[_k(kax, kmax + 1, kmax + 2)] = 0.0
    return out


# ----------------------------------------------------------------------
# horizontal pieces
# ----------------------------------------------------------------------
def horizontal_flux(ctx, rho, rvx, rvy, rvz, rho_pl, rvx_pl, rvy_pl, rvz_pl, dt):
    """Face mass fluxes and mass-centroid positions (mod_src_tracer.py:1376-1533)."""
    geo = ctx.geo
    t = geo.topo
    g, kall, L = t.gall_1d, t.kall, t.lall
    gt, ga = geo.gmtr_t, geo.gmtr_a                       # (g,g,l,2,5), (g,g,l,3,12)
    rarea = geo.gmtr_p[..., 1]                            # P_RAREA (g, g, l)
    xr, xr_pl = geo.xr, geo.xr_pl                         # GRD_xr (mod_grd.py:625-660)
    W1, W2, W3 = 2, 3, 4
    A_HN = 0
    flx = np.zeros((g, g, kall, L, 6))
    xc = np.zeros((g, g, kall, L, 3, 3))
    s, sp = slice(0, g - 1), slice(1, g)

    def w(T, c):
        return gt[s, s, :, T, c][:, :, None, :]

    def tri(f):
        ti = f[s, s] * w(TI, W1) + f[sp, s] * w(TI, W2)
        tj = f[s, s] * w(TJ, W1)
        ti = ti + f[sp, sp] * w(TI, W3)
        tj = tj + (f[sp, sp] * w(TJ, W2) + f[s, sp] * w(TJ, W3))
        return ti, tj

    tr = [tri(f) for f in (rho, rvx, rvy, rvz)]            # each (TI, TJ) on (g-1, g-1, k, l)
    m0, m1 = ctx.pnt_m0, ctx.pnt_m1
    for q in range(4):
        ti, tj = tr[q]
        ti[0, 0] = ti[0, 0] * m0 + tj[1, 0] * m1

    def hn(arc, si, sj):
        return [ga[si, sj, :, arc, A_HN + d][:, :, None, :] for d in range(3)]

    def rar(si, sj):
        return rarea[si, sj][:, :, None, :]

    def arc(a_idx, b_idx, arcname, si, sj, face, face_opp, oi, oj):
        # a_idx/b_idx: (T, slice_i, slice_j) of the two triangles
        rs = tr[0][a_idx[0]][a_idx[1], a_idx[2]] + tr[0][b_idx[0]][b_idx[1], b_idx[2]]
        rr = 1.0 / np.maximum(rs, EPS)
        v2 = [tr[q][a_idx[0]][a_idx[1], a_idx[2]] + tr[q][b_idx[0]][b_idx[1], b_idx[2]] for q in (1, 2, 3)]
        h = hn(arcname, si, sj)
        fl = 0.5 * (v2[0] * h[0] + v2[1] * h[1] + v2[2] * h[2])
        flx[si, sj, :, :, face] = fl * rar(si, sj) * dt
        flx[oi, oj, :, :, face_opp] = -fl * rar(oi, oj) * dt
        for d in range(3):
            xc[si, sj, :, :, arcname, d] = xr[si, sj, :, arcname, d][:, :, None, :] - v2[d] * rr * dt * 0.5

    # AI arc at (i, j), i = 0..g-2, j = 1..g-2: TJ(i, j-1) + TI(i, j)
    arc((TJ, slice(0, g - 1), slice(0, g - 2)), (TI, slice(0, g - 1), slice(1, g - 1)), AI,
        slice(0, g - 1), slice(1, g - 1), 0, 3, slice(1, g), slice(1, g - 1))
    # AIJ arc at (i, j), i, j = 0..g-2: TI(i, j) + TJ(i, j)
    arc((TI, slice(0, g - 1), slice(0, g - 1)), (TJ, slice(0, g - 1), slice(0, g - 1)), AIJ,
        slice(0, g - 1), slice(0, g - 1), 1, 4, slice(1, g), slice(1, g))
    # AJ arc at (i, j), i = 1..g-2, j = 0..g-2: TJ(i, j) + TI(i-1, j)
    arc((TJ, slice(1, g - 1), slice(0, g - 1)), (TI, slice(0, g - 2), slice(0, g - 1)), AJ,
        slice(1, g - 1), slice(0, g - 1), 2, 5, slice(1, g - 1), slice(1, g))
    flx[1, 1, :, :, 5] *= m0
    # pole
    flx_pl = np.zeros((6, kall, 2))
    xc_pl = np.zeros((6, kall, 2, 3))
    tpl, apl = geo.gmtr_t_pl, geo.gmtr_a_pl                 # (6, 2, 5), (6, 2, 18)
    rarea_pl = geo.gmtr_p_pl[0, :, 1]                       # (2,)
    for l in range(2):
        rt = np.zeros((6, kall, 4))
        for v in range(1, 6):
            vp = 1 if v == 5 else v + 1
            for q, f in enumerate((rho_pl, rvx_pl, rvy_pl, rvz_pl)):
                rt[v, :, q] = f[0, :, l] * tpl[v, l, W1] + f[v, :, l] * tpl[v, l, W2] + f[vp, :, l] * tpl[v, l, W3]
        for v in range(1, 6):
            vm = 5 if v == 1 else v - 1
            rr = 1.0 / np.maximum(rt[vm, :, 0] + rt[v, :, 0], EPS)
            v2 = [rt[vm, :, q] + rt[v, :, q] for q in (1, 2, 3)]
            fl = 0.5 * (v2[0] * apl[v, l, A_HN] + v2[1] * apl[v, l, A_HN + 1] + v2[2] * apl[v, l, A_HN + 2])
            flx_pl[v, :, l] = fl * rarea_pl[l] * dt
            for d in range(3):
                xc_pl[v, :, l, d] = xr_pl[v, l, d] - v2[d] * rr * dt * 0.5
    return flx, xc, flx_pl, xc_pl


def horizontal_remap(ctx, q, q_pl, cmask, cmask_pl, xc, xc_pl):
    """Upwind face values from the centre values and gradients (mod_src_tracer.py:1568-1843)."""
    gq, gq_pl = gradient(ctx, q, q_pl)
    ctx.halo(gq, gq_pl)
    return remap_faces(ctx, q, q_pl, gq, gq_pl, cmask, cmask_pl, xc, xc_pl)


def remap_faces(ctx, q, q_pl, gq, gq_pl, cmask, cmask_pl, xc, xc_pl):
    """Face values from the centre values and their (halo-filled) gradients: the
    reconstruction part of horizontal_remap (mod_src_tracer.py:1640-1843)."""
    geo = ctx.geo
    g = geo.topo.gall_1d
    x = geo.GRD_x[:, :, 0][:, :, None]                    # (g, g, 1, l, 3)
    x_pl = geo.GRD_x_pl[:, 0]                              # (6, 2, 3)
    q_a = np.zeros(q.shape + (6,))

    def ext(ci, cj, xi, xj, arcname):
        return (q[ci, cj] + gq[ci, cj, ..., XDIR] * (xc[xi, xj, ..., arcname, XDIR] - x[ci, cj, ..., XDIR])
                + gq[ci, cj, ..., YDIR] * (xc[xi, xj, ..., arcname, YDIR] - x[ci, cj, ..., YDIR])
                + gq[ci, cj, ..., ZDIR] * (xc[xi, xj, ..., arcname, ZDIR] - x[ci, cj, ..., ZDIR]))

    s, sp = slice(0, g - 1), slice(1, g)
    ss, sm = slice(1, g - 1), slice(0, g - 2)
    ap = [None] * 6
    am = [None] * 6
    ap[0], am[0] = ext(s, s, s, s, AI), ext(sp, s, s, s, AI)
    ap[1], am[1] = ext(s, s, s, s, AIJ), ext(sp, sp, s, s, AIJ)
    ap[2], am[2] = ext(s, s, s, s, AJ), ext(s, sp, s, s, AJ)
    for m in range(3):
        q_a[s, s, ..., m] = cmask[s, s, ..., m] * am[m] + (1.0 - cmask[s, s, ..., m]) * ap[m]
    z = np.zeros(q[s, s].shape)
    ap4_full, am4_full = z.copy(), z.copy()
    ap4_full[1:, 1:] = ext(sm, ss, sm, ss, AI)
    am4_full[1:, 1:] = ext(ss, ss, sm, ss, AI)
    ap5_full, am5_full = z.copy(), z.copy()
    ap5_full[1:, 1:] = ext(sm, sm, sm, sm, AIJ)
    am5_full[1:, 1:] = ext(ss, ss, sm, sm, AIJ)
    ap6_full, am6_full = z.copy(), z.copy()
    ap6_full[1:, 1:] = ext(ss, sm, ss, sm, AJ)
    am6_full[1:, 1:] = ext(ss, ss, ss, sm, AJ)
    for m, (a_p, a_m) in zip((3, 4, 5), ((ap4_full, am4_full), (ap5_full, am5_full), (ap6_full, am6_full))):
        q_a[s, s, ..., m] = cmask[s, s, ..., m] * a_m + (1.0 - cmask[s, s, ..., m]) * a_p
    q_a_pl = np.zeros(q_pl.shape)
    for v in range(1, 6):
        qap = (q_pl[0] + gq_pl[0, ..., XDIR] * (xc_pl[v, ..., XDIR] - x_pl[0, :, XDIR])
               + gq_pl[0, ..., YDIR] * (xc_pl[v, ..., YDIR] - x_pl[0, :, YDIR])
               + gq_pl[0, ..., ZDIR] * (xc_pl[v, ..., ZDIR] - x_pl[0, :, ZDIR]))
        qam = (q_pl[v] + gq_pl[v, ..., XDIR] * (xc_pl[v, ..., XDIR] - x_pl[v, :, XDIR])
               + gq_pl[v, ..., YDIR] * (xc_pl[v, ..., YDIR] - x_pl[v, :, YDIR])
               + gq_pl[v, ..., ZDIR] * (xc_pl[v, ..., ZDIR] - x_pl[v, :, ZDIR]))
        q_a_pl[v] = cmask_pl[v] * qam + (1.0 - cmask_pl[v]) * qap
    return q_a, q_a_pl


def horizontal_limiter(ctx, q_a, q_a_pl, q, q_pl, d, d_pl, ch, ch_pl, cmask, cmask_pl):
    """Thuburn limiter on the face values (mod_src_tracer.py:2190-2660), in place on q_a/q_a_pl."""
    t = ctx.topo
    g = t.gall_1d
    Qin = np.full(q.shape + (2, 6), -9.9999e30)
    G = g - 1
    for i in range(G):
        for j in range(G):
            q0 = q[i, j]
            q1 = q[i, j - 1] if j > 0 else q[0, 0]
            q2 = q[i + 1, j]
            q3 = q[i + 1, j + 1]
            qj1 = q[i, j + 1]
            q4 = q[i - 1, j] if i > 0 else q[0, 0]
            mins = (np.minimum(np.minimum(np.minimum(q0, q1), q2), q3),
                    np.minimum(np.minimum(np.minimum(q0, q2), q3), qj1),
                    np.minimum(np.minimum(np.minimum(q0, q3), qj1), q4))
            maxs = (np.maximum(np.maximum(np.maximum(q0, q1), q2), q3),
                    np.maximum(np.maximum(np.maximum(q0, q2), q3), qj1),
                    np.maximum(np.maximum(np.maximum(q0, q3), qj1), q4))
            for m in range(3):
                c = cmask[i, j, ..., m]
                Qin[i, j, ..., 0, m] = c * mins[m] + (1.0 - c) * BIG
                Qin[i, j, ..., 1, m] = c * maxs[m] + (1.0 - c) * (-BIG)
            for m, (ii, jj) in zip((3, 4, 5), ((i + 1, j), (i + 1, j + 1), (i, j + 1))):
                c = cmask[i, j, ..., m - 3]
                Qin[ii, jj, ..., 0, m] = c * BIG + (1.0 - c) * mins[m - 3]
                Qin[ii, jj, ..., 1, m] = c * (-BIG) + (1.0 - c) * maxs[m - 3]
    # pentagon corner: the (0, 0) AIJ face sees (0,0) (1,1) (2,1) (0,1)
    for l in np.nonzero(t.have_sgp)[0]:
        qs = [q[0, 0, :, l], q[1, 1, :, l], q[2, 1, :, l], q[0, 1, :, l]]
        mn = np.minimum(np.minimum(np.minimum(qs[0], qs[1]), qs[2]), qs[3])
        mx = np.maximum(np.maximum(np.maximum(qs[0], qs[1]), qs[2]), qs[3])
        c1 = cmask[0, 0, :, l, 1]
        Qin[0, 0, :, l, 0, 1] = np.where(c1 == 1.0, mn, BIG)
        Qin[1, 1, :, l, 0, 4] = np.where(c1 == 1.0, BIG, mn)
        Qin[0, 0, :, l, 1, 1] = np.where(c1 == 1.0, mx, -BIG)
        Qin[1, 1, :, l, 1, 4] = np.where(c1 == 1.0, -BIG, mx)
    s = slice(1, g - 1)
    Qout = np.zeros(q.shape + (2,))
    nmin = q[s, s]
    nmax = q[s, s]
    for m in range(6):
        nmin = np.minimum(nmin, Qin[s, s, ..., 0, m])
        nmax = np.maximum(nmax, Qin[s, s, ..., 1, m])
    chs = ch[s, s]
    chm = np.minimum(chs, 0.0)
    Cin = _seqsum([chm[..., m] for m in range(6)])
    Cout = _seqsum([chs[..., m] - chm[..., m] for m in range(6)])
    CQmin = _seqsum([chm[..., m] * Qin[s, s, ..., 0, m] for m in range(6)])
    CQmax = _seqsum([chm[..., m] * Qin[s, s, ..., 1, m] for m in range(6)])
    zsw = 0.5 - np.copysign(0.5, np.abs(Cout) - EPS)
    qq, dd = q[s, s], d[s, s]
    # a face whose Courant number lies in (0, EPS) counts as inflow for its owner and as
    # outflow here; its +-BIG bound then overflows the quotient, as in the reference
    with np.errstate(over="ignore", invalid="ignore"):
        Qout[s, s, ..., 0] = (qq - CQmax - nmax * (1.0 - Cin - Cout + dd)) / (Cout + zsw) * (1.0 - zsw) + qq * zsw
        Qout[s, s, ..., 1] = (qq - CQmin - nmin * (1.0 - Cin - Cout + dd)) / (Cout + zsw) * (1.0 - zsw) + qq * zsw
    for e in (0, g - 1):
        Qout[:, e, ..., 0] = q[:, e]
        Qout[:, e, ..., 1] = q[:, e]
        Qout[e, 1:g - 1, ..., 0] = q[e, 1:g - 1]
        Qout[e, 1:g - 1, ..., 1] = q[e, 1:g - 1]
    # pole
    Qin_pl = np.zeros(q_pl.shape + (2, 2))
    Qout_pl = np.zeros(q_pl.shape + (2,))
    for v in range(1, 6):
        vp = 1 if v == 5 else v + 1
        vm = 5 if v == 1 else v - 1
        mn = np.minimum(np.minimum(np.minimum(q_pl[0], q_pl[v]), q_pl[vm]), q_pl[vp])
        mx = np.maximum(np.maximum(np.maximum(q_pl[0], q_pl[v]), q_pl[vm]), q_pl[vp])
        cm = cmask_pl[v]
        Qin_pl[v, ..., 0, 0] = np.where(cm == 1.0, mn, BIG)
        Qin_pl[v, ..., 0, 1] = np.where(cm == 1.0, BIG, mn)
        Qin_pl[v, ..., 1, 0] = np.where(cm == 1.0, mx, -BIG)
        Qin_pl[v, ..., 1, 1] = np.where(cm == 1.0, -BIG, mx)
    nmin_p = q_pl[0].copy()
    nmax_p = q_pl[0].copy()
    for v in range(1, 6):
        nmin_p = np.minimum(nmin_p, Qin_pl[v, ..., 0, 0])
        nmax_p = np.maximum(nmax_p, Qin_pl[v, ..., 1, 0])
    Cin_p = np.zeros(q_pl.shape[1:])
    Cout_p = np.zeros(q_pl.shape[1:])
    CQmin_p = np.zeros(q_pl.shape[1:])
    CQmax_p = np.zeros(q_pl.shape[1:])
    for v in range(1, 6):
        chm_v = cmask_pl[v] * ch_pl[v]
        Cin_p = Cin_p + chm_v
        Cout_p = Cout_p + (ch_pl[v] - chm_v)
        CQmin_p = CQmin_p + chm_v * Qin_pl[v, ..., 0, 0]
        CQmax_p = CQmax_p + chm_v * Qin_pl[v, ..., 1, 0]
    zsw = 0.5 - np.copysign(0.5, np.abs(Cout_p) - EPS)
    den = Cout_p + zsw
    fac = 1.0 - zsw
    q0, d0 = q_pl[0], d_pl[0]
    Qout_pl[0, ..., 0] = (q0 - CQmax_p - nmax_p * (1.0 - Cin_p - Cout_p + d0)) / den * fac + q0 * zsw
    Qout_pl[0, ..., 1] = (q0 - CQmin_p - nmin_p * (1.0 - Cin_p - Cout_p + d0)) / den * fac + q0 * zsw
    ctx.halo(Qout, Qout_pl)
    # apply
    G = g - 1
    a, p = slice(0, G), slice(1, g)
    for m, (oi, oj), mo in ((0, (p, a), 3), (1, (p, p), 4), (2, (a, p), 5)):
        c = cmask[a, a, ..., m]
        qa = q_a[a, a, ..., m]
        v1 = c * np.minimum(np.maximum(qa, Qin[a, a, ..., 0, m]), Qin[a, a, ..., 1, m]) \
            + (1.0 - c) * np.minimum(np.maximum(qa, Qin[oi, oj, ..., 0, mo]), Qin[oi, oj, ..., 1, mo])
        v2 = c * np.maximum(np.minimum(v1, Qout[oi, oj, ..., 1]), Qout[oi, oj, ..., 0]) \
            + (1.0 - c) * np.maximum(np.minimum(v1, Qout[a, a, ..., 1]), Qout[a, a, ..., 0])
        q_a[a, a, ..., m] = v2
        q_a[oi, oj, ..., mo] = v2
    for v in range(1, 6):
        cm = cmask_pl[v]
        qv = q_a_pl[v]
        q0 = np.minimum(np.maximum(qv, Qin_pl[v, ..., 0, 0]), Qin_pl[v, ..., 1, 0])
        q1 = np.minimum(np.maximum(qv, Qin_pl[v, ..., 0, 1]), Qin_pl[v, ..., 1, 1])
        qv = cm * q0 + (1.0 - cm) * q1
        q2 = np.maximum(np.minimum(qv, Qout_pl[v, ..., 1]), Qout_pl[v, ..., 0])
        q3 = np.maximum(np.minimum(qv, Qout_pl[0, ..., 1]), Qout_pl[0, ..., 0])
        q_a_pl[v] = cm * q2 + (1.0 - cm) * q3


def _seqsum(terms):
    """Left-to-right sum, as numpy's reduction over a short axis (starting from -0.0)."""
    acc = -0.0 + terms[0]
    for x in terms[1:]:
        acc = acc + x
    return acc


# ----------------------------------------------------------------------
# the whole transport step
# ----------------------------------------------------------------------
def tracer_advection(ctx, rhogq, rhogq_pl, rhog_in, rhog_in_pl, pm, pm_pl, frhog, frhog_pl, dt,
                     limiter=True):
    """src_tracer_advection, MIURA2004, in place on rhogq (g,g,kall,l,nq) / rhogq_pl (6,kall,2,nq).

    pm / pm_pl: the time-mean mass fluxes (g,g,kall,l,5) / (6,kall,2,5) in the order
    rhog, rhogvx, rhogvy, rhogvz, rhogw (PROG_mean of the last RK stage).
    """
    t = ctx.topo
    g = t.gall_1d
    kmin, kmax = ctx.kmin, ctx.kmax
    nq = rhogq.shape[-1]
    # ---- vertical half step
    m = ctx.m
    flx_v, ck, d, rhog = vert_flux(ctx, pm, rhog_in, frhog, dt, 2, m["C2WfactGz"], m["RGAMH"])
    flx_vp, ckp, dp, rhog_p = vert_flux(ctx, pm_pl, rhog_in_pl, frhog_pl, dt, 1, m["C2WfactGz_pl"], m["RGAMH_pl"])
    for iq in range(nq):
        rhogq[..., iq] = vertical_step(ctx, rhogq[..., iq], rhog_in, flx_v, ck, d, 2, limiter)
        rhogq_pl[..., iq] = vertical_step(ctx, rhogq_pl[..., iq], rhog_in_pl, flx_vp, ckp, dp, 1, limiter)
    # ---- horizontal step
    d = B2 * frhog / rhog * dt
    rvx, rvy, rvz = pm[..., 1] * 1.0, pm[..., 2] * 1.0, pm[..., 3] * 1.0
    rvxp, rvyp, rvzp = pm_pl[..., 1] * 1.0, pm_pl[..., 2] * 1.0, pm_pl[..., 3] * 1.0
    flx_h, xc, flx_hp, xc_p = horizontal_flux(ctx, pm[..., 0], rvx, rvy, rvz, pm_pl[..., 0], rvxp, rvyp, rvzp, dt)
    ch = flx_h / rhog[..., None]
    cmask = 0.5 - np.copysign(0.5, ch - EPS)
    ch_p = np.zeros_like(flx_hp)
    cmask_p = np.zeros_like(flx_hp)
    ch_p[1:6] = flx_hp[1:6] / rhog_p[0]
    cmask_p[1:6] = 0.5 - np.copysign(0.5, ch_p[1:6] - EPS)
    d_p = B2 * frhog_pl / rhog_p * dt
    s = slice(1, g - 1)
    for iq in range(nq):
        q = rhogq[..., iq] / rhog
        q_pl = rhogq_pl[..., iq] / rhog_p
        q_a, q_a_pl = horizontal_remap(ctx, q, q_pl, cmask, cmask_p, xc, xc_p)
        if limiter:
            horizontal_limiter(ctx, q_a, q_a_pl, q, q_pl, d, d_p, ch, ch_p, cmask, cmask_p)
        f, a = flx_h[s, s], q_a[s, s]
        rhogq[s, s, ..., iq] -= (f[..., 0] * a[..., 0] + f[..., 1] * a[..., 1] + f[..., 2] * a[..., 2]
                                 + f[..., 3] * a[..., 3] + f[..., 4] * a[..., 4] + f[..., 5] * a[..., 5])
        for v in range(1, 6):
            rhogq_pl[0, :, :, iq] -= flx_hp[v] * q_a_pl[v]
    f = flx_h[s, s]
    rhog[s, s] -= (f[..., 0] + f[..., 1] + f[..., 2] + f[..., 3] + f[..., 4] + f[..., 5]) - (B2 * frhog[s, s] * dt)
    for v in range(1, 6):
        rhog_p[0] -= flx_hp[v]
    rhog_p[0] += B2 * frhog_pl[0] * dt
    # ---- vertical half step
    d = B3 * frhog / rhog * dt
    ck = courant_v(ctx, flx_v, rhog, 2)
    d_p = B3 * frhog_pl / rhog_p * dt
    ckp = courant_v(ctx, flx_vp, rhog_p, 1)
    for iq in range(nq):
        rhogq[..., iq] = vertical_step(ctx, rhogq[..., iq], rhog, flx_v, ck, d, 2, limiter)
        rhogq_pl[..., iq] = vertical_step(ctx, rhogq_pl[..., iq], rhog_p, flx_vp, ckp, d_p, 1, limiter)
        r = rhogq[:, :, kmin:kmax + 1, :, iq]
        r[(r > -1.0e-10) & (r < 0.0)] = 0.0
        r = rhogq_pl[..., iq]
        r[(r > -1.0e-10) & (r < 0.0)] = 0.0
