"""Horizontal icosahedral grid and vertical levels (host setup, FP64).

Host-side restatement of the reference's grid preparation: recursive
icosahedron subdivision (prep/hgrid/mod_mkgrd.py:66-193, 697-724), triangle
gravity centres (:568-611), the pole grid (share/mod_grd.py:764-938) and the
vertical factors (share/mod_grd.py:171-222).  Arrays use the reference's
(i, j, k, l, ...) layout; the HIP side transposes once at upload.
"""
from __future__ import annotations

import numpy as np

from .topology import (Topology, I_N, I_S, I_NPL, I_SPL, I_RGNID, VLINK, GALL_PL,
                       GSLF_PL, GMIN_PL, GMAX_PL)

TI, TJ = 0, 1
XDIR, YDIR, ZDIR = 0, 1, 2


def _subdivide(g0: np.ndarray) -> np.ndarray:
    """One bisection level: new points are normalised chord midpoints.

    Same construction as MKGRD decomposition (mod_mkgrd.py:697-724): the
    finer (2n-1)^2 lattice keeps the old points, adds sums of neighbours along
    i, j and the i=j diagonal, then projects everything to the unit sphere.
    """
    n0 = g0.shape[0]
    n1 = 2 * (n0 - 1) + 1
    g1 = np.zeros((n1, n1, 3))
    g1[0::2, 0::2] = g0
    g1[1::2, 0::2] = g0[1:, :] + g0[:-1, :]
    g1[0::2, 1::2] = g0[:, 1:] + g0[:, :-1]
    g1[1::2, 1::2] = g0[1:, 1:] + g0[:-1, :-1]
    r = np.sqrt(g1[..., 0] ** 2 + g1[..., 1] ** 2 + g1[..., 2] ** 2)
    return g1 / r[..., None]


def _diamond_corners(dmd: int) -> np.ndarray:
    """Corner points (2, 2, 3) of icosahedral diamond ``dmd`` (mod_mkgrd.py:76-124)."""
    alpha2 = 2.0 * np.pi / 5.0
    phi = np.arcsin(np.cos(alpha2) / (1.0 - np.cos(alpha2)))
    r0 = np.zeros((2, 2, 3))

    def pt(lat, lon):
        return np.array([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])

    if dmd <= 4:
        d = float(dmd)
        r0[0, 0] = pt(phi, alpha2 * d)
        r0[1, 0] = pt(-phi, alpha2 * (d + 0.5))
        r0[0, 1] = [0.0, 0.0, 1.0]
        r0[1, 1] = pt(phi, alpha2 * (d + 1.0))
    else:
        d = float(dmd - 5)
        r0[0, 0] = pt(-phi, -alpha2 * (d + 0.5))
        r0[1, 0] = [0.0, 0.0, -1.0]
        r0[0, 1] = pt(phi, -alpha2 * d)
        r0[1, 1] = pt(-phi, -alpha2 * (d - 0.5))
    return r0


def _arc(u, v):
    """Great-circle arc vector cross(u,v)/|cross(u,v)| * angle(u,v)."""
    c = np.cross(u, v)
    d = np.sum(u * v, axis=-1)
    s = np.sqrt(np.sum(c * c, axis=-1))
    with np.errstate(invalid="ignore", divide="ignore"):
        return c * (np.arctan2(s, d) / s)[..., None]


def _normalize(a):
    with np.errstate(invalid="ignore", divide="ignore"):
        return a / np.sqrt(np.sum(a * a, axis=-1))[..., None]


class HGrid:
    """Unit-sphere grid: x (g, g, 1, lall, 3), xt (g, g, 1, lall, 2, 3) and pole arrays."""

    def __init__(self, topo: Topology, x, xt, x_pl, xt_pl):
        self.topo = topo
        self.x, self.xt, self.x_pl, self.xt_pl = x, xt, x_pl, xt_pl


def generate_icosahedral(topo: Topology) -> HGrid:
    """Standard (unrelaxed) icosahedral grid, halos filled from the gather map."""
    g = topo.gall_1d
    rl = topo.rlevel
    nreg_1d = 2 ** rl
    x = np.zeros((g, g, 1, topo.lall, 3))
    x_pl = np.zeros((GALL_PL, 1, 2, 3))
    diamond = {}
    for rgn in range(topo.lall):
        dmd = rgn // (nreg_1d * nreg_1d)
        if dmd not in diamond:
            r0 = _diamond_corners(dmd)
            for _ in range(rl):
                r0 = _subdivide(r0)
            diamond[dmd] = r0
        r0 = diamond[dmd]
        loc = rgn % (nreg_1d * nreg_1d)
        ir, jr = loc % nreg_1d, loc // nreg_1d
        g0 = r0[ir:ir + 2, jr:jr + 2].copy()
        for _ in range(rl, topo.glevel):
            g0 = _subdivide(g0)
        x[topo.gmin:topo.gmax + 1, topo.gmin:topo.gmax + 1, 0, rgn] = g0[:topo.nmax, :topo.nmax]
    x_pl[GSLF_PL, 0, I_NPL] = [0.0, 0.0, 1.0]
    x_pl[GSLF_PL, 0, I_SPL] = [0.0, 0.0, -1.0]
    topo.halo_fill_host(x, x_pl)
    xt = _center_to_vertex(topo, x)
    xt_pl = _pole_vertices(topo, xt, x_pl)
    return HGrid(topo, x, xt, x_pl, xt_pl)


def _center_to_vertex(topo: Topology, x: np.ndarray) -> np.ndarray:
    """Triangle gravity centres from cell centres (mod_mkgrd.py:568-599)."""
    g = topo.gall_1d
    gmin, gmax = topo.gmin, topo.gmax
    xt = np.full((g, g, 1, topo.lall, 2, 3), -999.0)
    sl = slice(gmin - 1, gmax + 1)
    sp = slice(gmin, gmax + 2)
    a = x[sl, sl, 0]
    b = x[sp, sl, 0]
    c = x[sp, sp, 0]
    d = x[sl, sp, 0]
    xt[sl, sl, 0, :, TI] = _normalize(_arc(a, b) + _arc(b, c) + _arc(c, a))
    xt[sl, sl, 0, :, TJ] = _normalize(_arc(a, c) + _arc(c, d) + _arc(d, a))
    # the two outer halo triangles that degenerate take their twin's value
    xt[gmax, gmin - 1, 0, :, TI] = xt[gmax, gmin - 1, 0, :, TJ]
    xt[gmin - 1, gmax, 0, :, TJ] = xt[gmin - 1, gmax, 0, :, TI]
    sgp = np.nonzero(topo.have_sgp)[0]
    xt[gmin - 1, gmin - 1, 0, sgp, TI] = xt[gmin, gmin - 1, 0, sgp, TJ]
    return xt


def _pole_vertices(topo: Topology, xt: np.ndarray, x_pl: np.ndarray) -> np.ndarray:
    """Pole triangle centres gathered from the surrounding regions.

    Restates GRD_gen_plgrid (share/mod_grd.py:764-938): the north pole takes
    each neighbour's TJ centre at (gmin, gmax) in walk order rotated by one,
    the south pole each neighbour's TI centre at (gmax, gmin).
    """
    gmin, gmax = topo.gmin, topo.gmax
    xt_pl = np.zeros((GALL_PL, 1, 2, 3))
    vn, vt = topo.vert_num, topo.vert_tab
    ln = int(np.nonzero(vn[I_N] == VLINK)[0][-1])
    ls = int(np.nonzero(vn[I_S] == VLINK)[0][-1])
    for n in range(VLINK):
        rn = vt[I_RGNID, I_N, ln, (n + 1) % VLINK]
        xt_pl[n + 1, 0, I_NPL] = xt[gmin, gmax, 0, rn, TJ]
        rs = vt[I_RGNID, I_S, ls, n]
        xt_pl[n + 1, 0, I_SPL] = xt[gmax, gmin, 0, rs, TI]
    xt_pl[GSLF_PL] = x_pl[GSLF_PL]
    return xt_pl


def from_arrays(topo: Topology, x, xt, x_pl=None) -> HGrid:
    """Grid from stored centre / triangle arrays (e.g. a reference grid file).

    Pole centres are the exact poles; pole triangle centres follow the
    GRD_gen_plgrid gathering; halos of x are refilled from the gather map.
    """
    x = np.array(x, dtype=np.float64)
    xt = np.array(xt, dtype=np.float64)
    if x_pl is None:
        x_pl = np.zeros((GALL_PL, 1, 2, 3))
        x_pl[GSLF_PL, 0, I_NPL] = [0.0, 0.0, 1.0]
        x_pl[GSLF_PL, 0, I_SPL] = [0.0, 0.0, -1.0]
    topo.halo_fill_host(x, x_pl)
    xt_pl = _pole_vertices(topo, xt, x_pl)
    return HGrid(topo, x, xt, x_pl, xt_pl)


# ----------------------------------------------------------------------
# vertical grid
# ----------------------------------------------------------------------
def stretched_interfaces(nlayer: int, ztop: float, dz_bottom: float) -> np.ndarray:
    """Interface heights 0..ztop with layer thickness growing geometrically from dz_bottom."""
    if abs(nlayer * dz_bottom - ztop) < 1e-9 * ztop:
        return np.linspace(0.0, ztop, nlayer + 1)
    lo, hi = 1.0 + 1e-12, 2.0
    f = lambda r: dz_bottom * (r ** nlayer - 1.0) / (r - 1.0) - ztop
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    r = 0.5 * (lo + hi)
    dz = dz_bottom * r ** np.arange(nlayer)
    zh = np.concatenate([[0.0], np.cumsum(dz)])
    zh[-1] = ztop
    return zh


def vgrid_from_interfaces(zh_in: np.ndarray):
    """Full/half level heights including the two dummy levels.

    Same completion rule as the reference layer generator
    (prep/vgrid/mkvlayer.py:56-72): one mirrored level below the ground, full
    levels at layer midpoints, and one extra level above the top.
    Returns (gz, gzh) each of length nlayer + 2.
    """
    nl = len(zh_in) - 1
    kall, kmin, kmax = nl + 2, 1, nl
    zh = np.zeros(kall)
    zc = np.zeros(kall)
    zh[kmin:kmax + 2] = zh_in
    zh[kmin - 1] = zh[kmin] - (zh[kmin + 1] - zh[kmin])
    for k in range(kmin - 1, kmax + 1):
        zc[k] = zh[k] + 0.5 * (zh[k + 1] - zh[k])
    zc[kmax + 1] = zh[kmax + 1] + 0.5 * (zh[kmax + 1] - zh[kmax])
    return zc, zh


class VGrid:
    """Vertical factors of share/mod_grd.py:171-222 (gz, gzh, dgz, afact..dfact)."""

    def __init__(self, gz: np.ndarray, gzh: np.ndarray):
        kall = len(gz)
        kmin, kmax = 1, kall - 2
        self.gz, self.gzh = np.asarray(gz, float), np.asarray(gzh, float)
        gz, gzh = self.gz, self.gzh
        dgz = np.zeros(kall)
        dgzh = np.zeros(kall)
        for k in range(kmin - 1, kmax + 1):
            dgz[k] = gzh[k + 1] - gzh[k]
        dgz[kmax + 1] = dgz[kmax]
        for k in range(kmin, kmax + 2):
            dgzh[k] = gz[k] - gz[k - 1]
        dgzh[kmin - 1] = dgzh[kmin]
        self.dgz, self.dgzh = dgz, dgzh
        self.rdgz = 1.0 / dgz
        self.rdgzh = 1.0 / dgzh
        self.htop = gzh[kmax + 1] - gzh[kmin]
        afact = np.zeros(kall)
        for k in range(kmin, kmax + 2):
            afact[k] = (gzh[k] - gz[k - 1]) / (gz[k] - gz[k - 1])
        afact[kmin - 1] = 1.0
        cfact = np.zeros(kall)
        for k in range(kmin, kmax + 1):
            cfact[k] = (gz[k] - gzh[k]) / (gzh[k + 1] - gzh[k])
        cfact[kmin - 1] = 1.0
        cfact[kmax + 1] = 0.0
        self.afact, self.bfact = afact, 1.0 - afact
        self.cfact, self.dfact = cfact, 1.0 - cfact
