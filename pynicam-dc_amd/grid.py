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


def spring_relax(topo: Topology, hg: HGrid, beta: float = 1.15, dt: float = 2.0e-2, criteria: float = 1.0e-4,
                 dump_coef: float = 1.0, max_iter: int = 10000001):
    """Spring-dynamics smoothing of the grid points (MKGRD_spring, prep/hgrid/mod_mkgrd.py:196-303).

    Every point is tied to its six neighbours by springs of natural length
    beta * lambda (lambda = 2 pi / (10 * 2^(glevel-1)), the mean grid length of
    the unit sphere) and moves with damping on the sphere; the twelve pentagon
    points stay fixed.  All points advance together from the previous iterate
    (then the halos are exchanged) until the largest spring force, divided by
    lambda, falls below ``criteria``.  Returns (HGrid with the relaxed points and
    their triangle gravity centres, number of iterations).
    """
    g, L = topo.gall_1d, topo.lall
    gmin, gmax = topo.gmin, topo.gmax
    lam = 2.0 * np.pi / (10.0 * 2.0 ** (topo.glevel - 1))
    dbar = beta * lam
    # var: position (3), velocity (3), |force| / lambda, kinetic energy
    var = np.zeros((g, g, 1, L, 8))
    var_pl = np.zeros((GALL_PL, 1, 2, 8))
    var[..., 0:3] = hg.x
    var_pl[..., 0:3] = hg.x_pl
    sl = slice(gmin, gmax + 1)
    nb = [(1, 0), (1, 1), (0, 1), (-1, 0), (-1, -1), (0, -1)]
    sgp = np.nonzero(topo.have_sgp)[0]

    def cross(u, v):
        return np.stack([u[..., 1] * v[..., 2] - u[..., 2] * v[..., 1],
                         u[..., 2] * v[..., 0] - u[..., 0] * v[..., 2],
                         u[..., 0] * v[..., 1] - u[..., 1] * v[..., 0]], axis=-1)

    def dot(u, v):
        return u[..., 0] * v[..., 0] + u[..., 1] * v[..., 1] + u[..., 2] * v[..., 2]

    def vabs(u):
        return np.sqrt(u[..., 0] * u[..., 0] + u[..., 1] * u[..., 1] + u[..., 2] * u[..., 2])

    ite = 0
    for ite in range(max_iter):
        R = var[:, :, 0, :, 0:3]                                  # (g, g, l, 3)
        P0 = R[sl, sl]
        Pm = [R[gmin + di:gmax + 1 + di, gmin + dj:gmax + 1 + dj] for (di, dj) in nb]
        # pentagon cells: the sixth neighbour is the first
        Pm[5] = Pm[5].copy()
        Pm[5][0, 0, sgp] = Pm[0][0, 0, sgp]
        Fsum = None
        for m in range(6):
            P0Pm = cross(P0, Pm[m])
            P0PmP0 = cross(P0Pm, P0)
            length = vabs(P0PmP0)
            distance = np.arctan2(vabs(P0Pm), dot(P0, Pm[m]))
            F = (distance - dbar)[..., None] * P0PmP0 / length[..., None]
            if m == 5:
                F[0, 0, sgp] = 0.0
            Fsum = F if Fsum is None else Fsum + F
        W = var[sl, sl, 0, :, 3:6]
        fixed = var[gmin, gmin, 0, sgp, 0:3].copy()
        R0 = P0 + W * dt
        R0 = R0 / vabs(R0)[..., None]
        W0 = W + (Fsum - dump_coef * W) * dt
        E = dot(R0, W0)
        W0 = W0 - E[..., None] * R0
        var[sl, sl, 0, :, 0:3] = R0
        var[sl, sl, 0, :, 3:6] = W0
        var[sl, sl, 0, :, 6] = vabs(Fsum) / lam
        var[sl, sl, 0, :, 7] = 0.5 * dot(W0, W0)
        var[gmin, gmin, 0, sgp, :] = 0.0
        var[gmin, gmin, 0, sgp, 0:3] = fixed
        topo.halo_fill_host(var, var_pl)
        fmax = max(var[sl, sl, 0, :, 6].max(), var_pl[GSLF_PL, 0, :, 6].max())
        if fmax < criteria:
            break
    x = var[..., 0:3].copy()
    x_pl = var_pl[..., 0:3].copy()
    topo.halo_fill_host(x, x_pl)
    xt = _center_to_vertex(topo, x)
    xt_pl = _pole_vertices(topo, xt, x_pl)
    return HGrid(topo, x, xt, x_pl, xt_pl), ite + 1


def _arc_guarded(u, v, eps):
    """_arc with degenerate arcs (|cross| < eps) contributing zero (the zerosw guard)."""
    c = np.cross(u, v)
    d = np.sum(u * v, axis=-1)
    s = np.sqrt(np.sum(c * c, axis=-1))
    ang = np.arctan2(s, d)
    zerosw = 0.5 - np.copysign(0.5, np.abs(s) - eps)
    return c * ((1.0 - zerosw) / (s + zerosw) * ang)[..., None]


def gravity_centre(topo: Topology, hg: HGrid) -> HGrid:
    """Move every point to the gravity centre of its control volume (MKGRD_gravcenter,
    prep/hgrid/mod_mkgrd.py:568-667): triangle gravity centres from the points, then
    the points from the hexagons (pentagons) those centres span, the halos refilled.
    The triangle centres of the result are those of the points before the move, as
    the reference's grid files store them."""
    g, gmin, gmax = topo.gall_1d, topo.gmin, topo.gmax
    eps = float(np.finfo(np.float64).eps)
    xt = _center_to_vertex(topo, hg.x)
    xt_pl = _pole_vertices(topo, xt, hg.x_pl)
    x = hg.x.copy()
    x_pl = hg.x_pl.copy()
    sl, slm = slice(gmin, gmax + 1), slice(gmin - 1, gmax)
    T = xt[:, :, 0]
    w1 = T[sl, slm, :, TJ]
    w2 = T[sl, sl, :, TI]
    w3 = T[sl, sl, :, TJ]
    w4 = T[slm, sl, :, TI]
    w5 = T[slm, slm, :, TJ]
    w6 = T[slm, slm, :, TI].copy()
    sgp = np.nonzero(topo.have_sgp)[0]
    w6[0, 0, sgp] = w1[0, 0, sgp]
    gc = (_arc_guarded(w1, w2, eps) + _arc_guarded(w2, w3, eps) + _arc_guarded(w3, w4, eps)
          + _arc_guarded(w4, w5, eps) + _arc_guarded(w5, w6, eps) + _arc_guarded(w6, w1, eps))
    x[sl, sl, 0] = _normalize(gc)
    for lp in range(2):
        wk = [xt_pl[v, 0, lp] for v in range(GMIN_PL, GMAX_PL + 1)]
        wk.append(wk[0])
        gcp = np.zeros(3)
        for v in range(VLINK):
            gcp = gcp + _arc(wk[v], wk[v + 1])
        x_pl[GSLF_PL, 0, lp] = -_normalize(gcp)
    topo.halo_fill_host(x, x_pl)
    return HGrid(topo, x, xt, x_pl, xt_pl)


_HGRID_KEYS = ("grd_x_x", "grd_x_y", "grd_x_z", "grd_xt_ix", "grd_xt_jx", "grd_xt_iy", "grd_xt_jy", "grd_xt_iz",
               "grd_xt_jz")


def _hgrid_file(prefix: str, rank: int, io_mode: str) -> str:
    return f"{prefix}{rank:08d}.{io_mode}"


def write_hgrid(prefix: str, hg: HGrid, nfiles: int = 1, io_mode: str = "npz"):
    """The grid files GRD_input_hgrid reads (MKGRD output, prep/hgrid/mod_mkgrd.py:669-695):
    per process ``<prefix><rank, 8 digits>.<json|npz>``, flat arrays (regions of that
    process, then j, then i fastest) of the points and the two triangle centres."""
    topo = hg.topo
    L = topo.lall
    if L % nfiles:
        raise ValueError(f"{L} regions do not divide into {nfiles} files")
    per = L // nfiles
    x, xt = hg.x[:, :, 0], hg.xt[:, :, 0]
    arrs = [x[..., 0], x[..., 1], x[..., 2], xt[..., TI, 0], xt[..., TJ, 0], xt[..., TI, 1], xt[..., TJ, 1],
            xt[..., TI, 2], xt[..., TJ, 2]]
    paths = []
    for r in range(nfiles):
        sl = slice(r * per, (r + 1) * per)
        data = {k: np.ascontiguousarray(a[:, :, sl].transpose(2, 1, 0)).reshape(per, -1)
                for k, a in zip(_HGRID_KEYS, arrs)}
        path = _hgrid_file(prefix, r, io_mode)
        if io_mode == "npz":
            np.savez(path, **data)
        elif io_mode == "json":
            import json
            with open(path, "w") as f:
                json.dump({k: v.tolist() for k, v in data.items()}, f)
        else:
            raise ValueError(f"io_mode must be 'npz' or 'json', not {io_mode!r}")
        paths.append(path)
    return paths


def read_hgrid(prefix: str, topo: Topology, io_mode: str = "json", nfiles: int | None = None,
               lrvertex: bool = True) -> HGrid:
    """Read the reference's grid files (Grd.GRD_input_hgrid, share/mod_grd.py:444-503):
    ``<prefix><rank, 8 digits>.<json|npz>`` for rank 0..nfiles-1 (default: every file
    present), their regions concatenated in rank order (the regions of one process
    are consecutive, as RGNMNG_lp2r lays them out).  The pole points and the pole
    triangle centres are generated as GRD_gen_plgrid does; without the triangle
    centres (``lrvertex=False``) they are computed from the points."""
    g = topo.gall_1d
    if nfiles is None:
        nfiles = 0
        while __import__("os").path.exists(_hgrid_file(prefix, nfiles, io_mode)):
            nfiles += 1
        if nfiles == 0:
            raise FileNotFoundError(_hgrid_file(prefix, 0, io_mode))
    parts = {k: [] for k in _HGRID_KEYS}
    for r in range(nfiles):
        path = _hgrid_file(prefix, r, io_mode)
        if io_mode == "json":
            import json
            with open(path) as f:
                d = json.load(f)
        elif io_mode == "npz":
            d = np.load(path, allow_pickle=False)
        else:
            raise ValueError(f"io_mode must be 'npz' or 'json', not {io_mode!r}")
        for k in _HGRID_KEYS:
            if k in d:
                a = np.asarray(d[k], dtype=np.float64)
                if a.size % (g * g):
                    raise ValueError(f"{path}: {k} holds {a.size} values, not whole regions of {g} x {g} points")
                parts[k].append(a.reshape(-1, g * g))
    cat = {k: np.concatenate(v) for k, v in parts.items() if v}
    L = cat["grd_x_x"].shape[0]
    if L != topo.lall:
        raise ValueError(f"grid files hold {L} regions of {g} x {g} points, the topology {topo.lall}")

    def unpack(a):
        return a.reshape(L, g, g).transpose(2, 1, 0)          # (i, j, l)
    x = np.zeros((g, g, 1, L, 3))
    for d_, k in enumerate(("grd_x_x", "grd_x_y", "grd_x_z")):
        x[:, :, 0, :, d_] = unpack(cat[k])
    if lrvertex and all(k in cat for k in _HGRID_KEYS[3:]):
        xt = np.zeros((g, g, 1, L, 2, 3))
        for k, (t_, d_) in zip(_HGRID_KEYS[3:], ((TI, 0), (TJ, 0), (TI, 1), (TJ, 1), (TI, 2), (TJ, 2))):
            xt[:, :, 0, :, t_, d_] = unpack(cat[k])
    else:
        x_pl = np.zeros((GALL_PL, 1, 2, 3))
        x_pl[GSLF_PL, 0, I_NPL] = [0.0, 0.0, 1.0]
        x_pl[GSLF_PL, 0, I_SPL] = [0.0, 0.0, -1.0]
        topo.halo_fill_host(x, x_pl)
        xt = _center_to_vertex(topo, x)
    return from_arrays(topo, x, xt)


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
