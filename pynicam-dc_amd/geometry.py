"""Grid geometry, operator coefficients and vertical metrics (host setup, FP64).

Host restatement of GMTR_setup (share/mod_gmtr.py:65-776), the OPRT_*_setup
coefficient builders (share/mod_oprt.py:125-1379) and VMTR_setup
(share/mod_vmtr.py:30-401).  Everything here runs once on the host in FP64;
the dynamics step only reads the results.

Array layout (reference order without the single-level axis):
  regional   (i, j, l, ...)        pole   (v, lpl, ...)
Operator stencil order around cell (i, j) (oprtlaplacian.py:46-54):
  0:(i,j) 1:(i+1,j) 2:(i+1,j+1) 3:(i,j+1) 4:(i-1,j) 5:(i-1,j-1) 6:(i,j-1)

The divergence/gradient/Laplacian/interpolation coefficients are assembled
from one description of the hexagonal control volume: its six triangles (with
the stencil index of each triangle vertex) and its six arcs (each arc between
two consecutive triangles, with the orientation sign of its normal).  The
reference spells out each coefficient term by term; the sums here have the
same terms, including the pentagon cells (5 triangles / 5 arcs) and the two
pole polygons.
"""
from __future__ import annotations

import numpy as np

from .grid import HGrid, VGrid, TI, TJ, XDIR, YDIR, ZDIR
from .topology import Topology, GALL_PL, GSLF_PL, GMIN_PL, GMAX_PL

AI, AIJ, AJ = 0, 1, 2
# GMTR_p components
P_AREA, P_RAREA, P_IX, P_IY, P_IZ, P_JX, P_JY, P_JZ = range(8)
# GMTR_t components
T_AREA, T_RAREA, T_W1, T_W2, T_W3 = range(5)
# GMTR_a components (HN, HT, TN, TT each xyz; pole adds TN2, TT2)
A_HN, A_HT, A_TN, A_TT, A_TN2, A_TT2 = 0, 3, 6, 9, 12, 15


# ----------------------------------------------------------------------
# spherical primitives (share/mod_vector.py)
# ----------------------------------------------------------------------
def _angle(a, c):
    """Angle between vectors a and c seen from the origin (VECTR_angle with b = 0)."""
    cx = a[..., 1] * c[..., 2] - a[..., 2] * c[..., 1]
    cy = a[..., 2] * c[..., 0] - a[..., 0] * c[..., 2]
    cz = a[..., 0] * c[..., 1] - a[..., 1] * c[..., 0]
    s = np.sqrt(cx * cx + cy * cy + cz * cz)
    d = a[..., 0] * c[..., 0] + a[..., 1] * c[..., 1] + a[..., 2] * c[..., 2]
    return np.arctan2(s, d)


def sph_triangle_area(a, b, c, radius):
    """Spherical triangle area by L'Huilier's formula (VECTR_triangle, ON_SPHERE)."""
    l1 = _angle(a, b) * 0.5
    l2 = _angle(b, c) * 0.5
    l3 = _angle(c, a) * 0.5
    s = 0.5 * (l1 + l2 + l3)
    x = np.tan(s) * np.tan(s - l1) * np.tan(s - l2) * np.tan(s - l3)
    x = np.maximum(x, 0.0)
    return 4.0 * np.arctan(np.sqrt(x)) * radius * radius


def tn_vectors(v_from, v_to, radius):
    """Tangential and normal vectors of the great-circle arc from->to.

    Both scaled to the arc length (GMTR_TNvec, share/mod_gmtr.py:674-776).
    """
    vt = v_to - v_from
    cx = v_from[..., 1] * v_to[..., 2] - v_from[..., 2] * v_to[..., 1]
    cy = v_from[..., 2] * v_to[..., 0] - v_from[..., 0] * v_to[..., 2]
    cz = v_from[..., 0] * v_to[..., 1] - v_from[..., 1] * v_to[..., 0]
    nv = np.stack([cx, cy, cz], axis=-1)
    s = np.sqrt(cx * cx + cy * cy + cz * cz)
    d = np.sum(v_from * v_to, axis=-1)
    dist = np.arctan2(s, d) * radius
    with np.errstate(invalid="ignore", divide="ignore"):
        lt = np.sqrt(np.sum(vt * vt, axis=-1))
        vt = vt * (dist / lt)[..., None]
        vn = nv * (dist / s)[..., None]
    return vt, vn


def xyz_to_latlon(x, y, z):
    """Latitude/longitude as VECTR_xyz2latlon (share/mod_vector.py:182-206)."""
    eps = np.finfo(np.float64).eps
    length = np.sqrt(x * x + y * y + z * z)
    small = length < eps
    zr = z / np.where(small, 1.0, length)
    lat = np.arcsin(np.clip(zr, -1.0, 1.0))
    lh = np.sqrt(x * x + y * y)
    small_h = lh < eps
    xr = x / np.where(small_h, 1.0, lh)
    lon = np.arccos(np.clip(xr, -1.0, 1.0))
    lon = np.where(y < 0.0, -lon, lon)
    lon = np.where(small_h, 0.0, lon)
    lat = np.where(small, 0.0, lat)
    lon = np.where(small, 0.0, lon)
    return lat, lon


# ----------------------------------------------------------------------
# control-volume description
# ----------------------------------------------------------------------
# six triangles around (i, j), counter-clockwise: (di, dj, type) and the
# stencil index of their W1, W2, W3 vertices
_TRIS = [
    ((0, 0, TI), (0, 1, 2)),
    ((0, 0, TJ), (0, 2, 3)),
    ((-1, 0, TI), (4, 0, 3)),
    ((-1, -1, TJ), (5, 0, 4)),
    ((-1, -1, TI), (5, 6, 0)),
    ((0, -1, TJ), (6, 1, 0)),
]
# six hexagon arcs: (di, dj, arc type, sign) and the two adjacent triangles
_ARCS = [
    ((0, 0, AI, +1.0), (5, 0)),
    ((0, 0, AIJ, +1.0), (0, 1)),
    ((0, 0, AJ, +1.0), (1, 2)),
    ((-1, 0, AI, -1.0), (2, 3)),
    ((-1, -1, AIJ, -1.0), (3, 4)),
    ((0, -1, AJ, -1.0), (4, 5)),
]
# pentagon cells drop triangle 4 (TI at (i-1,j-1)) and arc 5; arc 4 then
# joins triangles 3 and 5
_ARCS_SGP = [
    ((0, 0, AI, +1.0), (5, 0)),
    ((0, 0, AIJ, +1.0), (0, 1)),
    ((0, 0, AJ, +1.0), (1, 2)),
    ((-1, 0, AI, -1.0), (2, 3)),
    ((-1, -1, AIJ, -1.0), (3, 5)),
]
# triangle edges along the loop W1->W2->W3->W1: (di, dj, arc type, sign)
# relative to the triangle's own (i, j)
_TRI_EDGES = {
    TI: [(0, 0, AI, +1.0), (1, 0, AJ, +1.0), (0, 0, AIJ, -1.0)],
    TJ: [(0, 0, AIJ, +1.0), (0, 1, AI, -1.0), (0, 0, AJ, -1.0)],
}


class Geometry:
    """GMTR + OPRT + VMTR for one grid, scaled to ``radius``."""

    def __init__(self, topo: Topology, hgrid: HGrid, vgrid: VGrid, radius: float, grav: float):
        self.topo, self.vgrid = topo, vgrid
        self.bsstate = None     # basic state (bsstate.BasicState) or None for NOBASE
        self.radius = radius
        self.grav = grav
        g = topo.gall_1d
        self.x = hgrid.x[:, :, 0] * radius            # (g, g, l, 3)
        self.xt = hgrid.xt[:, :, 0] * radius          # (g, g, l, 2, 3)
        self.x_pl = hgrid.x_pl[:, 0] * radius         # (6, 2, 3)
        self.xt_pl = hgrid.xt_pl[:, 0] * radius       # (6, 2, 3)
        # GRD_x as used by the dynamics (with the single-level axis restored)
        self.GRD_x = self.x[:, :, None]                # (g, g, 1, l, 3)
        self.GRD_x_pl = self.x_pl[:, None]             # (6, 1, 2, 3)
        self.lat, self.lon = xyz_to_latlon(self.x[..., 0], self.x[..., 1], self.x[..., 2])
        self.lat_pl, self.lon_pl = xyz_to_latlon(self.x_pl[..., 0], self.x_pl[..., 1], self.x_pl[..., 2])
        with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
            self._gmtr_p()
            self._gmtr_t()
            self._gmtr_a()
            self._oprt()
        self._arc_midpoints()
        self._vmtr()

    # ------------------------------------------------------------------
    def _arc_midpoints(self):
        """GRD_xr: chord midpoints of the hexagon arcs between triangle centres (mod_grd.py:625-660).

        AI (i, j >= 1): TJ(i, j-1)..TI(i, j); AIJ: TI(i, j)..TJ(i, j); AJ (i >= 1, j):
        TJ(i, j)..TI(i-1, j); pole arc v: triangle v-1 .. triangle v.  Used for the
        mass-centroid positions of the tracer transport.
        """
        t = self.topo
        g, gmax = t.gall_1d, t.gmax
        xt = self.xt
        xr = np.zeros((g, g, t.lall, 3, 3))
        i0, j1, jm = slice(0, gmax + 1), slice(1, gmax + 1), slice(0, gmax)
        xr[i0, j1, :, AI] = 0.5 * (xt[i0, jm, :, TJ] + xt[i0, j1, :, TI])
        s = slice(0, gmax + 1)
        xr[s, s, :, AIJ] = 0.5 * (xt[s, s, :, TI] + xt[s, s, :, TJ])
        i1, im, j0 = slice(1, gmax + 1), slice(0, gmax), slice(0, gmax + 1)
        xr[i1, j0, :, AJ] = 0.5 * (xt[i1, j0, :, TJ] + xt[im, j0, :, TI])
        xr_pl = np.zeros((GALL_PL, 2, 3))
        for l in range(2):
            for v in range(GMIN_PL, GMAX_PL + 1):
                vm = GMAX_PL if v == GMIN_PL else v - 1
                xr_pl[v, l] = 0.5 * (self.xt_pl[vm, l] + self.xt_pl[v, l])
        self.xr, self.xr_pl = xr, xr_pl

    def _gmtr_p(self):
        """Cell areas and local east/north unit vectors (GMTR_p_setup, mod_gmtr.py:185-335)."""
        t = self.topo
        g, gmin, gmax, R = t.gall_1d, t.gmin, t.gmax, self.radius
        p = np.zeros((g, g, t.lall, 8))
        ii = slice(gmin, gmax + 1)
        im = slice(gmin - 1, gmax)
        x, xt = self.x / R, self.xt / R
        ring = [xt[ii, im, :, TJ], xt[ii, ii, :, TI], xt[ii, ii, :, TJ], xt[im, ii, :, TI],
                xt[im, im, :, TJ], xt[im, im, :, TI]]
        ring = [r.copy() for r in ring]
        ring.append(ring[0].copy())
        sgp = np.nonzero(t.have_sgp)[0]
        ring[5][0, 0, sgp] = ring[0][0, 0, sgp]
        ring[6][0, 0, sgp] = ring[0][0, 0, sgp]
        area = 0.0
        c0 = x[ii, ii]
        for v in range(6):
            area = area + sph_triangle_area(c0, ring[v], ring[v + 1], R)
        p[ii, ii, :, P_AREA] = area
        p[ii, ii, :, P_RAREA] = 1.0 / area
        sl, cl = np.sin(self.lon[ii, ii]), np.cos(self.lon[ii, ii])
        p[ii, ii, :, P_IX] = -sl
        p[ii, ii, :, P_IY] = cl
        p[ii, ii, :, P_IZ] = 0.0
        p[ii, ii, :, P_JX] = -(self.x[ii, ii, :, ZDIR] * cl) / R
        p[ii, ii, :, P_JY] = -(self.x[ii, ii, :, ZDIR] * sl) / R
        p[ii, ii, :, P_JZ] = (self.x[ii, ii, :, XDIR] * cl + self.x[ii, ii, :, YDIR] * sl) / R

        ppl = np.zeros((GALL_PL, 2, 8))
        xpl, xtpl = self.x_pl / R, self.xt_pl / R
        for l in range(2):
            area = 0.0
            for v in range(1, 6):
                vp = v + 1 if v < 5 else 1
                area += sph_triangle_area(xpl[0, l], xtpl[v, l], xtpl[vp, l], R)
            ppl[0, l, P_AREA] = area
            ppl[0, l, P_RAREA] = 1.0 / area
            sl_, cl_ = np.sin(self.lon_pl[0, l]), np.cos(self.lon_pl[0, l])
            ppl[0, l, P_IX] = -sl_
            ppl[0, l, P_IY] = cl_
            ppl[0, l, P_JX] = -(self.x_pl[0, l, ZDIR] * cl_) / R
            ppl[0, l, P_JY] = -(self.x_pl[0, l, ZDIR] * sl_) / R
            ppl[0, l, P_JZ] = (self.x_pl[0, l, XDIR] * cl_ + self.x_pl[0, l, YDIR] * sl_) / R
        # halo values come from the neighbouring regions
        pf = p[:, :, None]                    # (g, g, 1, l, 8)
        pfl = ppl[:, None]                    # (6, 1, 2, 8)
        t.halo_fill_host(pf, pfl)
        self.gmtr_p, self.gmtr_p_pl = pf[:, :, 0], pfl[:, 0]
        self.area, self.area_pl = self.gmtr_p[..., P_AREA], self.gmtr_p_pl[..., P_AREA]

    def _gmtr_t(self):
        """Triangle areas and barycentric weights (GMTR_t_setup, mod_gmtr.py:337-452)."""
        t = self.topo
        g, gmin, gmax, R = t.gall_1d, t.gmin, t.gmax, self.radius
        x, xt = self.x / R, self.xt / R
        tt = np.zeros((g, g, t.lall, 2, 5))
        s = slice(gmin - 1, gmax + 1)
        sp = slice(gmin, gmax + 2)
        wk = np.zeros((4, g, g, t.lall, 2, 3))
        wk[0, s, s, :, TI] = xt[s, s, :, TI]
        wk[1, s, s, :, TI] = x[s, s]
        wk[2, s, s, :, TI] = x[sp, s]
        wk[3, s, s, :, TI] = x[sp, sp]
        wk[0, s, s, :, TJ] = xt[s, s, :, TJ]
        wk[1, s, s, :, TJ] = x[s, s]
        wk[2, s, s, :, TJ] = x[sp, sp]
        wk[3, s, s, :, TJ] = x[s, sp]
        wk[:, gmax, gmin - 1, :, TI] = wk[:, gmax, gmin - 1, :, TJ]
        wk[:, gmin - 1, gmax, :, TJ] = wk[:, gmin - 1, gmax, :, TI]
        sgp = np.nonzero(t.have_sgp)[0]
        wk[:, gmin - 1, gmin - 1, sgp, TI] = wk[:, gmin, gmin - 1, sgp, TJ]
        w = wk[:, s, s]
        a1 = sph_triangle_area(w[0], w[2], w[3], R)
        a2 = sph_triangle_area(w[0], w[3], w[1], R)
        a3 = sph_triangle_area(w[0], w[1], w[2], R)
        area = a1 + a2 + a3
        tt[s, s, :, :, T_AREA] = area
        tt[s, s, :, :, T_RAREA] = 1.0 / area
        tt[s, s, :, :, T_W1] = a1 / area
        tt[s, s, :, :, T_W2] = a2 / area
        tt[s, s, :, :, T_W3] = a3 / area
        self.gmtr_t = tt

        tpl = np.zeros((GALL_PL, 2, 5))
        xpl, xtpl = self.x_pl / R, self.xt_pl / R
        for l in range(2):
            for v in range(GMIN_PL, GMAX_PL + 1):
                vp = v + 1 if v < GMAX_PL else GMIN_PL
                w0, w1, w2, w3 = xtpl[v, l], xpl[GSLF_PL, l], xpl[v, l], xpl[vp, l]
                a1 = sph_triangle_area(w0, w2, w3, R)
                a2 = sph_triangle_area(w0, w3, w1, R)
                a3 = sph_triangle_area(w0, w1, w2, R)
                area = a1 + a2 + a3
                tpl[v, l] = [area, 1.0 / area, a1 / area, a2 / area, a3 / area]
        self.gmtr_t_pl = tpl

    def _gmtr_a(self):
        """Arc normal/tangent vectors (GMTR_a_setup, mod_gmtr.py:455-672)."""
        t = self.topo
        g, gmin, gmax, R = t.gall_1d, t.gmin, t.gmax, self.radius
        x, xt = self.x, self.xt
        a = np.zeros((g, g, t.lall, 3, 12))
        sgp = np.nonzero(t.have_sgp)[0]

        def put(arc, comp, si, sj, vf, vto):
            vt, vn = tn_vectors(vf, vto, R)
            a[si, sj, :, arc, comp:comp + 3] = vn
            a[si, sj, :, arc, comp + 3:comp + 6] = vt

        # arcs between cell centres (triangle edges): TN / TT
        si, sj = slice(gmin - 1, gmax + 1), slice(gmin - 1, gmax + 2)
        f = x[si, sj].copy()
        to = x[gmin:gmax + 2, sj].copy()
        f[gmax - (gmin - 1), 0] = x[gmax, gmin - 1]
        to[gmax - (gmin - 1), 0] = x[gmax, gmin]
        f[0, gmax + 1 - (gmin - 1)] = x[gmin, gmax + 1]
        to[0, gmax + 1 - (gmin - 1)] = x[gmin, gmax]
        f[0, 0, sgp] = x[gmin, gmin - 1, sgp]
        to[0, 0, sgp] = x[gmin + 1, gmin, sgp]
        put(AI, A_TN, si, sj, f, to)

        s = slice(gmin - 1, gmax + 1)
        put(AIJ, A_TN, s, s, x[s, s], x[gmin:gmax + 2, gmin:gmax + 2])

        si, sj = slice(gmin - 1, gmax + 2), slice(gmin - 1, gmax + 1)
        f = x[si, sj].copy()
        to = x[si, gmin:gmax + 2].copy()
        f[gmax + 1 - (gmin - 1), 0] = x[gmax + 1, gmin]
        to[gmax + 1 - (gmin - 1), 0] = x[gmax, gmin]
        f[0, gmax - (gmin - 1)] = x[gmin - 1, gmax]
        to[0, gmax - (gmin - 1)] = x[gmin, gmax]
        put(AJ, A_TN, si, sj, f, to)

        # arcs between triangle centres (hexagon edges): HN / HT
        si, sj = slice(gmin - 1, gmax + 1), slice(gmin, gmax + 1)
        put(AI, A_HN, si, sj, xt[si, sj, :, TI], xt[si, gmin - 1:gmax, :, TJ])

        s = slice(gmin - 1, gmax + 1)
        f = xt[s, s, :, TJ].copy()
        to = xt[s, s, :, TI].copy()
        f[gmax - (gmin - 1), 0] = xt[gmax, gmin - 1, :, TJ]
        to[gmax - (gmin - 1), 0] = xt[gmax, gmin, :, TI]
        f[0, gmax - (gmin - 1)] = xt[gmin, gmax, :, TJ]
        to[0, gmax - (gmin - 1)] = xt[gmin - 1, gmax, :, TI]
        put(AIJ, A_HN, s, s, f, to)

        si, sj = slice(gmin, gmax + 1), slice(gmin - 1, gmax + 1)
        f = xt[gmin - 1:gmax, sj, :, TI].copy()
        to = xt[si, sj, :, TJ].copy()
        f[0, 0, sgp] = xt[gmin, gmin, sgp, TI]
        to[0, 0, sgp] = xt[gmin, gmin - 1, sgp, TJ]
        put(AJ, A_HN, si, sj, f, to)
        self.gmtr_a = a

        apl = np.zeros((GALL_PL, 2, 18))
        for l in range(2):
            for v in range(GMIN_PL, GMAX_PL + 1):
                vp = v + 1 if v < GMAX_PL else GMIN_PL
                vm = v - 1 if v > GMIN_PL else GMAX_PL
                vt, vn = tn_vectors(self.x_pl[GSLF_PL, l], self.x_pl[v, l], R)
                apl[v, l, A_TN:A_TN + 3], apl[v, l, A_TT:A_TT + 3] = vn, vt
                vt, vn = tn_vectors(self.x_pl[v, l], self.x_pl[vp, l], R)
                apl[v, l, A_TN2:A_TN2 + 3], apl[v, l, A_TT2:A_TT2 + 3] = vn, vt
                vt, vn = tn_vectors(self.xt_pl[vm, l], self.xt_pl[v, l], R)
                apl[v, l, A_HN:A_HN + 3], apl[v, l, A_HT:A_HT + 3] = vn, vt
        self.gmtr_a_pl = apl

    # ------------------------------------------------------------------
    def _oprt(self):
        """Operator coefficients (OPRT_divergence/gradient/laplacian/diffusion_setup)."""
        t = self.topo
        g, gmin, gmax, L = t.gall_1d, t.gmin, t.gmax, t.lall
        P, T, A = self.gmtr_p, self.gmtr_t, self.gmtr_a
        ii = slice(gmin, gmax + 1)
        n = gmax - gmin + 1
        sgp = np.nonzero(t.have_sgp)[0]

        def Tsl(di, dj, typ, comp):
            return T[gmin + di:gmin + di + n, gmin + dj:gmin + dj + n, :, typ, comp]

        def Asl(di, dj, arc, comp):
            return A[gmin + di:gmin + di + n, gmin + dj:gmin + dj + n, :, arc, comp]

        rarea = P[ii, ii, :, P_RAREA]

        def assemble(arcs, cell_mask=None):
            """div-type coefficients (n, n, l, xyz, 7) for one arc/triangle layout."""
            c = np.zeros((n, n, L, 3, 7))
            for d in range(3):
                for (adi, adj, atyp, sgn), (ta, tb) in arcs:
                    hn = sgn * Asl(adi, adj, atyp, A_HN + d)
                    for tri in (ta, tb):
                        (tdi, tdj, ttyp), verts = _TRIS[tri]
                        for w, m in enumerate(verts):
                            c[..., d, m] += Tsl(tdi, tdj, ttyp, T_W1 + w) * hn
                c[..., d, :] *= (0.5 * rarea)[..., None]
            return c

        def arc_sum(arcs):
            s = np.zeros((n, n, L, 3))
            for d in range(3):
                for (adi, adj, atyp, sgn), _ in arcs:
                    s[..., d] += sgn * Asl(adi, adj, atyp, A_HN + d)
            return s * (0.5 * rarea)[..., None]

        div = assemble(_ARCS)
        div_sgp = assemble(_ARCS_SGP)
        grad = div.copy()
        grad[..., 0] -= 2.0 * arc_sum(_ARCS)
        grad_sgp = div_sgp.copy()
        grad_sgp[..., 0] -= 2.0 * arc_sum(_ARCS_SGP)
        div[0, 0, sgp] = div_sgp[0, 0, sgp]
        grad[0, 0, sgp] = grad_sgp[0, 0, sgp]
        self.coef_div = np.zeros((g, g, L, 3, 7))
        self.coef_div[ii, ii] = div
        self.coef_grad = np.zeros((g, g, L, 3, 7))
        self.coef_grad[ii, ii] = grad

        # Laplacian: triangle-wise gradient (edges along the triangle loop)
        # dotted with the two hexagon arcs bounding that triangle
        def lap_assemble(arcs, tris):
            c = np.zeros((n, n, L, 7))
            hsum = {k: np.zeros((n, n, L, 3)) for k in tris}
            for (adi, adj, atyp, sgn), (ta, tb) in arcs:
                for d in range(3):
                    h = sgn * Asl(adi, adj, atyp, A_HN + d)
                    hsum[ta][..., d] += h
                    hsum[tb][..., d] += h
            for tri in tris:
                (tdi, tdj, ttyp), verts = _TRIS[tri]
                edges = []
                for (edi, edj, etyp, esgn) in _TRI_EDGES[ttyp]:
                    edges.append(np.stack([esgn * Asl(tdi + edi, tdj + edj, etyp, A_TN + d)
                                           for d in range(3)], axis=-1))
                L1, L2, L3 = edges
                gvec = (2.0 * L2 - L1 - L3, 2.0 * L3 - L1 - L2, 2.0 * L1 - L2 - L3)
                tr = Tsl(tdi, tdj, ttyp, T_RAREA)
                for w, m in enumerate(verts):
                    c[..., m] += tr * np.sum(gvec[w] * hsum[tri], axis=-1)
            return c * (rarea / 12.0)[..., None]

        lap = lap_assemble(_ARCS, range(6))
        lap_sgp = lap_assemble(_ARCS_SGP, [0, 1, 2, 3, 5])
        lap[0, 0, sgp] = lap_sgp[0, 0, sgp]
        self.coef_lap = np.zeros((g, g, L, 7))
        self.coef_lap[ii, ii] = lap

        # vertex interpolation of a triangle-mean gradient (OPRT_diffusion_setup)
        s = slice(gmin - 1, gmax + 1)
        m1 = gmax - gmin + 2
        intp = np.zeros((g, g, L, 3, 3, 2))
        for ttyp in (TI, TJ):
            tr = T[s, s, :, ttyp, T_RAREA]
            edges = []
            for (edi, edj, etyp, esgn) in _TRI_EDGES[ttyp]:
                edges.append(esgn * A[gmin - 1 + edi:gmin - 1 + edi + m1,
                                      gmin - 1 + edj:gmin - 1 + edj + m1, :, etyp, A_TN:A_TN + 3])
            L1, L2, L3 = edges
            for w, e in enumerate((L1 + L3, L1 + L2, L2 + L3)):
                intp[s, s, :, :, w, ttyp] = -e * (0.5 * tr)[..., None]
        self.coef_intp = intp

        diff = np.zeros((g, g, L, 3, 6))
        order = [1, 2, 3, 4, 5, 0]   # coef_diff slot q uses hexagon arc order[q]
        for q, ai in enumerate(order):
            (adi, adj, atyp, sgn), _ = _ARCS[ai]
            for d in range(3):
                diff[ii, ii, :, d, q] = sgn * Asl(adi, adj, atyp, A_HN + d) * 0.5 * rarea
        diff[gmin, gmin, sgp, :, 4] = 0.0
        self.coef_diff = diff

        # ---- pole polygons ------------------------------------------------
        Ppl, Tpl, Apl = self.gmtr_p_pl, self.gmtr_t_pl, self.gmtr_a_pl
        div_pl = np.zeros((GALL_PL, 2, 3))
        grad_pl = np.zeros((GALL_PL, 2, 3))
        lap_pl = np.zeros((GALL_PL, 2))
        intp_pl = np.zeros((GALL_PL, 2, 3, 3))
        diff_pl = np.zeros((GALL_PL, 2, 3))
        vs = range(GMIN_PL, GMAX_PL + 1)

        def nxt(v):
            return v + 1 if v < GMAX_PL else GMIN_PL

        def prv(v):
            return v - 1 if v > GMIN_PL else GMAX_PL

        for l in range(2):
            ra = Ppl[GSLF_PL, l, P_RAREA]
            hn = {v: Apl[v, l, A_HN:A_HN + 3] for v in vs}
            tn = {v: Apl[v, l, A_TN:A_TN + 3] for v in vs}
            tn2 = {v: Apl[v, l, A_TN2:A_TN2 + 3] for v in vs}
            w1 = {v: Tpl[v, l, T_W1] for v in vs}
            w2 = {v: Tpl[v, l, T_W2] for v in vs}
            w3 = {v: Tpl[v, l, T_W3] for v in vs}
            tra = {v: Tpl[v, l, T_RAREA] for v in vs}
            # pole triangle v spans arcs v and v+1
            div_pl[GSLF_PL, l] = sum(w1[v] * (hn[v] + hn[nxt(v)]) for v in vs) * 0.5 * ra
            grad_pl[GSLF_PL, l] = sum(2.0 * (w1[v] - 1.0) * hn[nxt(v)] for v in vs) * 0.5 * ra
            for v in vs:
                c = (w3[prv(v)] * (hn[prv(v)] + hn[v]) + w2[v] * (hn[v] + hn[nxt(v)])) * 0.5 * ra
                div_pl[v, l] = c
                grad_pl[v, l] = c
            # Laplacian: the pole triangles run clockwise, so the bounding arcs
            # enter with the opposite sign
            lap = np.zeros(GALL_PL)
            for v in vs:
                e1, e2, e3 = tn[v], tn2[v], -tn[nxt(v)]
                gv = (2.0 * e2 - e1 - e3, 2.0 * e3 - e1 - e2, 2.0 * e1 - e2 - e3)
                h = -(hn[v] + hn[nxt(v)])
                for w, m in enumerate((GSLF_PL, v, nxt(v))):
                    lap[m] += tra[v] * np.dot(gv[w], h)
            lap_pl[:, l] = lap * ra / 12.0
            for v in vs:
                e1, e2, e3 = tn[v], tn2[v], -tn[nxt(v)]
                for w, e in enumerate((e1 + e3, e1 + e2, e2 + e3)):
                    intp_pl[v, l, :, w] = e * 0.5 * tra[v]
                diff_pl[v, l] = hn[v] * 0.5 * ra
        self.coef_div_pl, self.coef_grad_pl, self.coef_lap_pl = div_pl, grad_pl, lap_pl
        self.coef_intp_pl, self.coef_diff_pl = intp_pl, diff_pl

    # ------------------------------------------------------------------
    def gradient(self, scl, scl_pl):
        """Horizontal gradient (compute_oprt_gradient); scl (g,g,k,l), scl_pl (6,k,2)."""
        g = self.topo.gall_1d
        c = self.coef_grad[1:g - 1, 1:g - 1][:, :, None]          # (.., 1, l, 3, 7)
        st = stencil7(scl)                                       # (.., k, l, 7)
        inner = np.stack([np.sum(c[..., d, :] * st, axis=-1) for d in range(3)], axis=-1)
        out = np.zeros(scl.shape + (3,))
        out[1:g - 1, 1:g - 1] = inner
        out_pl = np.zeros(scl_pl.shape + (3,))
        cp = self.coef_grad_pl                                   # (6, 2, 3)
        out_pl[GSLF_PL] = np.sum(cp[:, None, :, :] * scl_pl[..., None], axis=0)
        return out, out_pl

    def horizontalize(self, vx, vy, vz, vx_pl, vy_pl, vz_pl):
        """Remove the radial component (OPRT_horizontalize_vec), interior + pole."""
        g, R = self.topo.gall_1d, self.radius
        s = slice(1, g - 1)
        gx = self.x[s, s, None, :, XDIR]
        gy = self.x[s, s, None, :, YDIR]
        gz = self.x[s, s, None, :, ZDIR]
        prd = (vx[s, s] * gx + vy[s, s] * gy + vz[s, s] * gz) / R
        vx[s, s] = vx[s, s] - prd * gx / R
        vy[s, s] = vy[s, s] - prd * gy / R
        vz[s, s] = vz[s, s] - prd * gz / R
        px = self.x_pl[:, None, :, XDIR]
        py = self.x_pl[:, None, :, YDIR]
        pz = self.x_pl[:, None, :, ZDIR]
        prd = vx_pl * px / R + vy_pl * py / R + vz_pl * pz / R
        vx_pl[...] = vx_pl - prd * px / R
        vy_pl[...] = vy_pl - prd * py / R
        vz_pl[...] = vz_pl - prd * pz / R

    # ------------------------------------------------------------------
    def _vmtr(self):
        """Vertical metrics for terrain-following height (VMTR_setup, mod_vmtr.py:30-401).

        Flat topography: z(i,j,k) = gz[k], zh = gzh[k] at every column.
        """
        t, vg = self.topo, self.vgrid
        g, kall, L = t.gall_1d, t.kall, t.lall
        kmin, kmax = t.kmin, t.kmax
        vz = np.zeros((g, g, kall, L, 2))
        vz_pl = np.zeros((GALL_PL, kall, 2, 2))
        ii = slice(t.gmin, t.gmax + 1)
        vz[ii, ii, :, :, 0] = vg.gz[None, None, :, None]
        vz[ii, ii, :, :, 1] = vg.gzh[None, None, :, None]
        vz_pl[GSLF_PL, :, :, 0] = vg.gz[:, None]
        vz_pl[GSLF_PL, :, :, 1] = vg.gzh[:, None]
        t.halo_fill_host(vz, vz_pl)
        self.vz, self.vz_pl = vz, vz_pl

        var = np.zeros((g, g, kall, L, 6))
        var_pl = np.zeros((GALL_PL, kall, 2, 6))
        for base, comp in ((0, 1), (3, 0)):      # JXH..JZH from zh, JX..JZ from z
            gr, gr_pl = self.gradient(vz[..., comp], vz_pl[..., comp])
            self.horizontalize(gr[..., 0], gr[..., 1], gr[..., 2],
                               gr_pl[..., 0], gr_pl[..., 1], gr_pl[..., 2])
            var[..., base:base + 3] = gr
            var_pl[..., base:base + 3] = gr_pl
        t.halo_fill_host(var, var_pl)

        def metrics(vzz, shape_k_axis):
            zc, zh = vzz[..., 0], vzz[..., 1]
            gs = np.full(zc.shape, np.nan)
            gsh = np.full(zc.shape, np.nan)
            with np.errstate(invalid="ignore", divide="ignore"):
                for k in range(kmin, kmax + 1):
                    _set_k(gs, k, (_get_k(zh, k + 1) - _get_k(zh, k)) / vg.dgz[k], shape_k_axis)
                _set_k(gs, kmin - 1, _get_k(gs, kmin), shape_k_axis)
                _set_k(gs, kmax + 1, _get_k(gs, kmax), shape_k_axis)
                for k in range(kmin, kmax + 2):
                    _set_k(gsh, k, (_get_k(zc, k) - _get_k(zc, k - 1)) / vg.dgzh[k], shape_k_axis)
                _set_k(gsh, kmin - 1, _get_k(gsh, kmin), shape_k_axis)
            return gs, gsh

        def _get_k(a, k):
            return a.take(k, axis=self._kax)

        def _set_k(a, k, v, _):
            idx = [slice(None)] * a.ndim
            idx[self._kax] = k
            a[tuple(idx)] = v

        out = {}
        for tag, vzz, vv, kax in (("", vz, var, 2), ("_pl", vz_pl, var_pl, 1)):
            self._kax = kax
            gsq, gsqh = metrics(vzz, kax)
            gam = np.ones_like(gsq)
            gamh = np.ones_like(gsq)
            with np.errstate(invalid="ignore", divide="ignore"):
                GAM2H = gamh ** 2
                GSGAM2 = gam ** 2 * gsq
                GSGAM2H = gamh ** 2 * gsqh
                RGSQRTH = 1.0 / gsqh
                RGAM = 1.0 / gam
                RGAMH = 1.0 / gamh
                RGSGAM2 = 1.0 / GSGAM2
                RGSGAM2H = 1.0 / GSGAM2H
            kshape = [1] * gsq.ndim
            kshape[kax] = kall
            af = vg.afact.reshape(kshape)
            bf = vg.bfact.reshape(kshape)
            cf = vg.cfact.reshape(kshape)
            df = vg.dfact.reshape(kshape)
            RGSGAM2_km1 = np.roll(RGSGAM2, 1, axis=kax)
            RGSGAM2H_kp1 = np.roll(RGSGAM2H, -1, axis=kax)
            C2W = np.stack([af * RGSGAM2 * GSGAM2H, bf * RGSGAM2_km1 * GSGAM2H], axis=-1)
            W2C = np.stack([cf * GSGAM2 * RGSGAM2H_kp1, df * GSGAM2 * RGSGAM2H], axis=-1)
            idx0 = [slice(None)] * C2W.ndim
            idx0[kax] = kmin - 1
            C2W[tuple(idx0)] = 0.0
            idxt = [slice(None)] * W2C.ndim
            idxt[kax] = kmax + 1
            W2C[tuple(idxt)] = 0.0
            with np.errstate(invalid="ignore", divide="ignore"):
                GZH = [-vv[..., c] / gsqh for c in range(3)]
            gz6 = np.stack([af * RGSGAM2 * GSGAM2H * GZH[0], bf * RGSGAM2_km1 * GSGAM2H * GZH[0],
                            af * RGSGAM2 * GSGAM2H * GZH[1], bf * RGSGAM2_km1 * GSGAM2H * GZH[1],
                            af * RGSGAM2 * GSGAM2H * GZH[2], bf * RGSGAM2_km1 * GSGAM2H * GZH[2]],
                           axis=-1)
            idx0 = [slice(None)] * gz6.ndim
            idx0[kax] = kmin - 1
            gz6[tuple(idx0)] = 0.0
            PHI = vzz[..., 0] * self.grav
            out.update({
                "GSQRT" + tag: gsq, "GSQRTH" + tag: gsqh, "GAM2H" + tag: GAM2H,
                "GSGAM2" + tag: GSGAM2, "GSGAM2H" + tag: GSGAM2H, "RGSQRTH" + tag: RGSQRTH,
                "RGAM" + tag: RGAM, "RGAMH" + tag: RGAMH, "RGSGAM2" + tag: RGSGAM2,
                "RGSGAM2H" + tag: RGSGAM2H, "C2Wfact" + tag: C2W, "W2Cfact" + tag: W2C,
                "C2WfactGz" + tag: gz6, "PHI" + tag: PHI,
            })
        self.vmtr = out
        area = self.area[:, :, None, :]
        self.VOLUME = area * out["GSGAM2"] * vg.dgz[None, None, :, None]
        self.VOLUME_pl = self.area_pl[:, None, :] * out["GSGAM2_pl"] * vg.dgz[None, :, None]


def stencil7(a):
    """Stack the 7 stencil neighbours of every interior point: (g-2, g-2, ..., 7)."""
    i, j = a.shape[0], a.shape[1]
    return np.stack([
        a[1:i - 1, 1:j - 1], a[2:i, 1:j - 1], a[2:i, 2:j], a[1:i - 1, 2:j],
        a[0:i - 2, 1:j - 1], a[0:i - 2, 0:j - 2], a[1:i - 1, 0:j - 2],
    ], axis=-1)
