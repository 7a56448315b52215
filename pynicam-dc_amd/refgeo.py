"""Geometry for ``HipContext`` from the reference's own setup objects.

``from_reference(adm, grd, gmtr, oprt, vmtr, cnst)`` takes the objects the
reference builds in its setup chain (share/mod_adm.py ``Adm``, mod_grd.py
``Grd``, mod_gmtr.py ``Gmtr``, mod_oprt.py ``Oprt``, mod_vmtr.py ``Vmtr``,
mod_const.py ``Const``) and returns a geometry object with the attributes the
HIP backend reads (``layout.geometry_arrays``, ``HipContext``), so that a run
set up by the reference can hand its coefficients to ``backend="hip"``
unchanged.  The arrays are the reference's, with its KNONE axis dropped
(``OPRT_coef_div`` (g, g, 1, l, 3, 7) -> (g, g, l, 3, 7), mod_oprt.py:93-110;
``GMTR_p/t/a`` mod_gmtr.py:200, 349, 464; ``GRD_x/xr/LAT/LON`` mod_grd.py:112-128;
the ``VMTR_*`` metrics mod_vmtr.py:72-94).  The halo gather map is this
package's (``topology.Topology``, pinned to the reference's region links and
grid-file halos); its dimensions must match the reference's ``Adm``.
"""
from __future__ import annotations

import numpy as np

from . import topology

_VMTR = ("GAM2H", "GSGAM2", "GSGAM2H", "RGSQRTH", "RGAM", "RGAMH", "RGSGAM2", "RGSGAM2H", "W2Cfact", "C2Wfact",
         "C2WfactGz", "PHI")


class _VGrid:
    """The vertical grid factors of ``Grd`` (mod_grd.py:172-222)."""

    def __init__(self, grd):
        for name in ("gz", "gzh", "dgz", "dgzh", "rdgz", "rdgzh", "afact", "bfact", "cfact", "dfact"):
            setattr(self, name, np.asarray(getattr(grd, "GRD_" + name), dtype=np.float64))


class ReferenceGeometry:
    """Duck-typed counterpart of ``geometry.Geometry`` over the reference's arrays."""

    def __init__(self, adm, grd, gmtr, oprt, vmtr, cnst, topo: topology.Topology | None = None):
        glevel, rlevel, vlayer = int(adm.ADM_glevel), int(adm.ADM_rlevel), int(adm.ADM_vlayer)
        self.topo = topo or topology.Topology(glevel, rlevel, vlayer)
        t = self.topo
        if (t.gall_1d, t.kall, t.lall) != (int(adm.ADM_gall_1d), int(adm.ADM_kall), int(adm.ADM_lall)):
            raise ValueError("the reference's Adm dimensions differ from the halo topology's")
        self.vgrid = _VGrid(grd)
        self.bsstate = None     # set a bsstate.BasicState for ref_type 'INPUT'
        self.radius = float(cnst.CONST_RADIUS)
        self.grav = float(cnst.CONST_GRAV)
        k0 = 0

        def reg(a):
            return np.ascontiguousarray(np.asarray(a, dtype=np.float64)[:, :, k0])

        def pl(a):
            return np.ascontiguousarray(np.asarray(a, dtype=np.float64)[:, k0])

        self.coef_div, self.coef_div_pl = reg(oprt.OPRT_coef_div), pl(oprt.OPRT_coef_div_pl)
        self.coef_grad, self.coef_grad_pl = reg(oprt.OPRT_coef_grad), pl(oprt.OPRT_coef_grad_pl)
        self.coef_lap, self.coef_lap_pl = reg(oprt.OPRT_coef_lap), pl(oprt.OPRT_coef_lap_pl)
        self.coef_intp, self.coef_intp_pl = reg(oprt.OPRT_coef_intp), pl(oprt.OPRT_coef_intp_pl)
        self.coef_diff, self.coef_diff_pl = reg(oprt.OPRT_coef_diff), pl(oprt.OPRT_coef_diff_pl)
        self.GRD_x = np.asarray(grd.GRD_x, dtype=np.float64)
        self.GRD_x_pl = np.asarray(grd.GRD_x_pl, dtype=np.float64)
        self.x, self.x_pl = self.GRD_x[:, :, k0], self.GRD_x_pl[:, k0]
        self.xr, self.xr_pl = reg(grd.GRD_xr), pl(grd.GRD_xr_pl)
        self.lat, self.lon = reg(grd.GRD_LAT), reg(grd.GRD_LON)
        self.lat_pl, self.lon_pl = pl(grd.GRD_LAT_pl), pl(grd.GRD_LON_pl)
        self.gmtr_p, self.gmtr_p_pl = reg(gmtr.GMTR_p), pl(gmtr.GMTR_p_pl)
        self.gmtr_t, self.gmtr_t_pl = reg(gmtr.GMTR_t), pl(gmtr.GMTR_t_pl)
        self.gmtr_a, self.gmtr_a_pl = reg(gmtr.GMTR_a), pl(gmtr.GMTR_a_pl)
        self.area, self.area_pl = self.gmtr_p[..., 0], self.gmtr_p_pl[..., 0]
        self.vmtr = {}
        for name in _VMTR:
            self.vmtr[name] = np.asarray(getattr(vmtr, "VMTR_" + name), dtype=np.float64)
            self.vmtr[name + "_pl"] = np.asarray(getattr(vmtr, "VMTR_" + name + "_pl"), dtype=np.float64)
        dgz = self.vgrid.dgz
        self.VOLUME = self.area[:, :, None, :] * self.vmtr["GSGAM2"] * dgz[None, None, :, None]
        self.VOLUME_pl = self.area_pl[:, None, :] * self.vmtr["GSGAM2_pl"] * dgz[None, :, None]


def from_reference(adm, grd, gmtr, oprt, vmtr, cnst, topo: topology.Topology | None = None) -> ReferenceGeometry:
    """Geometry for ``hipdyn.HipContext(geo, const)`` from the reference's setup objects."""
    return ReferenceGeometry(adm, grd, gmtr, oprt, vmtr, cnst, topo)
