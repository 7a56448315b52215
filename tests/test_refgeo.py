"""The geometry adapter from the reference's setup objects (refgeo.from_reference).

The reference's own setup chain cannot run here (it needs toml, zarr and its
grid files), so the objects are stand-ins that carry the reference's attribute
names and array shapes (the KNONE axis of the K0 arrays, mod_adm.py:172-182;
OPRT_coef_* mod_oprt.py:93-110, GMTR_p/t/a mod_gmtr.py:200/349/464, GRD_*
mod_grd.py:112-185, VMTR_* mod_vmtr.py:72-94), filled from this package's
geometry.  The adapter must hand the host build of the kernels exactly the
same coefficients, and a step through it must equal the step through the
package's own geometry bit for bit.
"""
import types

import numpy as np

from conftest import mod

hipdyn = mod("hipdyn")
layout = mod("layout")
refgeo = mod("refgeo")


def _reference_objects(case):
    geo, t, vg, c = case.geo, case.topo, case.vgrid, case.const
    K = lambda a: np.asarray(a)[:, :, None]           # noqa: E731  (g, g, l, ...) -> (g, g, 1, l, ...)
    Kp = lambda a: np.asarray(a)[:, None]             # noqa: E731  (6, 2, ...) -> (6, 1, 2, ...)
    adm = types.SimpleNamespace(ADM_glevel=t.glevel, ADM_rlevel=t.rlevel, ADM_vlayer=t.kall - 2,
                                ADM_gall_1d=t.gall_1d, ADM_kall=t.kall, ADM_lall=t.lall)
    grd = types.SimpleNamespace(GRD_x=geo.GRD_x, GRD_x_pl=geo.GRD_x_pl, GRD_xr=K(geo.xr), GRD_xr_pl=Kp(geo.xr_pl),
                                GRD_LAT=K(geo.lat), GRD_LON=K(geo.lon), GRD_LAT_pl=Kp(geo.lat_pl),
                                GRD_LON_pl=Kp(geo.lon_pl))
    for n in ("gz", "gzh", "dgz", "dgzh", "rdgz", "rdgzh", "afact", "bfact", "cfact", "dfact"):
        setattr(grd, "GRD_" + n, getattr(vg, n))
    gmtr = types.SimpleNamespace(GMTR_p=K(geo.gmtr_p), GMTR_p_pl=Kp(geo.gmtr_p_pl), GMTR_t=K(geo.gmtr_t),
                                 GMTR_t_pl=Kp(geo.gmtr_t_pl), GMTR_a=K(geo.gmtr_a), GMTR_a_pl=Kp(geo.gmtr_a_pl))
    oprt = types.SimpleNamespace()
    for n in ("div", "grad", "lap", "intp", "diff"):
        setattr(oprt, f"OPRT_coef_{n}", K(getattr(geo, f"coef_{n}")))
        setattr(oprt, f"OPRT_coef_{n}_pl", Kp(getattr(geo, f"coef_{n}_pl")))
    vmtr = types.SimpleNamespace(**{"VMTR_" + k: v for k, v in geo.vmtr.items()})
    cnst = types.SimpleNamespace(CONST_RADIUS=geo.radius, CONST_GRAV=geo.grav)
    return adm, grd, gmtr, oprt, vmtr, cnst


def test_reference_geometry_gives_the_same_coefficients_and_step(case_gl04, state_gl04_tracers, hostemu_lib):
    rg = refgeo.from_reference(*_reference_objects(case_gl04), topo=case_gl04.topo)
    a, b = layout.geometry_arrays(case_gl04.geo), layout.geometry_arrays(rg)
    assert sorted(a) == sorted(b)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(rg.VOLUME, case_gl04.geo.VOLUME)
    out = []
    for geo in (case_gl04.geo, rg):
        ctx = hipdyn.HipContext(geo, case_gl04.const, device="cpu", lib_path=hostemu_lib)
        dyn = hipdyn.HipDynamics(ctx, state_gl04_tracers)
        dyn.dynamics_step()
        out.append(dyn.state())
        ctx.close()
    for key in ("PROG", "PROG_pl", "PROGq", "PROGq_pl"):
        assert np.array_equal(out[0][key], out[1][key]), key


def test_reference_geometry_checks_dimensions(case_gl04):
    objs = list(_reference_objects(case_gl04))
    objs[0] = types.SimpleNamespace(**{**vars(objs[0]), "ADM_lall": 10})
    try:
        refgeo.from_reference(*objs, topo=case_gl04.topo)
    except ValueError as e:
        assert "dimensions" in str(e)
    else:
        raise AssertionError("a dimension mismatch was accepted")
