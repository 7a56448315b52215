"""CPU oracle for the global mass and energy budgets.

TEST INFRASTRUCTURE ONLY: the checker for ``nh_budget``; the product path
never imports it.

Restates ``Embudget._gsum`` and ``Embudget._store_var``
(pynicamdc/share/mod_embudget.py:74-139) for a run whose only water tracer is
qv (tracer 0), in the reference layout (i, j, k, l).
"""
from __future__ import annotations

import numpy as np

from . import dynamics as od


def gsum(ctx, var):
    """Volume-weighted interior sum, pole omitted (mod_embudget.py:74-80)."""
    g, kmin, kmax = ctx.g, ctx.kmin, ctx.kmax
    sl = (slice(1, g - 1), slice(1, g - 1), slice(kmin, kmax + 1), slice(None))
    return float(np.sum(var[sl] * ctx.geo.VOLUME[sl]))


def store_var(ctx, PROG, PROGq, lhv):
    """Energies (pot, int, kin, tot) and masses (dry, vap): mod_embudget.py:82-139.

    The diagnostic state is re-derived from the prognostic one as
    ``compute_diagnostics`` does: rho = rhog / GSGAM2, ein = rhoge / rhog,
    q = rhogq / rhog, qd = 1 - q, tem = ein / (qd CVdry + q CVW).
    """
    c, m = ctx.c, ctx.m
    GSGAM2 = m["GSGAM2"]
    rhog = PROG[..., od.I_RHOG]
    with np.errstate(divide="ignore", invalid="ignore"):
        rho = rhog / GSGAM2
        ein = PROG[..., od.I_RHOGE] / rhog
        q = PROGq[..., 0] / rhog
    qd = 1.0 - q
    cv = q * c.CVW[0] + qd * c.CVdry
    tem = ein / cv

    pot = gsum(ctx, rho * m["PHI"])
    eint = qd * c.CVdry * tem
    eint = eint + q * c.CVW[0] * tem
    eint = eint + q * lhv
    eint = gsum(ctx, rho * eint)
    with np.errstate(divide="ignore", invalid="ignore"):
        kin = od.rhogkin(ctx, rhog, PROG[..., od.I_RHOGVX], PROG[..., od.I_RHOGVY], PROG[..., od.I_RHOGVZ],
                         PROG[..., od.I_RHOGW], "", 2)
    ekin = gsum(ctx, kin * m["RGSGAM2"])
    mdry = gsum(ctx, rho * qd)
    mvap = gsum(ctx, rho * q)
    return {"pot": pot, "int": eint, "kin": ekin, "tot": pot + eint + ekin, "dry": mdry, "vap": mvap}


def sfc_pressure(ctx, PROG, PROGq):
    """Surface density and pressure per column: Bndc.BNDCND_pre_sfc
    (pynicamdc/nhm/share/mod_bndcnd.py:521-555) with a flat surface z_srf = 0.

    rho_srf is the quadratic Lagrange polynomial through (z, rho) at the lowest
    three full levels, evaluated at z_srf; pre_srf = pre(kmin) +
    0.5 (rho_srf + rho(kmin)) GRAV (z(kmin) - z_srf).  rho and pre are the
    diagnostic density and pressure of ``compute_diagnostics``.  Returns
    (rho_srf, pre_srf), each (i, j, l).
    """
    c, kmin = ctx.c, ctx.kmin
    GSGAM2 = ctx.m["GSGAM2"]
    rhog = PROG[..., od.I_RHOG]
    rho = rhog / GSGAM2
    ein = PROG[..., od.I_RHOGE] / rhog
    q = PROGq[..., 0] / rhog
    qd = 1.0 - q
    tem = ein / (q * c.CVW[0] + qd * c.CVdry)
    pre = rho * tem * (qd * c.Rdry + q * c.Rvap)
    z = ctx.geo.vgrid.gz
    z1, z2, z3, zs = z[kmin], z[kmin + 1], z[kmin + 2], 0.0
    r1, r2, r3 = rho[:, :, kmin], rho[:, :, kmin + 1], rho[:, :, kmin + 2]
    rho_srf = (((zs - z2) * (zs - z3)) / ((z1 - z2) * (z1 - z3)) * r1
               + ((zs - z1) * (zs - z3)) / ((z2 - z1) * (z2 - z3)) * r2
               + ((zs - z1) * (zs - z2)) / ((z3 - z1) * (z3 - z2)) * r3)
    pre_srf = pre[:, :, kmin] + 0.5 * (rho_srf + r1) * c.GRAV * (z1 - zs)
    return rho_srf, pre_srf
