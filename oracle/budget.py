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
