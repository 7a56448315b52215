"""CPU oracle of the Held-Suarez forcing step.

TEST INFRASTRUCTURE ONLY: the checker for ``nh_forcing_hs``; the product path
never imports it.

Restates, in the reference layout (i, j, k, l):
  AfHeldsuarez.AF_heldsuarez   nhm/forcing/mod_af_heldsuarez.py:43-95
  hs_apply_core                nhm/forcing/mod_forcing.py:12-29
  Dyn.forcing_step (HELD-SUAREZ branch: diagnostics of the stepped state, the
  forcing of the regular grid, the halo exchange; the pole is not forced)
                               nhm/dynamics/mod_dynamics.py:597-720
"""
from __future__ import annotations

import numpy as np

from . import dynamics as od

# the forcing's constants (mod_af_heldsuarez.py:17-27, 31-39)
SIGMA_B = 0.7
KF = 1.0 / (1.0 * 86400.0)
DTH_Z = 10.0
KA = 1.0 / (40.0 * 86400.0)
KS = 1.0 / (4.0 * 86400.0)


def eq_params(moist_case: bool = False):
    """(T_eq0, DT_y) of AF_heldsuarez_init: Held and Suarez (1994) dry, Thatcher and
    Jablonowski (2016) moist."""
    return (294.0, 65.0) if moist_case else (315.0, 60.0)


def heldsuarez_tendencies(lat, pre, tem, vx, vy, vz, kmin, kmax, c, moist_case=False):
    """fvx, fvy, fvz, fe with the ghost levels zeroed (mod_af_heldsuarez.py:43-95).
    lat: (i, j, l); the fields (i, j, k, l)."""
    T_eq0, DT_y = eq_params(moist_case)
    Rd, Cp, CVdry, PRE00 = c.Rdry, c.CPdry, c.CVdry, c.PRE00
    latk = lat[:, :, None, :]
    psref = (0.5 * (pre[:, :, kmin, :] + pre[:, :, kmin - 1, :]))[:, :, None, :]
    sigma = pre / psref
    factor = np.maximum((sigma - SIGMA_B) / (1.0 - SIGMA_B), 0.0)
    fvx = -KF * factor * vx
    fvy = -KF * factor * vy
    fvz = -KF * factor * vz
    sinlat = np.abs(np.sin(latk))
    coslat = np.abs(np.cos(latk))
    ap0 = np.abs(pre / PRE00)
    Kt = KA + (KS - KA) * factor * coslat ** 4
    T_eq = np.maximum(200.0, (T_eq0 - DT_y * sinlat ** 2
                              - DTH_Z * np.log(np.maximum(ap0, 1e-30)) * coslat ** 2) * ap0 ** (Rd / Cp))
    fe = -Kt * (tem - T_eq) * CVdry
    kk = np.arange(pre.shape[2]).reshape(1, 1, -1, 1)
    valid = (kk >= kmin) & (kk <= kmax)
    return tuple(np.where(valid, f, 0.0) for f in (fvx, fvy, fvz, fe))


def forcing_step_hs(ctx, PROG, PROG_pl, PROGq, PROGq_pl, dt, moist_case=False):
    """The forced state after one step of dt: diagnostics of PROG (compute_diagnostics,
    kernels/diag.py:53), the tendencies applied as PROG += dt f rho GSGAM2
    (hs_apply_core), then the halo exchange of the forced prognostics."""
    c, m = ctx.c, ctx.m
    kmin, kmax = ctx.kmin, ctx.kmax
    P = PROG.copy()
    rhog = P[..., od.I_RHOG]
    rho = rhog / m["GSGAM2"]
    vx, vy, vz = P[..., 1] / rhog, P[..., 2] / rhog, P[..., 3] / rhog
    ein = P[..., 5] / rhog
    q = PROGq[..., 0] / rhog
    qd = 1.0 - q
    cv = q * c.CVW[0] + qd * c.CVdry
    tem = ein / cv
    pre = rho * tem * (qd * c.Rdry + q * c.Rvap)
    lat = ctx.geo.lat
    fvx, fvy, fvz, fe = heldsuarez_tendencies(lat, pre, tem, vx, vy, vz, kmin, kmax, c, moist_case)
    G2 = m["GSGAM2"]
    P[..., 1] = P[..., 1] + dt * fvx * rho * G2
    P[..., 2] = P[..., 2] + dt * fvy * rho * G2
    P[..., 3] = P[..., 3] + dt * fvz * rho * G2
    P[..., 5] = P[..., 5] + dt * fe * rho * G2
    Pp = PROG_pl.copy()
    ctx.halo(P, Pp)
    return P, Pp
