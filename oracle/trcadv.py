"""CPU oracle for the offline tracer runs (prescribed winds + MIURA2004 transport).

TEST INFRASTRUCTURE ONLY: the checker for ``nh_trcadv_winds`` and
``nh_tracer_step``; the product path never imports it.

* ``test11_velocity`` restates Trcadv.test11_velocity
  (nhm/forcing/mod_af_trcadv.py:23-98, DCMIP 2012 case 1-1);
* ``deform_velocity`` is the deformational flow of the tracer benchmark, the
  non-divergent case 1 of Nair and Lauritzen (2010) with peak speed U:
  u = U sin^2(lon/2) sin(2 lat) cos(pi t/T), v = U/2 sin(lon) cos(lat) cos(pi t/T),
  w = w0 sin(pi z/ztop) cos(2 pi t/T) on the inner half levels;
* ``set_winds`` restates Forcing.forcing_update (nhm/forcing/mod_forcing.py:113-206);
* ``OfflineTracer.step`` restates the TRCADV branch of Dyn.dynamics_step
  (nhm/dynamics/mod_dynamics.py:2030-2076) followed by the closing halo fill.
"""
from __future__ import annotations

import numpy as np

from . import dynamics as od
from .tracer import tracer_advection

PI = 3.141592653589793238


def _unit_east_north(lon, lat):
    """Sp_Unit_East / Sp_Unit_North (mod_af_trcadv.py:151-170)."""
    east = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)], axis=-1)
    north = np.stack([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)], axis=-1)
    return east, north


def test11_velocity(time, lon, lat, zf, zh):
    """(vx, vy, vz) at full levels zf and w at half levels zh (mod_af_trcadv.py:23-98)."""
    a, Rd, g = 6371220.0, 287.0, 9.80616
    tau = 12.0 * 86400.0
    u0 = 2.0 * PI * a / tau
    k0 = 10.0 * a / tau
    omega0 = 23000.0 * PI / tau
    T0 = 300.0
    H = Rd * T0 / g
    p0 = 1000.e2
    dlon = 2.0 * PI * time / tau
    lonp = lon - dlon
    bs = 0.2
    p = p0 * np.exp(-zf / H)
    ptop = p0 * np.exp(-12000.0 / H)
    ud = (omega0 * a) / (bs * ptop) * np.cos(lonp) * np.cos(lat) ** 2 * np.cos(dlon) * (
        -np.exp((p - p0) / (bs * ptop)) + np.exp((ptop - p) / (bs * ptop)))
    u = k0 * np.sin(2.0 * lat) * np.sin(lonp) ** 2 * np.cos(0.5 * dlon) + u0 * np.cos(lat) + ud
    v = k0 * np.sin(2.0 * lonp) * np.cos(lat) * np.cos(0.5 * dlon)
    east, nrth = _unit_east_north(lon, lat)
    vxyz = [east[..., d] * u + nrth[..., d] * v for d in range(3)]
    p = p0 * np.exp(-zh / H)
    s = (1.0 + np.exp((ptop - p0) / (bs * ptop)) - np.exp((p - p0) / (bs * ptop))
         - np.exp((ptop - p) / (bs * ptop)))
    w = -(Rd * T0) / (g * p) * omega0 * np.sin(lonp) * np.cos(lat) * np.cos(dlon) * s
    return vxyz[0], vxyz[1], vxyz[2], w


def deform_velocity(time, lon, lat, zh, inner, params):
    """The benchmark's deformational flow; ``inner`` marks the half levels kmin+1..kmax."""
    T, umax, w0, ztop = params
    ct = np.cos(PI * time / T)
    u = umax * np.sin(0.5 * lon) ** 2 * np.sin(2.0 * lat) * ct
    v = 0.5 * umax * np.sin(lon) * np.cos(lat) * ct
    east, nrth = _unit_east_north(lon, lat)
    vxyz = [east[..., d] * u + nrth[..., d] * v for d in range(3)]
    w = np.where(inner, w0 * np.sin(PI * zh / ztop) * np.cos(2.0 * PI * time / T), 0.0)
    return vxyz[0], vxyz[1], vxyz[2], w


def set_winds(geo, PROG, PROG_pl, time, wind="dcmip11", params=None):
    """rhogv = v rhog, rhogw = w rhog at every point (mod_forcing.py:193-204)."""
    t = geo.topo
    vg = geo.vgrid
    k = np.arange(t.kall)
    inner = (k > t.kmin) & (k <= t.kmax)
    for P, lat, lon, kax in ((PROG, geo.lat, geo.lon, 2), (PROG_pl, geo.lat_pl, geo.lon_pl, 1)):
        lat3 = np.expand_dims(lat, kax)
        lon3 = np.expand_dims(lon, kax)
        shp = [1] * (lat3.ndim)
        shp[kax] = t.kall
        zf, zh, inn = vg.gz.reshape(shp), vg.gzh.reshape(shp), inner.reshape(shp)
        if wind == "dcmip11":
            vx, vy, vz, w = test11_velocity(time, lon3, lat3, zf, zh)
        else:
            vx, vy, vz, w = deform_velocity(time, lon3, lat3, zh, inn, params)
        full = np.broadcast_shapes(lat3.shape, zf.shape)
        rhog = P[..., od.I_RHOG]
        P[..., od.I_RHOGVX] = np.broadcast_to(vx, full) * rhog
        P[..., od.I_RHOGVY] = np.broadcast_to(vy, full) * rhog
        P[..., od.I_RHOGVZ] = np.broadcast_to(vz, full) * rhog
        P[..., od.I_RHOGW] = np.broadcast_to(w, full) * rhog


class OfflineTracer:
    """TRCADV large step: transport with PROG's fluxes, then the winds of the new time."""

    def __init__(self, ctx, state, wind="dcmip11", params=None, time0=0.0):
        self.ctx = ctx
        self.PROG = state["PROG"].copy()
        self.PROG_pl = state["PROG_pl"].copy()
        self.PROGq = state["PROGq"].copy()
        self.PROGq_pl = state["PROGq_pl"].copy()
        self.wind, self.params = wind, params
        self.time = float(time0)
        set_winds(ctx.geo, self.PROG, self.PROG_pl, self.time, wind, params)

    def step(self):
        ctx, c = self.ctx, self.ctx.c
        dtl = c.DTL
        P, Pp = self.PROG, self.PROG_pl
        f0, f0p = np.zeros(P.shape[:-1]), np.zeros(Pp.shape[:-1])
        tracer_advection(ctx, self.PROGq, self.PROGq_pl, P[..., od.I_RHOG].copy(), Pp[..., od.I_RHOG].copy(),
                         P[..., 0:5].copy(), Pp[..., 0:5].copy(), f0, f0p, dtl, limiter=c.THUBURN_LIM)
        ctx.halo(self.PROGq, self.PROGq_pl)
        self.time += dtl
        set_winds(ctx.geo, P, Pp, self.time, self.wind, self.params)
