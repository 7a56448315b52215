"""Initial states of the benchmark family (host setup, FP64).

The benchmark's dry test atmosphere is not one of the reference's idealised
cases (nhm/share/mod_ideal_init.py only offers them as templates), so it is
defined here:

* reference profile: T = 288 K - 6.5 K/km up to 11 km, isothermal above,
  surface pressure 1e5 Pa;
* balanced zonal wind u = 35 m/s cos(lat): geopotential on pressure surfaces
  shifted by -(a*Omega*u0 + u0^2/2) sin^2(lat) (gradient-wind balance of
  solid-body rotation), i.e. each column is the reference column lifted by
  that shift divided by g;
* potential-temperature bumps: 8 Gaussians of 1 K amplitude and 500 km
  e-folding radius at seeded random positions;
* pressure re-integrated upward in the model's discrete hydrostatic form
  (p_k - p_{k-1}) / dz_h = -g (a_k rho_k + b_k rho_{k-1}), the balance the
  vertical implicit solver sees.

Prognostics follow cnvvar_diag2prg (nhm/share/mod_cnvvar.py:20): rhog = rho *
GSGAM2, rhogv = rhog * v, rhogw = 0, rhoge = rhog * cv * T, tracers zero.
"""
from __future__ import annotations

import numpy as np

from .config import Const
from .geometry import Geometry


def reference_profile(z, ps=1.0e5, t0=288.0, lapse=0.0065, ztrop=11000.0, grav=9.80616, rd=287.0):
    z = np.asarray(z, dtype=np.float64)
    t_trop = t0 - lapse * ztrop
    p_trop = ps * (t_trop / t0) ** (grav / (rd * lapse))
    tem = np.where(z < ztrop, t0 - lapse * z, t_trop)
    p_low = ps * (np.maximum(t0 - lapse * z, 1.0) / t0) ** (grav / (rd * lapse))
    p_high = p_trop * np.exp(-grav * (z - ztrop) / (rd * t_trop))
    return tem, np.where(z < ztrop, p_low, p_high)


def theta_bumps(lat, lon, centres, amp=1.0, radius_m=5.0e5, a=6.371e6):
    """Sum of Gaussian bumps exp(-(d/r)^2) with great-circle distance d."""
    out = np.zeros(np.shape(lat))
    for (clat, clon) in centres:
        cosd = np.sin(lat) * np.sin(clat) + np.cos(lat) * np.cos(clat) * np.cos(lon - clon)
        d = a * np.arccos(np.clip(cosd, -1.0, 1.0))
        out = out + amp * np.exp(-(d / radius_m) ** 2)
    return out


def bump_centres(seed: int, n: int = 8):
    rng = np.random.default_rng(seed)
    lat = np.arcsin(rng.uniform(-0.8, 0.8, n))
    lon = rng.uniform(-np.pi, np.pi, n)
    return list(zip(lat, lon))


def baroclinic_bumps_state(geo: Geometry, c: Const, seed: int = 0, u0: float = 35.0, amp: float = 1.0):
    """Dry balanced state with potential-temperature bumps; returns prognostic arrays.

    PROG (g, g, kall, l, 6), PROG_pl (6, kall, 2, 6), PROGq / PROGq_pl zeros.
    u0 = 0 and amp = 0 give the resting, horizontally uniform atmosphere.
    """
    t = geo.topo
    vg = geo.vgrid
    kall, kmin = t.kall, t.kmin
    centres = bump_centres(seed)
    out = {}
    for reg in (True, False):
        lat = geo.lat if reg else geo.lat_pl            # (g, g, l) or (6, 2)
        lon = geo.lon if reg else geo.lon_pl
        shift = -(c.RADIUS * c.OHM * u0 + 0.5 * u0 * u0) * np.sin(lat) ** 2 / c.GRAV
        z = vg.gz.reshape((1,) * lat.ndim + (kall,))
        zz = z - shift[..., None]
        tem, pre = reference_profile(zz, grav=c.GRAV, rd=c.Rdry)
        th_b = theta_bumps(lat, lon, centres, amp=amp, a=c.RADIUS)[..., None]
        kappa = c.Rdry / c.CPdry
        tem = tem + th_b * (pre / c.PRE00) ** kappa
        # discrete hydrostatic re-integration from the first level
        p = pre.copy()
        rdg = c.GRAV
        for k in range(kmin + 1, kall):
            dz = vg.dgzh[k]
            a_, b_ = vg.afact[k], vg.bfact[k]
            rho_km1 = p[..., k - 1] / (c.Rdry * tem[..., k - 1])
            p[..., k] = (p[..., k - 1] - rdg * dz * b_ * rho_km1) / (1.0 + rdg * dz * a_ / (c.Rdry * tem[..., k]))
        k = kmin - 1
        dz = vg.dgzh[kmin]
        rho_k = p[..., kmin] / (c.Rdry * tem[..., kmin])
        a_, b_ = vg.afact[kmin], vg.bfact[kmin]
        p[..., k] = (p[..., kmin] + rdg * dz * a_ * rho_k) / (1.0 - rdg * dz * b_ / (c.Rdry * tem[..., k]))
        rho = p / (c.Rdry * tem)
        u = u0 * np.cos(lat)[..., None] * np.ones(kall)
        vx = -u * np.sin(lon)[..., None]
        vy = u * np.cos(lon)[..., None]
        vz = np.zeros_like(u)
        # move the k axis to the reference position
        if reg:
            def kfirst(a):
                return np.transpose(a, (0, 1, 3, 2))     # (g, g, l, k) -> (g, g, k, l)
            gsg = geo.vmtr["GSGAM2"]
        else:
            def kfirst(a):
                return np.transpose(a, (0, 2, 1))        # (6, 2, k) -> (6, k, 2)
            gsg = geo.vmtr["GSGAM2_pl"]
        rho, tem, vx, vy, vz = (kfirst(a) for a in (rho, tem, vx, vy, vz))
        rhog = rho * gsg
        prog = np.stack([rhog, rhog * vx, rhog * vy, rhog * vz, np.zeros_like(rhog),
                         rhog * c.CVdry * tem], axis=-1)
        out["PROG" if reg else "PROG_pl"] = np.nan_to_num(prog)
    PROGq = np.zeros(out["PROG"].shape[:-1] + (c.NTRACER,))
    PROGq_pl = np.zeros(out["PROG_pl"].shape[:-1] + (c.NTRACER,))
    t.halo_fill_host(out["PROG"], out["PROG_pl"])
    t.halo_fill_host(PROGq, PROGq_pl)
    return {"PROG": out["PROG"], "PROG_pl": out["PROG_pl"], "PROGq": PROGq, "PROGq_pl": PROGq_pl}


def cosine_blobs(lat, lon, centres, radius_m=1.0e6, a=6.371e6):
    """Sum of compact cosine bells 0.5 (1 + cos(pi d / r)) for d < r, capped at 1."""
    out = np.zeros(np.shape(lat))
    for (clat, clon) in centres:
        cosd = np.sin(lat) * np.sin(clat) + np.cos(lat) * np.cos(clat) * np.cos(lon - clon)
        d = a * np.arccos(np.clip(cosd, -1.0, 1.0))
        out = out + np.where(d < radius_m, 0.5 * (1.0 + np.cos(np.pi * d / radius_m)), 0.0)
    return np.minimum(out, 1.0)


def blob_tracers(geo: Geometry, state: dict, ntracer: int = 4, seed: int = 0, nblob: int = 6,
                 radius_m: float = 1.0e6, with_qv: bool = True, with_constant: bool = False):
    """Passive tracers for the transport tests: each a seeded sum of compact cosine blobs in [0, 1].

    Returns a copy of ``state`` whose PROGq/PROGq_pl hold qv = 0 (when
    ``with_qv``) followed by ``ntracer`` tracer masses rhog * q on the physical
    levels (zero on the ghost levels), halos filled.  ``with_constant`` appends
    a tracer with q = 1: transported with the same mass fluxes it carries the
    density the transport implies, the reference for mixing-ratio bounds when
    the prescribed fluxes do not conserve the state's own density.
    """
    t = geo.topo
    kmin, kmax = t.kmin, t.kmax
    rng = np.random.default_rng(seed + 1000)
    P, Pp = state["PROG"], state["PROG_pl"]
    nq = ntracer + (1 if with_qv else 0) + (1 if with_constant else 0)
    Q = np.zeros(P.shape[:-1] + (nq,))
    Qp = np.zeros(Pp.shape[:-1] + (nq,))
    k = slice(kmin, kmax + 1)
    for n in range(ntracer):
        centres = list(zip(np.arcsin(rng.uniform(-0.9, 0.9, nblob)), rng.uniform(-np.pi, np.pi, nblob)))
        q = cosine_blobs(geo.lat, geo.lon, centres, radius_m, geo.radius)
        qp = cosine_blobs(geo.lat_pl, geo.lon_pl, centres, radius_m, geo.radius)
        iq = n + (1 if with_qv else 0)
        Q[:, :, k, :, iq] = P[:, :, k, :, 0] * q[:, :, None, :]
        Qp[:, k, :, iq] = Pp[:, k, :, 0] * qp[:, None, :]
    if with_constant:
        Q[:, :, k, :, -1] = P[:, :, k, :, 0]
        Qp[:, k, :, -1] = Pp[:, k, :, 0]
    t.halo_fill_host(Q, Qp)
    return dict(state, PROGq=Q, PROGq_pl=Qp)
