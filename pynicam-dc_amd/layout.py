"""Conversion between the reference array layout and the device layout.

Reference (mod_adm.py:172-182): regional (i, j, k, l[, v]) and pole
(v, k, lpl[, var]) arrays in C order.  Device: one field is NF doubles,
regional point ((l*kall + k)*g + j)*g + i followed by the pole points
R + (lpl*kall + k)*6 + v, i.e. the axes reversed so a level of a region is
contiguous in i and a column is strided by g*g.
"""
from __future__ import annotations

import numpy as np


def field_size(g: int, kall: int, lall: int) -> int:
    return lall * kall * g * g + 2 * kall * 6


def pack(var: np.ndarray, var_pl: np.ndarray) -> np.ndarray:
    """(g, g, kall, l, nv) + (6, kall, 2, nv) -> (nv, NF) device-layout array."""
    g, _, kall, lall, nv = var.shape
    R = lall * kall * g * g
    out = np.empty((nv, R + 2 * kall * 6), dtype=np.float64)
    out[:, :R] = np.transpose(var, (4, 3, 2, 1, 0)).reshape(nv, R)
    out[:, R:] = np.transpose(var_pl, (3, 2, 1, 0)).reshape(nv, 2 * kall * 6)
    return out


def unpack(flat: np.ndarray, g: int, kall: int, lall: int):
    """(nv, NF) -> ((g, g, kall, l, nv), (6, kall, 2, nv))."""
    nv = flat.shape[0]
    R = lall * kall * g * g
    var = np.ascontiguousarray(np.transpose(flat[:, :R].reshape(nv, lall, kall, g, g), (4, 3, 2, 1, 0)))
    var_pl = np.ascontiguousarray(np.transpose(flat[:, R:].reshape(nv, 2, kall, 6), (3, 2, 1, 0)))
    return var, var_pl


def pack1(scl: np.ndarray, scl_pl: np.ndarray) -> np.ndarray:
    """(g, g, kall, l) + (6, kall, 2) -> (NF,)."""
    return pack(scl[..., None], scl_pl[..., None])[0]


def unpack1(flat: np.ndarray, g: int, kall: int, lall: int):
    var, var_pl = unpack(flat[None], g, kall, lall)
    return var[..., 0], var_pl[..., 0]


def _soa(a: np.ndarray) -> np.ndarray:
    """(g, g, l, *comp) -> (prod(comp), l*g*g) with column index (l*g + j)*g + i."""
    g, _, lall = a.shape[:3]
    comp = a.shape[3:]
    nc = int(np.prod(comp)) if comp else 1
    b = a.reshape(g, g, lall, nc)
    return np.ascontiguousarray(np.transpose(b, (3, 2, 1, 0)).reshape(nc, lall * g * g))


def vertical_profiles(geo, tol_gz: float = 1.0e-9) -> dict:
    """k-profiles of the vertical metrics; raise when they are not horizontally uniform.

    With flat topography GSQRT = GAM = 1 (mod_vmtr.py:176-197), C2Wfact,
    W2Cfact and PHI depend on k only and C2WfactGz is round-off; the device
    kernels rely on exactly that.
    """
    m = geo.vmtr
    t = geo.topo
    kmin, kmax = t.kmin, t.kmax
    vg = geo.vgrid
    g = t.gall_1d
    s = slice(1, g - 1)

    def uniform(name, ref):
        # ref: (kall,) or (kall, 2); compared on interior points and all pole points
        for arr in (np.moveaxis(m[name][s, s], 2, 0), np.moveaxis(m[name + "_pl"], 1, 0)):
            tgt = ref.reshape((ref.shape[0],) + (1,) * (arr.ndim - ref.ndim) + ref.shape[1:])
            if not np.array_equal(arr, np.broadcast_to(tgt, arr.shape)):
                raise ValueError(f"vertical metric {name} is not horizontally uniform "
                                 "(the HIP backend supports flat topography only)")

    for name in ("GSGAM2", "GSGAM2H", "RGSGAM2", "RGSGAM2H", "RGAM", "RGAMH", "GAM2H", "RGSQRTH"):
        uniform(name, np.ones(t.kall))
    c2w = m["C2Wfact"][1, 1, :, 0, :].copy()
    w2c = m["W2Cfact"][1, 1, :, 0, :].copy()
    phi = m["PHI"][1, 1, :, 0].copy()
    uniform("C2Wfact", c2w)
    uniform("W2Cfact", w2c)
    uniform("PHI", phi)
    gzmax = max(np.max(np.abs(m["C2WfactGz"][s, s])), np.max(np.abs(m["C2WfactGz_pl"])))
    if not gzmax < tol_gz:
        raise ValueError(f"C2WfactGz is not negligible ({gzmax:g}); terrain is not flat")
    return {
        "afact": vg.afact, "bfact": vg.bfact, "cfact": vg.cfact, "dfact": vg.dfact,
        "rdgz": vg.rdgz, "rdgzh": vg.rdgzh,
        "c2w0": c2w[:, 0], "c2w1": c2w[:, 1], "w2c0": w2c[:, 0], "w2c1": w2c[:, 1], "phi": phi,
    }


def geometry_arrays(geo) -> dict:
    """Operator coefficients in the device's structure-of-arrays order."""
    x = geo.GRD_x[:, :, 0]                                   # (g, g, l, 3)
    out = {
        "coef_div": _soa(geo.coef_div),                      # (g, g, l, 3, 7) -> 21
        "coef_grad": _soa(geo.coef_grad),
        "coef_lap": _soa(geo.coef_lap),                      # 7
        "coef_intp": _soa(geo.coef_intp),                    # (g, g, l, 3, 3, 2) -> 18
        "coef_diff": _soa(geo.coef_diff),                    # (g, g, l, 3, 6) -> 18
        "grd_x": _soa(x),                                    # 3
        "coef_div_pl": np.ascontiguousarray(geo.coef_div_pl).reshape(-1),     # (6, 2, 3)
        "coef_grad_pl": np.ascontiguousarray(geo.coef_grad_pl).reshape(-1),   # (6, 2, 3)
        "coef_lap_pl": np.ascontiguousarray(geo.coef_lap_pl).reshape(-1),     # (6, 2)
        "coef_intp_pl": np.ascontiguousarray(geo.coef_intp_pl).reshape(-1),   # (6, 2, 3, 3)
        "coef_diff_pl": np.ascontiguousarray(geo.coef_diff_pl).reshape(-1),   # (6, 2, 3)
        "grd_x_pl": np.ascontiguousarray(geo.GRD_x_pl[:, 0]).reshape(-1),     # (6, 2, 3)
        "sgp": np.ascontiguousarray(geo.topo.have_sgp.astype(np.int32)),
        # tracer transport (GMTR_t weights W1..W3, GMTR_a hexagon normals, GMTR_p RAREA, GRD_xr)
        "tri_w": _soa(geo.gmtr_t[..., 2:5]),                 # (g, g, l, 2, 3) -> 6
        "arc_hn": _soa(geo.gmtr_a[..., :, 0:3]),             # (g, g, l, 3, 3) -> 9
        "rarea": _soa(geo.gmtr_p[..., 1]),                   # 1
        "arc_xr": _soa(geo.xr),                              # (g, g, l, 3, 3) -> 9
        "tri_w_pl": np.ascontiguousarray(geo.gmtr_t_pl[..., 2:5]).reshape(-1),   # (6, 2, 3)
        "arc_hn_pl": np.ascontiguousarray(geo.gmtr_a_pl[..., 0:3]).reshape(-1),  # (6, 2, 3)
        "rarea_pl": np.ascontiguousarray(geo.gmtr_p_pl[0, :, 1]).reshape(-1),   # (2,)
        "arc_xr_pl": np.ascontiguousarray(geo.xr_pl).reshape(-1),               # (6, 2, 3)
        # budgets (GMTR_p P_AREA per column, layer thickness)
        "area": _soa(geo.gmtr_p[..., 0]),
        "dgz": np.ascontiguousarray(geo.vgrid.dgz, dtype=np.float64),
        # prescribed tracer winds (GRD_LAT/GRD_LON, level heights)
        "lat": _soa(geo.lat),
        "lon": _soa(geo.lon),
        "lat_pl": np.ascontiguousarray(geo.lat_pl).reshape(-1),                 # (6, 2)
        "lon_pl": np.ascontiguousarray(geo.lon_pl).reshape(-1),
        "gz": np.ascontiguousarray(geo.vgrid.gz, dtype=np.float64),
        "gzh": np.ascontiguousarray(geo.vgrid.gzh, dtype=np.float64),
    }
    sizes = {"coef_div_pl": 36, "coef_grad_pl": 36, "coef_lap_pl": 12, "coef_intp_pl": 108,
             "coef_diff_pl": 36, "grd_x_pl": 36, "tri_w_pl": 36, "arc_hn_pl": 36, "rarea_pl": 2,
             "arc_xr_pl": 36, "lat_pl": 12, "lon_pl": 12}
    for k, n in sizes.items():
        if out[k].size != n:
            raise ValueError(f"{k} has {out[k].size} entries, expected {n}")
    for k in ("coef_div", "coef_grad", "coef_lap", "coef_intp", "coef_diff", "grd_x", "tri_w", "arc_hn",
              "rarea", "arc_xr", "tri_w_pl", "arc_hn_pl", "rarea_pl", "arc_xr_pl", "area", "lat", "lon",
              "lat_pl", "lon_pl"):
        out[k] = np.ascontiguousarray(out[k], dtype=np.float64)
    out.update({k: np.ascontiguousarray(v, dtype=np.float64) for k, v in vertical_profiles(geo).items()})
    # basic state k-profiles (zeros for NOBASE)
    bs = getattr(geo, "bsstate", None)
    kall = geo.topo.kall
    for k, name in (("pre_bs", "pre"), ("tem_bs", "tem"), ("rho_bs", "rho")):
        out[k] = np.ascontiguousarray(getattr(bs, name) if bs is not None else np.zeros(kall), dtype=np.float64)
    return out
