"""Seeded inputs of the per-kernel golden vectors (``kernels.npz``).

No reference code here: this module only draws the random inputs, in the
reference's array layouts, on a small synthetic grid:

  regional (i, j, k, l) = (6, 6, 8, 2): gmin = 1, gmax = 4, kmin = 1, kmax = 6
  pole     (v, k, l)    = (6, 8, 2):    gslf = 0, gmin = 1, gmax = 5

Region 0 carries the pentagon (singular point) mask, region 1 does not.
Operator coefficients, geometry and vertical metrics are random (not a real
grid), so that a transposed index, a swapped coefficient or a dropped term
changes the result.  ``make_kernel_vectors.py`` feeds these inputs to the
reference's kernels and stores the outputs; ``tests/test_oracle_kernels.py``
feeds the same inputs to the oracle and compares.
"""
from __future__ import annotations

import numpy as np

G, KALL, LALL = 6, 8, 2
GMIN, GMAX = 1, G - 2
KMIN, KMAX = 1, KALL - 2
GPL, LPL = 6, 2
RADIUS = 6.371e6
# constants of the benchmark family (mod_const.py with the [cnstparam] overrides)
GRAV, RDRY, CPDRY, RVAP, CPVAP = 9.80616, 287.0, 1004.5, 461.46, 1845.60
CVDRY = CPDRY - RDRY
CVVAP = CPVAP - RVAP
OHM, ALPHA = 7.292e-5, 1.0
ALPHA_D = 1.0e16 / 8.0
DT_SMALL, DT_LARGE = 100.0, 600.0
HAVE_SGP = (True, False)


def _unit(rng, shape):
    v = rng.standard_normal(shape + (3,))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def make_inputs(seed: int = 0) -> dict:
    """Every input array of every kernel vector, reference layouts, float64."""
    rng = np.random.default_rng(seed)
    u = rng.uniform
    n = rng.standard_normal
    R3 = (G, G, KALL, LALL)
    P3 = (GPL, KALL, LPL)
    d: dict = {}
    # --- operator coefficients (KNONE k axis kept, as the reference stores them)
    d["coef_lap"] = n((G, G, 1, LALL, 7))
    d["coef_lap_pl"] = n((GPL, 1, LPL))
    d["coef_div"] = n((G, G, 1, LALL, 3, 7))
    d["coef_div_pl"] = n((GPL, 1, LPL, 3))
    d["coef_grad"] = n((G, G, 1, LALL, 3, 7))
    d["coef_grad_pl"] = n((GPL, 1, LPL, 3))
    d["coef_intp"] = n((G, G, 1, LALL, 3, 3, 2))
    d["coef_intp_pl"] = n((GPL, 1, LPL, 3, 3))
    d["coef_diff"] = n((G, G, 1, LALL, 3, 6))
    d["coef_diff_pl"] = n((GPL, 1, LPL, 3))
    d["GRD_x"] = RADIUS * _unit(rng, (G, G, 1, LALL))
    d["GRD_x_pl"] = RADIUS * _unit(rng, (GPL, 1, LPL))
    pm = np.zeros((1, LALL, 2), dtype=np.int32)
    for l in range(LALL):
        pm[0, l] = (0, 1) if HAVE_SGP[l] else (1, 0)
    d["pntmask"] = pm
    # --- horizontal transport geometry (GMTR_t / GMTR_a / GMTR_p / GRD_xr)
    d["GMTR_t"] = u(0.05, 0.5, (G, G, 1, LALL, 2, 5))
    d["GMTR_a"] = n((G, G, 1, LALL, 3, 12))
    d["GMTR_p"] = u(0.5, 2.0, (G, G, 1, LALL, 8))
    d["GRD_xr"] = RADIUS * _unit(rng, (G, G, 1, LALL, 3))
    d["GMTR_t_pl"] = u(0.05, 0.5, (GPL, 1, LPL, 5))
    d["GMTR_a_pl"] = n((GPL, 1, LPL, 18))
    d["GMTR_p_pl"] = u(0.5, 2.0, (GPL, 1, LPL, 8))
    d["GRD_xr_pl"] = RADIUS * _unit(rng, (GPL, 1, LPL))
    # --- vertical grid factors (kall,)
    for k in ("afact", "bfact", "cfact", "dfact"):
        d[k] = u(0.3, 0.7, KALL)
    d["rdgz"] = u(2e-3, 1e-2, KALL)
    d["rdgzh"] = u(2e-3, 1e-2, KALL)
    # --- vertical metrics (full arrays, as the reference keeps them)
    for k in ("RGAM", "RGAMH", "RGSQRTH", "GAM2H", "RGSGAM2", "GSGAM2", "RGSGAM2H", "GSGAM2H"):
        d[k] = u(0.8, 1.25, R3)
        d[k + "_pl"] = u(0.8, 1.25, P3)
    d["C2Wfact"] = u(0.3, 0.7, R3 + (2,))
    d["C2Wfact_pl"] = u(0.3, 0.7, P3 + (2,))
    d["W2Cfact"] = u(0.3, 0.7, R3 + (2,))
    d["W2Cfact_pl"] = u(0.3, 0.7, P3 + (2,))
    d["C2WfactGz"] = 0.1 * n(R3 + (6,))
    d["C2WfactGz_pl"] = 0.1 * n(P3 + (6,))
    d["PHI"] = u(0.0, 3.0e5, R3)
    d["PHI_pl"] = u(0.0, 3.0e5, P3)
    # --- prognostic / diagnostic state
    for tag, shp in (("", R3), ("_pl", P3)):
        rhog = u(0.5, 1.5, shp)
        prog = np.stack([rhog, 10.0 * n(shp), 10.0 * n(shp), 10.0 * n(shp), n(shp),
                         rhog * u(2.0e5, 2.6e5, shp)], axis=-1)
        d["PROG" + tag] = prog
        d["PROGq" + tag] = (rhog * u(0.0, 1.0e-2, shp))[..., None]
        d["DIAG" + tag] = n(shp + (6,))
        # generic scalar / vector fields for the operators
        for k in ("s1", "s2", "s3", "s4", "s5"):
            d[k + tag] = n(shp)
        for k in ("pos1", "pos2", "pos3"):
            d[k + tag] = u(0.5, 1.5, shp)
        for k in ("v1", "v2", "v3", "v4", "v5", "v6", "v7", "v8"):
            d[k + tag] = 10.0 * n(shp)
        d["pres" + tag] = u(5.0e3, 1.0e5, shp)
        d["eth" + tag] = u(2.5e5, 3.2e5, shp)
        d["ethh" + tag] = u(2.5e5, 3.2e5, shp)
        d["gtilde" + tag] = u(5.0, 15.0, shp)
        d["kh" + tag] = u(1.0e14, 1.0e16, shp)
        d["gTEND" + tag] = n(shp + (6,))
        # a diagonally dominant tridiagonal system (full-k arrays, interior rows used)
        d["Mc" + tag] = u(4.0, 6.0, shp)
        d["Mu" + tag] = 0.5 * n(shp)
        d["Ml" + tag] = 0.5 * n(shp)
        d["PSPLIT" + tag] = 0.1 * n(shp + (6,))
        d["vtmp2" + tag] = n(shp + (3,))
        # tracer transport
        d["q" + tag] = u(0.0, 1.0, shp)
        d["qh" + tag] = u(0.0, 1.0, shp)
        d["dq" + tag] = 0.05 * n(shp)
        d["gradq" + tag] = 1.0e-6 * n(shp + (3,))
    # latitude per column and temperature for the Held-Suarez forcing
    d["lat"] = u(-0.5 * np.pi, 0.5 * np.pi, (G, G, LALL))
    d["tem"] = u(200.0, 310.0, R3)
    # tracer Courant numbers with both signs and a few exact zeros
    ck = 0.3 * n(R3 + (2,))
    ck[0, 0, 2, 0, 0] = 0.0
    d["ck"] = ck
    d["ck_pl"] = 0.3 * n(P3 + (2,))
    ch = 0.2 * n(R3 + (6,))
    d["ch"] = ch
    d["cmask"] = 0.5 - np.copysign(0.5, ch - np.finfo(np.float64).eps)
    chp = 0.2 * n(P3)
    d["ch_pl"] = chp
    d["cmask_pl"] = 0.5 - np.copysign(0.5, chp - np.finfo(np.float64).eps)
    d["q_a"] = u(0.0, 1.0, R3 + (6,))
    d["q_a_pl"] = u(0.0, 1.0, P3)
    d["grd_xc"] = RADIUS * (1.0 + 1e-3 * n(R3 + (3, 3)))
    d["grd_xc_pl"] = RADIUS * (1.0 + 1e-3 * n(P3 + (3,)))
    return d
