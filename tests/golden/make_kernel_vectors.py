"""Write ``kernels.npz``: outputs of the reference's own kernels on seeded inputs.

Runs ONLY in a container where the reference is mounted at /root/reference
(it never travels anywhere else, and no file it writes contains reference
code: the npz holds output arrays only):

    python tests/golden/make_kernel_vectors.py

The reference's pure kernels (pynicamdc/nhm/dynamics/kernels/*.py) import only
numpy, dataclasses and share/mod_backend.py, so they are imported from the
reference tree as they are, without writing bytecode there.  Inputs come from
``kernel_cases.make_inputs(seed=0)``.  Where the reference composes several
kernels in a host routine (mod_src.py, mod_vi.py, mod_numfilter.py,
mod_bndcnd.py, mod_src_tracer.py), the composition below calls them in that
routine's order, cited per case; the COMM halo exchanges between them are left
out (every case is one COMM-free island).  Two reference functions live in
modules that cannot be imported here (they import ``toml``):
``Numf._hdiff_tendency_host`` (mod_numfilter.py:1519) is compiled from its own
source text, unchanged, and called with the inputs it reads.

A few of the jax-form kernels (the Thuburn limiters) write their results with
``array.at[idx].set(v)``; they are run on numpy arrays through ``_AtArray``, a
numpy array whose ``.at[idx].set(v)`` returns an updated copy, which is what
the jax method does.
"""
from __future__ import annotations

import ast
import os
import sys
import textwrap
import types

import numpy as np

sys.dont_write_bytecode = True
REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, REF)

import kernel_cases as kc  # noqa: E402

from pynicamdc.nhm.dynamics.kernels import (  # noqa: E402
    advconv, advconvmom, bndcnd, buoyancy, diag, divdamppostcomm, fluxconv, horizontalflux,
    horizontalizevec, horizontallimiter, horizontalremap, oprt3ddivdamp, oprtdiffusion,
    oprtdivdamp, oprtgradient, oprtlaplacian, presgrad, rhogkin, thrmdyn, tracervertadv,
    tracervertflux, verticallimiter, vimain, vimatrix, vipath1, vipath2, virhowsolver,
)

X, Y, Z = 0, 1, 2
I_SRC_H, I_SRC_D = 1, 3           # mod_src.py:24-26
UNDEF = -9.9999e30
EPS = float(np.finfo(np.float64).eps)
BIG = float(np.finfo(np.float64).max)
KMIN, KMAX = kc.KMIN, kc.KMAX


# ----------------------------------------------------------------------
# numpy arrays with the jax ``.at[idx].set`` method
# ----------------------------------------------------------------------
class _AtArray(np.ndarray):
    @property
    def at(self):
        arr = self

        class _Idx:
            def __getitem__(self, idx):
                class _Set:
                    def set(self, v):
                        out = np.array(arr, copy=True).view(_AtArray)
                        out[idx] = v
                        return out
                return _Set()
        return _Idx()


def _wrap(x):
    return x.view(_AtArray) if isinstance(x, np.ndarray) else x


class _AtNumpy(types.SimpleNamespace):
    """numpy namespace whose array results are ``_AtArray``."""

    def __getattr__(self, name):
        f = getattr(np, name)
        if not callable(f):
            return f

        def g(*a, **k):
            r = f(*a, **k)
            if isinstance(r, tuple):
                return tuple(_wrap(x) for x in r)
            return _wrap(r)
        return g


XPA = _AtNumpy()


def at(a):
    return np.asarray(a).view(_AtArray)


# ----------------------------------------------------------------------
# configurations of the benchmark family
# ----------------------------------------------------------------------
def flux_cfg():
    return fluxconv.FluxConvCfg(kmin=KMIN, kmax=KMAX, have_pl=True, XDIR=X, YDIR=Y, ZDIR=Z, gslf_pl=0,
                                gmax_pl=5, I_SRC_default=I_SRC_D, I_SRC_horizontal=I_SRC_H)


def adv_cfg():
    return advconv.AdvConvCfg(kmin=KMIN, kmax=KMAX, have_pl=True, I_SRC_default=I_SRC_D)


def sol_cfg():
    return virhowsolver.ViSolverCfg(kmin=KMIN, kmax=KMAX, have_pl=True, GRAV=kc.GRAV, Rdry=kc.RDRY,
                                    CVdry=kc.CVDRY, alpha=kc.ALPHA)


def bnd_cfg():
    # TEM top/bottom, FREE top / RIGID bottom momentum (mod_bndcnd.py:31-180 defaults)
    return bndcnd.BndCfg(kmin=KMIN, kmax=KMAX, have_pl=True, is_top_tem=True, is_top_epl=False,
                         is_btm_tem=True, is_btm_epl=False, is_top_rigid=False, is_top_free=True,
                         is_btm_rigid=True, is_btm_free=False, GRAV=kc.GRAV, Rdry=kc.RDRY)


def kin_cfg():
    return rhogkin.RhogkinCfg(kmin=KMIN, kmax=KMAX, have_pl=True, UNDEF=UNDEF)


def presgrad_cfg():
    return presgrad.PresGradCfg(kmin=KMIN, kmax=KMAX, have_pl=True, XDIR=X, YDIR=Y, ZDIR=Z, gslf_pl=0, gmax_pl=5,
                                nxyz=3, first_layer_remedy=True, rscale=kc.RADIUS, plmask=1, horizontalize=True,
                                I_SRC_default=I_SRC_D, I_SRC_horizontal=I_SRC_H)


def divdamp_cfg():
    return oprtdivdamp.OprtDivdampCfg(have_pl=True, gmax=kc.GMAX, gslf_pl=0, gmin_pl=1, gmax_pl=5, k0=0, TI=0,
                                      TJ=1, XDIR=X, YDIR=Y, ZDIR=Z)


def hz_cfg():
    return horizontalizevec.HorizontalizeVecCfg(have_pl=True, XDIR=X, YDIR=Y, ZDIR=Z)


def metrics(d):
    keys = ["RGAM", "RGAMH", "RGSQRTH", "GAM2H", "RGSGAM2", "GSGAM2", "RGSGAM2H", "GSGAM2H", "C2Wfact",
            "W2Cfact", "C2WfactGz", "PHI"]
    out = {}
    for k in keys:
        out[k], out[k + "_pl"] = d[k], d[k + "_pl"]
    for k in ("coef_div", "coef_div_pl", "coef_grad", "coef_grad_pl", "GRD_x", "GRD_x_pl", "rdgz", "rdgzh",
              "afact", "bfact", "cfact", "dfact"):
        out[k] = d[k]
    return out


# ----------------------------------------------------------------------
# the cases
# ----------------------------------------------------------------------
def case_laplacian(d):
    cfg = oprtlaplacian.OprtLaplacianCfg(have_pl=True, gslf_pl=0, gmax_pl=5)
    o, op = oprtlaplacian.compute_oprt_laplacian(d["s1"], d["s1_pl"], d["coef_lap"], d["coef_lap_pl"], cfg, np)
    return {"out": o, "out_pl": op}


def case_divergence(d):
    cfg = flux_cfg()
    o = fluxconv._divergence(d["v1"], d["v2"], d["v3"], d["coef_div"], cfg, np)
    op = fluxconv._divergence_pl(d["v1_pl"], d["v2_pl"], d["v3_pl"], d["coef_div_pl"], cfg, np)
    return {"out": o, "out_pl": op}


def case_gradient(d):
    cfg = oprtgradient.OprtGradientCfg(have_pl=True, gslf_pl=0, gmax_pl=5, k0=0, XDIR=X, YDIR=Y, ZDIR=Z)
    o, op = oprtgradient.compute_oprt_gradient(d["s1"], d["s1_pl"], d["coef_grad"], d["coef_grad_pl"], cfg, np)
    return {"out": o, "out_pl": op}


def case_horizontalize(d):
    r = horizontalizevec.compute_horizontalize_vec(d["v1"], d["v2"], d["v3"], d["v1_pl"], d["v2_pl"], d["v3_pl"],
                                                   d["GRD_x"], d["GRD_x_pl"], kc.RADIUS, hz_cfg(), np)
    return {"vx": r[0], "vy": r[1], "vz": r[2], "vx_pl": r[3], "vy_pl": r[4], "vz_pl": r[5]}


def case_divdamp2d(d):
    r = oprtdivdamp.compute_oprt_divdamp(d["v1"], d["v2"], d["v3"], d["v1_pl"], d["v2_pl"], d["v3_pl"],
                                         d["coef_intp"], d["coef_diff"], d["coef_intp_pl"], d["coef_diff_pl"],
                                         divdamp_cfg(), np)
    return {"dx": r[0], "dy": r[1], "dz": r[2], "dx_pl": r[3], "dy_pl": r[4], "dz_pl": r[5]}


def case_diffusion(d):
    cfg = oprtdiffusion.OprtDiffusionCfg(have_pl=True, gmin=kc.GMIN, gmax=kc.GMAX, nxyz=3, gslf_pl=0, gmin_pl=1,
                                         gmax_pl=5, k0=0, TI=0, TJ=1)
    o, op = oprtdiffusion.compute_oprt_diffusion(d["s1"], d["s1_pl"], d["kh"], d["kh_pl"], d["coef_intp"],
                                                 d["coef_intp_pl"], d["coef_diff"], d["coef_diff_pl"], d["pntmask"],
                                                 cfg, np)
    return {"out": o, "out_pl": op}


def _flux_conv(d, vx, vy, vz, w, vxp, vyp, vzp, wp, fluxtype):
    return fluxconv.compute_flux_convergence(
        vx, vy, vz, w, vxp, vyp, vzp, wp, d["RGAM"], d["RGAMH"], d["RGSQRTH"], d["C2WfactGz"], d["coef_div"],
        d["rdgz"], d["RGAM_pl"], d["RGAMH_pl"], d["RGSQRTH_pl"], d["C2WfactGz_pl"], d["coef_div_pl"], fluxtype,
        flux_cfg(), np)


def _adv_conv(d, vx, vy, vz, w, s, vxp, vyp, vzp, wp, sp, fluxtype):
    """mod_src.py:499 src_advection_convergence: scaled fluxes, then the flux convergence."""
    f = advconv.compute_scaled_fluxes(vx, vy, vz, w, s, vxp, vyp, vzp, wp, sp, d["afact"], d["bfact"], fluxtype,
                                      adv_cfg(), np)
    return _flux_conv(d, *f, fluxtype)


def case_flux_convergence(d):
    out = {}
    for name, ft in (("default", I_SRC_D), ("horizontal", I_SRC_H)):
        g, gp = _flux_conv(d, d["v1"], d["v2"], d["v3"], d["v4"], d["v1_pl"], d["v2_pl"], d["v3_pl"], d["v4_pl"], ft)
        out[name], out[name + "_pl"] = g, gp
    return out


def case_advection_convergence(d):
    out = {}
    for name, ft in (("default", I_SRC_D), ("horizontal", I_SRC_H)):
        g, gp = _adv_conv(d, d["v1"], d["v2"], d["v3"], d["v4"], d["s2"], d["v1_pl"], d["v2_pl"], d["v3_pl"],
                          d["v4_pl"], d["s2_pl"], ft)
        out[name], out[name + "_pl"] = g, gp
    return out


def case_pres_gradient(d):
    out = {}
    for name, ft in (("default", I_SRC_D), ("horizontal", I_SRC_H)):
        r = presgrad.compute_pres_gradient(
            d["pres"], d["pres_pl"], d["RGAM"], d["RGAMH"], d["C2WfactGz"], d["coef_grad"], d["GRD_x"], d["rdgz"],
            d["rdgzh"], d["GAM2H"], d["RGSGAM2"], d["RGAM_pl"], d["RGAMH_pl"], d["C2WfactGz_pl"],
            d["coef_grad_pl"], d["GRD_x_pl"], d["GAM2H_pl"], d["RGSGAM2_pl"], ft, presgrad_cfg(), np)
        for n_, a in zip(("grad", "gradw", "grad_pl", "gradw_pl"), r):
            out[f"{name}_{n_}"] = a
    return out


def case_buoyancy(d):
    cfg = buoyancy.BuoyCfg(kmin=KMIN, kmax=KMAX, GRAV=kc.GRAV, have_pl=True)
    b, bp = buoyancy.compute_buoyancy(d["pos1"], d["pos1_pl"], d["C2Wfact"], d["C2Wfact_pl"], cfg, np)
    return {"out": b, "out_pl": bp}


def case_rhogkin(d):
    k = rhogkin.compute_rhogkin_reg(d["pos1"], d["v1"], d["v2"], d["v3"], d["v4"], d["C2Wfact"], d["W2Cfact"],
                                    kin_cfg(), np)
    kp = rhogkin.compute_rhogkin_pl(d["pos1_pl"], d["v1_pl"], d["v2_pl"], d["v3_pl"], d["v4_pl"], d["C2Wfact_pl"],
                                    d["W2Cfact_pl"], kin_cfg(), np)
    return {"out": k, "out_pl": kp}


def case_vi_matrix(d):
    cfg = vimatrix.ViMatrixCfg(kmin=KMIN, kmax=KMAX, have_pl=True, GRAV=kc.GRAV, Rdry=kc.RDRY, CVdry=kc.CVDRY,
                               alpha=kc.ALPHA)
    c = vimatrix.precombine_matrix_reg(d["RGSGAM2"], d["GAM2H"], d["RGAMH"], d["rdgz"], d["dfact"], d["cfact"], cfg)
    cp = vimatrix.precombine_matrix_pl(d["RGSGAM2_pl"], d["GAM2H_pl"], d["RGAMH_pl"], d["rdgz"], d["dfact"],
                                       d["cfact"], cfg)
    M = vimatrix.compute_rhow_matrix_reg(d["ethh"], d["gtilde"], d["RGSQRTH"], d["rdgzh"], d["dfact"], d["cfact"],
                                         c, kc.DT_SMALL, cfg, np)
    Mp = vimatrix.compute_rhow_matrix_pl(d["ethh_pl"], d["gtilde_pl"], d["RGSQRTH_pl"], d["rdgzh"], d["dfact"],
                                         d["cfact"], cp, kc.DT_SMALL, cfg, np)
    return {"Mc": M[0], "Mu": M[1], "Ml": M[2], "Mc_pl": Mp[0], "Mu_pl": Mp[1], "Ml_pl": Mp[2]}


def case_vi_solver(d):
    r = virhowsolver.compute_rhow_solver_reg(
        d["v4"], d["v5"], d["pres"], d["pos1"], d["s1"], d["s2"], d["s3"], d["Mc"], d["Mu"], d["Ml"],
        d["RGAMH"], d["RGSGAM2"], d["RGAM"], d["RGSGAM2H"], d["GSGAM2H"], d["rdgzh"], d["afact"], d["bfact"],
        kc.DT_SMALL, sol_cfg(), np)
    rp = virhowsolver.compute_rhow_solver_pl(
        d["v4_pl"], d["v5_pl"], d["pres_pl"], d["pos1_pl"], d["s1_pl"], d["s2_pl"], d["s3_pl"], d["Mc_pl"],
        d["Mu_pl"], d["Ml_pl"], d["RGAMH_pl"], d["RGSGAM2_pl"], d["RGAM_pl"], d["RGSGAM2H_pl"], d["GSGAM2H_pl"],
        d["rdgzh"], d["afact"], d["bfact"], kc.DT_SMALL, sol_cfg(), np)
    return {"out": r, "out_pl": rp}


def vimain_inputs(d):
    """P and C of compute_vi_main, wired as mod_vi.py:2055-2090 wires them."""
    P = {}
    for tag in ("", "_pl"):
        PS, PR, gT = d["PSPLIT" + tag], d["PROG" + tag], d["gTEND" + tag]
        P["rhogvx_s1" + tag], P["rhogvy_s1" + tag], P["rhogvz_s1" + tag] = d["v6" + tag], d["v7" + tag], d["v8" + tag]
        P["rhog_s0" + tag], P["rhogvx_s0" + tag], P["rhogvy_s0" + tag] = PS[..., 0], PS[..., 1], PS[..., 2]
        P["rhogvz_s0" + tag], P["rhogw_s0" + tag], P["rhoge_s0" + tag] = PS[..., 3], PS[..., 4], PS[..., 5]
        P["preg_s0" + tag] = d["s4" + tag]
        P["rhog0" + tag], P["rhogvx0" + tag], P["rhogvy0" + tag] = PR[..., 0], PR[..., 1], PR[..., 2]
        P["rhogvz0" + tag], P["rhogw0" + tag] = PR[..., 3], PR[..., 4]
        P["eth0" + tag] = d["eth" + tag]
        P["grhog" + tag], P["grhoge" + tag] = gT[..., 0], gT[..., 5]
        P["grhogw" + tag] = d["s5" + tag]
        P["grhogetot" + tag] = d["s3" + tag]
        P["Mc" + tag], P["Mu" + tag], P["Ml" + tag] = d["Mc" + tag], d["Mu" + tag], d["Ml" + tag]
    return P


def case_vi_main(d):
    cfg = vimain.VimainCfg(kmin=KMIN, kmax=KMAX, have_pl=True, TIME_split=True, Rdry=kc.RDRY, CVdry=kc.CVDRY,
                           I_SRC_default=I_SRC_D, I_SRC_horizontal=I_SRC_H, flux=flux_cfg(), adv=adv_cfg(),
                           sol=sol_cfg(), bnd=bnd_cfg(), kin=kin_cfg())
    return dict(vimain.compute_vi_main(vimain_inputs(d), metrics(d), kc.DT_SMALL, cfg, np))


def case_vi_path1(d):
    cfg = vipath1.ViPath1Cfg(kmin=KMIN, kmax=KMAX, have_pl=True, TIME_split=True, alpha=kc.ALPHA, XDIR=X, YDIR=Y,
                             ZDIR=Z, I_RHOGVX=1, I_RHOGVY=2, I_RHOGVZ=3, I_RHOGW=4, gradtype=I_SRC_H,
                             presgrad=presgrad_cfg(), bnd=bnd_cfg())
    P = {}
    for tag in ("", "_pl"):
        z = np.zeros_like(d["s1" + tag])
        P["preg" + tag] = d["pres" + tag]
        P["g_TEND" + tag] = d["gTEND" + tag]
        P["ddivdvx" + tag], P["ddivdvy" + tag], P["ddivdvz" + tag] = d["v1" + tag], d["v2" + tag], d["v3" + tag]
        P["ddivdw" + tag] = z                     # vertical divergence damping off (mod_numfilter.py:2238)
        P["ddivdvx_2d" + tag] = P["ddivdvy_2d" + tag] = P["ddivdvz_2d" + tag] = z   # 2-D damping off
        PS = d["PSPLIT" + tag]
        P["psvx" + tag], P["psvy" + tag], P["psvz" + tag] = PS[..., 1], PS[..., 2], PS[..., 3]
        P["prog_rhog" + tag] = d["PROG" + tag][..., 0]
    return dict(vipath1.compute_vi_path1(P, metrics(d), kc.DT_SMALL, cfg, np))


def case_vi_path2(d):
    cfg = vipath2.ViPath2Cfg(have_pl=True, I_RHOG=0, I_RHOGVX=1, I_RHOGVY=2, I_RHOGVZ=3, I_RHOGW=4, I_RHOGE=5)
    P = {"rweight_itr": 1.0 / 3.0}
    for tag in ("", "_pl"):
        P["diff_vh" + tag] = d["vtmp2" + tag]
        P["diff_we" + tag] = d["PSPLIT" + tag][..., 3:6]
        P["PROG_mean" + tag] = d["PROG" + tag][..., 0:5]
    return dict(vipath2.compute_vi_path2_update(P, None, cfg, np))


def case_prepost(d):
    """compute_diagnostics, BNDCND_all (mod_bndcnd.py:182-262, for the regional and the pole
    arrays) and THRMDYN_eth, in the order of mod_dynamics.py:965-1283 (NOBASE: the
    perturbations are the full pressure and density times GSGAM2)."""
    cfg = diag.DiagCfg(I_RHOG=0, I_RHOGVX=1, I_RHOGVY=2, I_RHOGVZ=3, I_RHOGW=4, I_RHOGE=5, I_pre=0, I_tem=1,
                       I_vx=2, I_vy=3, I_vz=4, I_w=5, kmin=KMIN, kmax=KMAX, nmin=0, nmax=0, iqv=0, Rdry=kc.RDRY,
                       Rvap=kc.RVAP, CVdry=kc.CVDRY)
    CVW = np.array([kc.CVVAP])
    out = {}
    for tag in ("", "_pl"):
        PROG = d["PROG" + tag].copy()
        PROGq, DIAG0 = d["PROGq" + tag], d["DIAG" + tag]
        GSGAM2, C2W, PHI, GZ = d["GSGAM2" + tag], d["C2Wfact" + tag], d["PHI" + tag], d["C2WfactGz" + tag]
        if tag:   # the pole arrays (v, k, l) as (v, 1, k, l): the same elementwise code
            PROG, PROGq, DIAG0 = PROG[:, None].copy(), PROGq[:, None], DIAG0[:, None]
            GSGAM2, C2W, PHI, GZ = GSGAM2[:, None], C2W[:, None], PHI[:, None], GZ[:, None]
        rho, DIAG, ein, q, _cv, _qd = diag.compute_diagnostics(PROG, PROGq, DIAG0, GSGAM2, C2W, CVW, cfg, np)
        rho, DIAG, ein = rho.copy(), DIAG.copy(), ein.copy()
        bc = bnd_cfg()
        kt, kb = KMAX + 1, KMIN - 1
        tt, tb, pt, pb, rt, rb = bndcnd.compute_bndcnd_thermo_reg(DIAG[..., 1], rho, DIAG[..., 0], PHI, bc, np)
        DIAG[:, :, kt, :, 1], DIAG[:, :, kb, :, 1] = tt, tb
        DIAG[:, :, kt, :, 0], DIAG[:, :, kb, :, 0] = pt, pb
        rho[:, :, kt], rho[:, :, kb] = rt, rb
        PROG[:, :, kt, :, 0] = rho[:, :, kt] * GSGAM2[:, :, kt]
        PROG[:, :, kb, :, 0] = rho[:, :, kb] * GSGAM2[:, :, kb]
        ein[:, :, kt] = kc.CVDRY * DIAG[:, :, kt, :, 1]
        ein[:, :, kb] = kc.CVDRY * DIAG[:, :, kb, :, 1]
        PROG[:, :, kt, :, 5] = PROG[:, :, kt, :, 0] * ein[:, :, kt]
        PROG[:, :, kb, :, 5] = PROG[:, :, kb, :, 0] * ein[:, :, kb]
        vt = bndcnd.compute_bndcnd_rhovxvyvz_reg(PROG[..., 0], PROG[..., 1], PROG[..., 2], PROG[..., 3], bc, np)
        for c in range(3):
            PROG[:, :, kt, :, 1 + c] = vt[c]
            PROG[:, :, kb, :, 1 + c] = vt[3 + c]
        for kk in (kt, kb):
            for c in range(3):
                DIAG[:, :, kk, :, 2 + c] = PROG[:, :, kk, :, 1 + c] / PROG[:, :, kk, :, 0]
        rw_t, rw_b = bndcnd.compute_bndcnd_rhow_reg(PROG[..., 1], PROG[..., 2], PROG[..., 3], GZ, bc, np)
        PROG[:, :, kt, :, 4] = rw_t
        PROG[:, :, KMIN, :, 4] = rw_b
        PROG[:, :, kb, :, 4] = 0.0
        rw, rg = PROG[..., 4], PROG[..., 0]
        DIAG[:, :, kt, :, 5] = rw[:, :, kt] / (C2W[:, :, kt, :, 0] * rg[:, :, kt] + C2W[:, :, kt, :, 1] * rg[:, :, KMAX])
        DIAG[:, :, KMIN, :, 5] = rw[:, :, KMIN] / (C2W[:, :, KMIN, :, 0] * rg[:, :, KMIN]
                                                    + C2W[:, :, KMIN, :, 1] * rg[:, :, kb])
        DIAG[:, :, kb, :, 5] = 0.0
        eth = thrmdyn.compute_thrmdyn_eth(ein, DIAG[..., 0], rho, np)
        pregd = (DIAG[..., 0] - 0.0) * GSGAM2
        rhogd = (rho - 0.0) * GSGAM2
        res = {"PROG": PROG, "DIAG": DIAG, "rho": rho, "ein": ein, "eth": eth, "pregd": pregd, "rhogd": rhogd,
               "q": q}
        for k, v in res.items():
            out[k + tag] = v[:, 0] if tag else v
    return out


def case_advection_momentum(d):
    """mod_src.py:96-497: merged velocity, three advection convergences, tendency."""
    cfg = advconvmom.AdvMomCfg(kmin=KMIN, kmax=KMAX, have_pl=True, XDIR=X, YDIR=Y, ZDIR=Z, rscale=kc.RADIUS,
                               ohm=kc.OHM, alpha=kc.ALPHA)
    D, Dp, P, Pp = d["DIAG"], d["DIAG_pl"], d["PROG"], d["PROG_pl"]
    vv = advconvmom.compute_merged_velocity_reg(D[..., 2], D[..., 3], D[..., 4], D[..., 5], d["cfact"], d["dfact"],
                                                d["GRD_x"], cfg, np)
    vvp = advconvmom.compute_merged_velocity_pl(Dp[..., 2], Dp[..., 3], Dp[..., 4], Dp[..., 5], d["cfact"],
                                                d["dfact"], d["GRD_x_pl"], cfg, np)
    dv = [_adv_conv(d, P[..., 1], P[..., 2], P[..., 3], P[..., 4], vv[c], Pp[..., 1], Pp[..., 2], Pp[..., 3],
                    Pp[..., 4], vvp[c], I_SRC_D) for c in range(3)]
    t = advconvmom.compute_momentum_tendency_reg(dv[0][0], dv[1][0], dv[2][0], P[..., 0], vv[0], vv[1], d["GRD_x"],
                                                 d["C2Wfact"], cfg, np)
    tp = advconvmom.compute_momentum_tendency_pl(dv[0][1], dv[1][1], dv[2][1], Pp[..., 0], vvp[0], vvp[1],
                                                 d["GRD_x_pl"], d["C2Wfact_pl"], cfg, np)
    out = {}
    for n_, a, ap in zip(("gvx", "gvy", "gvz", "gw"), t, tp):
        out[n_], out[n_ + "_pl"] = a, ap
    return out


def case_divdamp3d(d):
    """numfilter_divdamp (mod_numfilter.py:2032-2268, 3-D, lap_order 2): OPRT3D_divdamp, then
    (after the halo exchange, left out here) the post-COMM kernel with divdamp_coef = alpha_d."""
    cfg3 = oprt3ddivdamp.Oprt3DDivdampCfg(have_pl=True, kmin=KMIN, kmax=KMAX, gmax=kc.GMAX, gslf_pl=0, gmin_pl=1,
                                          gmax_pl=5, k0=0, TI=0, TJ=1, XDIR=X, YDIR=Y, ZDIR=Z)
    r = oprt3ddivdamp.compute_oprt3d_divdamp(
        d["v1"], d["v2"], d["v3"], d["v4"], d["v1_pl"], d["v2_pl"], d["v3_pl"], d["v4_pl"], d["coef_intp"],
        d["coef_diff"], d["coef_intp_pl"], d["coef_diff_pl"], d["C2WfactGz"], d["RGAMH"], d["RGSQRTH"], d["RGAM"],
        d["C2WfactGz_pl"], d["RGAMH_pl"], d["RGSQRTH_pl"], d["RGAM_pl"], d["rdgz"], d["pntmask"], cfg3, np)
    vt2 = np.stack(r[0:3], axis=-1)
    vt2p = np.stack(r[3:6], axis=-1)
    coef = np.full(d["s1"].shape, kc.ALPHA_D)
    coefp = np.full(d["s1_pl"].shape, kc.ALPHA_D)
    g = divdamppostcomm.compute_divdamp_post_comm(vt2, vt2p, coef, coefp, d["coef_intp"], d["coef_diff"],
                                                  d["coef_intp_pl"], d["coef_diff_pl"], d["GRD_x"], d["GRD_x_pl"],
                                                  kc.RADIUS, divdamp_cfg(), hz_cfg(), np)
    out = {"vtmp2": vt2, "vtmp2_pl": vt2p}
    for n_, a in zip(("gdx", "gdy", "gdz", "gdx_pl", "gdy_pl", "gdz_pl"), g):
        out[n_] = a
    return out


def _ref_function(path, cls, name, **globs):
    """Compile one method of a reference module from its own source text (``globs``:
    the module-level names it reads)."""
    src = open(os.path.join(REF, path)).read()
    tree = ast.parse(src)
    for node in ast.walk(tree):
        if isinstance(node, ast.ClassDef) and node.name == cls:
            for f in node.body:
                if isinstance(f, ast.FunctionDef) and f.name == name:
                    code = textwrap.dedent(ast.get_source_segment(src, f))
                    ns = {"np": np, "adm": types.SimpleNamespace(ADM_have_pl=True)}
                    ns.update(globs)
                    exec(compile(code, f"{path}:{f.lineno}", "exec"), ns)
                    return ns[name]
    raise LookupError(name)


def case_hdiff_tendency(d):
    """Numf._hdiff_tendency_host (mod_numfilter.py:1519-1578) with the Laplacian-1 filter
    off (its coefficient zero) and the half-level Kh of the linear branch (:1210)."""
    fn = _ref_function("pynicamdc/nhm/dynamics/mod_numfilter.py", "Numf", "_hdiff_tendency_host")
    kh, khp = d["kh"], d["kh_pl"]
    self = types.SimpleNamespace(Kh_coef=kh, Kh_coef_lap1=np.zeros_like(kh), Kh_coef_pl=khp,
                                 Kh_coef_lap1_pl=np.zeros_like(khp))
    rcnf = types.SimpleNamespace(I_RHOG=0, I_RHOGVX=1, I_RHOGVY=2, I_RHOGVZ=3, I_RHOGW=4, I_RHOGE=5)
    vt = np.stack([d["v1"], d["v2"], d["v3"], d["v4"], d["s1"], d["s2"]], axis=-1)
    vtp = np.stack([d["v1_pl"], d["v2_pl"], d["v3_pl"], d["v4_pl"], d["s1_pl"], d["s2_pl"]], axis=-1)
    T = np.zeros(vt.shape)
    Tp = np.zeros(vtp.shape)
    fn(self, T, Tp, vt, vtp, np.zeros_like(vt), np.zeros_like(vtp), kh, np.zeros_like(kh), khp, np.zeros_like(khp),
       d["pos1"], d["pos2"], d["pos1_pl"], d["pos2_pl"], rcnf, np.float64)
    return {"out": T, "out_pl": Tp}


def case_tracer_vert_flux(d):
    cfg = tracervertflux.TracerVertFluxCfg(kmin=KMIN, kmax=KMAX, have_pl=True)
    ones, ones_p = np.ones_like(d["s1"]), np.ones_like(d["s1_pl"])
    r = tracervertflux.compute_tracer_vert_flux(
        d["v1"], d["v2"], d["v3"], d["v4"], d["v1_pl"], d["v2_pl"], d["v3_pl"], d["v4_pl"], d["pos1"], d["pos1_pl"],
        d["s1"], d["s1_pl"], d["C2WfactGz"], d["RGAMH"], ones, d["C2WfactGz_pl"], d["RGAMH_pl"], ones_p, d["rdgz"],
        kc.DT_LARGE, 0.0, cfg, np)
    return dict(zip(("flx", "ck", "d", "rhog", "flx_pl", "ck_pl", "d_pl", "rhog_pl"), r))


def case_tracer_vertical_step(d):
    """mod_src_tracer.py:420-533 for one tracer: q and q_h, the vertical limiter, the update."""
    cfg = tracervertadv.TracerVertAdvCfg(kmin=KMIN, kmax=KMAX, have_pl=True)
    out = {}
    for tag in ("", "_pl"):
        rq, rden, flx, ck, dd = d["s3" + tag] ** 2, d["pos1" + tag], d["v5" + tag], d["ck" + tag], d["dq" + tag]
        if tag:
            q, q_h = tracervertadv.compute_vert_qh_pl(rq, rden, d["afact"], d["bfact"], cfg, np)
        else:
            q, q_h = tracervertadv.compute_vert_qh(rq, rden, d["afact"], d["bfact"], cfg, np)
        shp = q.shape if not tag else (q.shape[0], 1) + q.shape[1:]
        vcfg = verticallimiter.VLimiterCfg(iall=shp[0], jall=shp[1], kall=shp[2], lall=shp[3], kmin=KMIN, kmax=KMAX,
                                           BIG=BIG, EPS=EPS)
        qh_l = verticallimiter.compute_vertical_limiter(at(q_h.reshape(shp)), at(q.reshape(shp)),
                                                        at(dd.reshape(shp)), at(ck.reshape(shp + (2,))), vcfg, XPA)
        qh_l = np.asarray(qh_l).reshape(q.shape)
        if tag:
            new = tracervertadv.compute_vert_update_pl(rq, flx, qh_l, d["rdgz"], cfg, np)
        else:
            new = tracervertadv.compute_vert_update(rq, flx, qh_l, d["rdgz"], cfg, np)
        out["q_h" + tag], out["q_h_limited" + tag], out["rhogq" + tag] = q_h, qh_l, new
    return out


def case_horizontal_flux(d):
    cfg = horizontalflux.HorizFluxCfg(iall=kc.G, jall=kc.G, K0=0, TI=0, TJ=1, AI=0, AIJ=1, AJ=2, W1=2, W2=3, W3=4,
                                      HNX=0, HNY=1, HNZ=2, P_RAREA=1, XDIR=X, YDIR=Y, ZDIR=Z, have_pl=True,
                                      gslf_pl=0, gmin_pl=1, gmax_pl=5, EPS=EPS)
    r = horizontalflux.compute_horizontal_flux(
        d["pos1"], d["v1"], d["v2"], d["v3"], d["pos1_pl"], d["v1_pl"], d["v2_pl"], d["v3_pl"], d["GMTR_t"],
        d["GMTR_a"], d["GMTR_p"], d["GRD_xr"], d["pntmask"], d["GMTR_t_pl"], d["GMTR_a_pl"], d["GMTR_p_pl"],
        d["GRD_xr_pl"], kc.DT_LARGE, cfg, np)
    return dict(zip(("flx", "xc", "flx_pl", "xc_pl"), r))


def case_horizontal_remap(d):
    cfg = horizontalremap.RemapCfg(AI=0, AIJ=1, AJ=2, XDIR=X, YDIR=Y, ZDIR=Z)
    cfgp = horizontalremap.RemapCfgPl(n=0, gmin=1, gmax=5, XDIR=X, YDIR=Y, ZDIR=Z)
    qa = horizontalremap.compute_horizontal_remap(d["q"], d["gradq"], d["grd_xc"], d["cmask"],
                                                  d["GRD_x"][:, :, 0], cfg, np)
    qap = horizontalremap.compute_horizontal_remap_pl(d["q_pl"], d["gradq_pl"], d["grd_xc_pl"], d["cmask_pl"],
                                                      d["GRD_x_pl"][:, 0], cfgp, np)
    return {"q_a": qa, "q_a_pl": qap}


def case_horizontal_limiter(d):
    """horizontal_limiter_thuburn (mod_src_tracer.py:2115-2662): Qin/Qout, (halo of Qout
    left out), then the apply; regional and pole kernels of kernels/horizontallimiter.py."""
    cfg = horizontallimiter.HLimiterCfg(iall=kc.G, jall=kc.G, lall=kc.LALL, I_min=0, I_max=1, BIG=BIG, EPS=EPS,
                                        have_sgp=kc.HAVE_SGP)
    cfgp = horizontallimiter.HLimiterCfgPl(n=0, gmin=1, gmax=5, I_min=0, I_max=1, BIG=BIG, EPS=EPS)
    with np.errstate(over="ignore", invalid="ignore"):
        Qin, Qout = horizontallimiter.compute_horizontal_limiter_qout(at(d["q"]), at(d["dq"]), at(d["ch"]),
                                                                      at(d["cmask"]), cfg, XPA)
        qa = horizontallimiter.compute_horizontal_limiter_apply(at(d["q_a"]), Qin, Qout, at(d["cmask"]), cfg, XPA)
        chp = d["ch_pl"].copy()
        chp[0] = 0.0
        Qinp, Qoutp = horizontallimiter.compute_horizontal_limiter_qout_pl(at(d["q_pl"]), at(d["dq_pl"]), at(chp),
                                                                           at(d["cmask_pl"]), cfgp, XPA)
        qap = horizontallimiter.compute_horizontal_limiter_apply_pl(at(d["q_a_pl"]), Qinp, Qoutp, at(d["cmask_pl"]),
                                                                    cfgp, XPA)
    return {"Qout": np.asarray(Qout), "q_a": np.asarray(qa), "Qout_pl": np.asarray(Qoutp), "q_a_pl": np.asarray(qap)}


def case_held_suarez(d):
    """AfHeldsuarez.AF_heldsuarez (nhm/forcing/mod_af_heldsuarez.py:43-95), dry and moist."""
    from pynicamdc.nhm.forcing.mod_af_heldsuarez import AfHeldsuarez
    cnst = types.SimpleNamespace(CONST_Rdry=kc.RDRY, CONST_CPdry=kc.CPDRY, CONST_CVdry=kc.CVDRY,
                                 CONST_PRE00=1.0e5)
    out = {}
    for name, moist in (("dry", False), ("moist", True)):
        hs = AfHeldsuarez()
        hs.AF_heldsuarez_init(moist_case=moist)
        r = hs.AF_heldsuarez(d["lat"], d["pres"], d["tem"], d["v1"], d["v2"], d["v3"], KMIN, KMAX, cnst, np.float64,
                             xp=np)
        for n_, a in zip(("fvx", "fvy", "fvz", "fe"), r):
            out[f"{name}_{n_}"] = a
    return out


def case_restart_pack(d):
    """Prgv._pack5d (nhm/share/mod_prgvar.py:150-154): the (ij, k, l) layout of a restart
    variable, ij = j * gall_1d + i."""
    adm = types.SimpleNamespace(ADM_gall_1d=kc.G, ADM_kall=kc.KALL, ADM_lall=kc.LALL)
    pack = _ref_function("pynicamdc/nhm/share/mod_prgvar.py", "Prgv", "_pack5d", adm=adm)
    return {"prog_rhogvx": pack(None, d["PROG"], 1), "prog_rhoge": pack(None, d["PROG"], 5)}


CASES = {n[5:]: f for n, f in sorted(globals().items()) if n.startswith("case_")}


def main():
    d = kc.make_inputs(seed=0)
    out = {}
    for name, fn in CASES.items():
        res = fn({k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in d.items()})
        for k, v in res.items():
            out[f"{name}/{k}"] = np.asarray(v, dtype=np.float64)
        print(f"{name}: {', '.join(sorted(res))}", flush=True)
    np.savez_compressed(os.path.join(HERE, "kernels.npz"), **out)


if __name__ == "__main__":
    main()
