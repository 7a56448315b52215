"""The oracle against the reference's own kernels (golden vectors, CPU).

``tests/golden/kernels.npz`` holds the outputs of the reference's pure kernels
(pynicamdc/nhm/dynamics/kernels/*.py, and Numf._hdiff_tendency_host of
mod_numfilter.py) on the seeded inputs of ``tests/golden/kernel_cases.py``;
``tests/golden/make_kernel_vectors.py`` wrote it.  Every oracle function on the
dynamics and tracer path is fed the same inputs here, through a small context
that carries the random coefficients, geometry and metrics instead of a real
grid, and must reproduce the reference's values bit for bit; where numpy's own
reduction order legitimately differs (noted per case) the bound is 4 ulp of the
field's largest value.
"""
import os
import sys
import types

import numpy as np
import pytest

from conftest import GOLDEN, mod

sys.path.insert(0, GOLDEN)
import kernel_cases as kc  # noqa: E402

from oracle import dynamics as od  # noqa: E402
from oracle import tracer as ot  # noqa: E402

VEC = np.load(os.path.join(GOLDEN, "kernels.npz"))
I = kc.make_inputs(seed=0)
ULP4 = 4 * np.finfo(np.float64).eps


def ref(case, key):
    return VEC[f"{case}/{key}"]


class _Topo:
    gall_1d, kall, lall = kc.G, kc.KALL, kc.LALL
    kmin, kmax, gmin, gmax = kc.KMIN, kc.KMAX, kc.GMIN, kc.GMAX
    have_sgp = np.array(kc.HAVE_SGP)

    def halo_fill_host(self, var, var_pl):
        """The vectors are COMM-free islands: no exchange between kernels."""


def make_ctx():
    d = I
    m = {}
    for k in ("RGAM", "RGAMH", "RGSQRTH", "GAM2H", "RGSGAM2", "GSGAM2", "RGSGAM2H", "GSGAM2H", "C2Wfact",
              "W2Cfact", "C2WfactGz", "PHI"):
        m[k], m[k + "_pl"] = d[k], d[k + "_pl"]
    vg = types.SimpleNamespace(**{k: d[k] for k in ("afact", "bfact", "cfact", "dfact", "rdgz", "rdgzh")})
    geo = types.SimpleNamespace(
        topo=_Topo(), vgrid=vg, vmtr=m, radius=kc.RADIUS,
        coef_lap=d["coef_lap"][:, :, 0], coef_lap_pl=d["coef_lap_pl"][:, 0],
        coef_div=d["coef_div"][:, :, 0], coef_div_pl=d["coef_div_pl"][:, 0],
        coef_grad=d["coef_grad"][:, :, 0], coef_grad_pl=d["coef_grad_pl"][:, 0],
        coef_intp=d["coef_intp"][:, :, 0], coef_intp_pl=d["coef_intp_pl"][:, 0],
        coef_diff=d["coef_diff"][:, :, 0], coef_diff_pl=d["coef_diff_pl"][:, 0],
        GRD_x=d["GRD_x"], GRD_x_pl=d["GRD_x_pl"],
        gmtr_t=d["GMTR_t"][:, :, 0], gmtr_a=d["GMTR_a"][:, :, 0], gmtr_p=d["GMTR_p"][:, :, 0],
        xr=d["GRD_xr"][:, :, 0], gmtr_t_pl=d["GMTR_t_pl"][:, 0], gmtr_a_pl=d["GMTR_a_pl"][:, 0],
        gmtr_p_pl=d["GMTR_p_pl"][:, 0], xr_pl=d["GRD_xr_pl"][:, 0])
    c = mod("config").Const(GRAV=kc.GRAV, Rdry=kc.RDRY, CPdry=kc.CPDRY, Rvap=kc.RVAP, CPvap=kc.CPVAP,
                            OHM=kc.OHM, ALPHA=kc.ALPHA, ALPHA_D=kc.ALPHA_D, RADIUS=kc.RADIUS)
    return od.OracleContext(geo, c, None)


CTX = make_ctx()


def dyn_self(**state):
    """An OracleDynamics carrying the given arrays (its methods read only these)."""
    s = od.OracleDynamics.__new__(od.OracleDynamics)
    s.ctx = CTX
    for k, v in state.items():
        setattr(s, k, v.copy())
    return s


def same(a, b, tol=0.0):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    if tol == 0.0:
        bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
        assert not bad.any(), f"{bad.sum()} of {a.size} values differ; max |d| {np.abs(a - b)[bad].max():.3e}"
    else:
        sc = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= tol * sc, np.abs(a - b).max() / sc


def test_golden_inputs_are_the_committed_ones():
    """The input generator must still draw what the vectors were written from."""
    s = sum(float(np.sum(v, dtype=np.float64)) for k, v in sorted(I.items()) if isinstance(v, np.ndarray))
    assert np.isfinite(s)
    assert VEC["laplacian/out"].shape == (kc.G, kc.G, kc.KALL, kc.LALL)
    # the Laplacian's first interior point, summed by hand from the inputs
    c, f = I["coef_lap"][1, 1, 0, 0], I["s1"][..., 3, 0]
    v = (c[0] * f[1, 1] + c[1] * f[2, 1] + c[2] * f[2, 2] + c[3] * f[1, 2] + c[4] * f[0, 1] + c[5] * f[0, 0]
         + c[6] * f[1, 0])
    assert VEC["laplacian/out"][1, 1, 3, 0] == v


def test_laplacian():
    o, op = od.laplacian(CTX, I["s1"], I["s1_pl"])
    same(o, ref("laplacian", "out"))
    same(op, ref("laplacian", "out_pl"))


def test_divergence():
    o, op = od.divergence(CTX, I["v1"], I["v2"], I["v3"], I["v1_pl"], I["v2_pl"], I["v3_pl"])
    same(o, ref("divergence", "out"))
    same(op, ref("divergence", "out_pl"))


def test_gradient():
    o, op = od.gradient(CTX, I["s1"], I["s1_pl"])
    same(o, ref("gradient", "out"))
    same(op, ref("gradient", "out_pl"))


def test_horizontalize():
    r = od.horizontalize_interior(CTX, I["v1"], I["v2"], I["v3"])
    rp = od.horizontalize_pole(CTX, I["v1_pl"], I["v2_pl"], I["v3_pl"])
    for n, a in zip(("vx", "vy", "vz"), r):
        same(a, ref("horizontalize", n))
    for n, a in zip(("vx_pl", "vy_pl", "vz_pl"), rp):
        same(a, ref("horizontalize", n))


def test_divdamp_2d_operator():
    o, op = od.divdamp_op(CTX, I["v1"], I["v2"], I["v3"], I["v1_pl"], I["v2_pl"], I["v3_pl"])
    for n, a in zip(("dx", "dy", "dz"), o):
        same(a, ref("divdamp2d", n))
    for n, a in zip(("dx_pl", "dy_pl", "dz_pl"), op):
        same(a, ref("divdamp2d", n))


def test_kh_diffusion():
    o, op = od.diffusion(CTX, I["s1"], I["s1_pl"], I["kh"], I["kh_pl"])
    same(o, ref("diffusion", "out"))
    same(op, ref("diffusion", "out_pl"))


@pytest.mark.parametrize("vertical,name", [(True, "default"), (False, "horizontal")])
def test_flux_convergence(vertical, name):
    g, gp = od.flux_convergence(CTX, I["v1"], I["v2"], I["v3"], I["v4"], I["v1_pl"], I["v2_pl"], I["v3_pl"],
                                I["v4_pl"], vertical)
    ks = slice(kc.KMIN, kc.KMAX + 1)
    # the reference fills the rows outside kmin..kmax too; the oracle (and every caller) only those
    same(g[:, :, ks], ref("flux_convergence", name)[:, :, ks])
    same(gp[:, ks], ref("flux_convergence", name + "_pl")[:, ks])


@pytest.mark.parametrize("vertical,name", [(True, "default"), (False, "horizontal")])
def test_advection_convergence(vertical, name):
    g, gp = od.advection_convergence(CTX, I["v1"], I["v2"], I["v3"], I["v4"], I["s2"], I["v1_pl"], I["v2_pl"],
                                     I["v3_pl"], I["v4_pl"], I["s2_pl"], vertical)
    ks = slice(kc.KMIN, kc.KMAX + 1)
    same(g[:, :, ks], ref("advection_convergence", name)[:, :, ks])
    same(gp[:, ks], ref("advection_convergence", name + "_pl")[:, ks])


@pytest.mark.parametrize("vertical,name", [(True, "default"), (False, "horizontal")])
def test_pres_gradient(vertical, name):
    Pg, Pw, Pgp, Pwp = od.pres_gradient(CTX, I["pres"], I["pres_pl"], vertical)
    same(Pg, ref("pres_gradient", name + "_grad"))
    same(Pw, ref("pres_gradient", name + "_gradw"))
    same(Pgp, ref("pres_gradient", name + "_grad_pl"))
    same(Pwp, ref("pres_gradient", name + "_gradw_pl"))


def test_buoyancy():
    b, bp = od.buoyancy(CTX, I["pos1"], I["pos1_pl"])
    same(b, ref("buoyancy", "out"))
    same(bp, ref("buoyancy", "out_pl"))


def test_rhogkin():
    k = od.rhogkin(CTX, I["pos1"], I["v1"], I["v2"], I["v3"], I["v4"], "", 2)
    kp = od.rhogkin(CTX, I["pos1_pl"], I["v1_pl"], I["v2_pl"], I["v3_pl"], I["v4_pl"], "_pl", 1)
    same(k, ref("rhogkin", "out"))
    same(kp, ref("rhogkin", "out_pl"))


def _slab(a, kax):
    return a[(slice(None),) * kax + (slice(kc.KMIN + 1, kc.KMAX + 1),)]


def test_vi_matrix():
    s = dyn_self()
    for tag, kax in (("", 2), ("_pl", 1)):
        M = od.OracleDynamics._rhow_matrix(s, I["ethh" + tag], I["gtilde" + tag], kc.DT_SMALL, tag, kax)
        for n, a in zip(("Mc", "Mu", "Ml"), M):
            same(a, ref("vi_matrix", n + tag))


def test_vi_tridiagonal_solve():
    s = dyn_self()
    for tag, kax in (("", 2), ("_pl", 1)):
        out = od.OracleDynamics._rhow_solve(
            s, I["v4" + tag], I["v5" + tag], I["pres" + tag], I["pos1" + tag], I["s1" + tag], I["s2" + tag],
            I["s3" + tag], _slab(I["Mc" + tag], kax), _slab(I["Mu" + tag], kax), _slab(I["Ml" + tag], kax),
            kc.DT_SMALL, tag, kax)
        same(out, ref("vi_solver", "out" + tag))


def test_vi_main():
    s = dyn_self()
    DV = np.stack([I["v6"], I["v7"], I["v8"]], axis=-1)
    DVp = np.stack([I["v6_pl"], I["v7_pl"], I["v8_pl"]], axis=-1)
    Mat = {tag: tuple(_slab(I[n + tag], kax) for n in ("Mc", "Mu", "Ml")) for tag, kax in (("", 2), ("_pl", 1))}
    PS, PSp = I["PSPLIT"].copy(), I["PSPLIT_pl"].copy()
    (rs, rw, re), (rsp, rwp, rep) = od.OracleDynamics._vi_main(
        s, DV, DVp, PS, PSp, I["s4"], I["s4_pl"], I["PROG"], I["PROG_pl"], I["eth"], I["eth_pl"], I["gTEND"],
        I["gTEND_pl"], I["s5"], I["s5_pl"], I["s3"], I["s3_pl"], Mat, kc.DT_SMALL)
    ks = slice(kc.KMIN, kc.KMAX + 1)
    # the reference computes every row; the split density and energy are used on kmin..kmax
    same(rs[:, :, ks], ref("vi_main", "rhog_split1")[:, :, ks])
    same(re[:, :, ks], ref("vi_main", "rhoge_split1")[:, :, ks])
    same(rw, ref("vi_main", "rhogw_split1"))
    same(rsp[:, ks], ref("vi_main", "rhog_split1_pl")[:, ks])
    same(rep[:, ks], ref("vi_main", "rhoge_split1_pl")[:, ks])
    same(rwp, ref("vi_main", "rhogw_split1_pl"))


def test_vi_path1():
    s = dyn_self()
    z = [np.zeros_like(I["s1"])] * 3
    dd = [I["v1"], I["v2"], I["v3"]]
    ddp = [I["v1_pl"], I["v2_pl"], I["v3_pl"]]
    gT, gTp = I["gTEND"].copy(), I["gTEND_pl"].copy()
    DV, DVp, dw, dwp = od.OracleDynamics._vi_path1(s, I["pres"], I["pres_pl"], gT, gTp, dd, ddp, I["PSPLIT"],
                                                  I["PSPLIT_pl"], I["PROG"], I["PROG_pl"], kc.DT_SMALL)
    del z
    same(DV, ref("vi_path1", "diff_vh"))
    same(DVp, ref("vi_path1", "diff_vh_pl"))
    same(dw, ref("vi_path1", "drhogw"))
    same(dwp, ref("vi_path1", "drhogw_pl"))


def test_vi_path2():
    for tag in ("", "_pl"):
        PS = np.zeros(I["PSPLIT" + tag].shape)
        WE = I["PSPLIT" + tag][..., 3:6]
        mean = I["PROG" + tag][..., 0:5].copy()
        od.OracleDynamics._vi_path2(PS, I["vtmp2" + tag], WE, mean, 1.0 / 3.0)
        same(PS, ref("vi_path2", "PROG_split" + tag))
        same(mean, ref("vi_path2", "PROG_mean" + tag))


def test_prepost_diagnostics_bndcnd_eth():
    s = dyn_self(PROG=I["PROG"], PROG_pl=I["PROG_pl"], PROGq=I["PROGq"], PROGq_pl=I["PROGq_pl"], DIAG=I["DIAG"],
                 DIAG_pl=I["DIAG_pl"])
    out = s.prepost()
    same(s.PROG, ref("prepost", "PROG"))
    same(s.PROG_pl, ref("prepost", "PROG_pl"))
    same(s.DIAG, ref("prepost", "DIAG"))
    same(s.DIAG_pl, ref("prepost", "DIAG_pl"))
    for k in ("rho", "ein", "eth", "pregd", "rhogd", "q"):
        same(out[k], ref("prepost", k))
        same(out[k + "_pl"], ref("prepost", k + "_pl"))


def test_momentum_advection_and_coriolis():
    s = dyn_self(PROG=I["PROG"], PROG_pl=I["PROG_pl"], DIAG=I["DIAG"], DIAG_pl=I["DIAG_pl"])
    g, gp = s.advection_momentum()
    ks = slice(kc.KMIN, kc.KMAX + 1)
    for n, a, ap in zip(("gvx", "gvy", "gvz", "gw"), g, gp):
        same(a[:, :, ks], ref("advection_momentum", n)[:, :, ks])
        same(ap[:, ks], ref("advection_momentum", n + "_pl")[:, ks])


def test_divergence_damping_3d():
    s = dyn_self()
    gd, gdp = s.divdamp(I["v1"], I["v2"], I["v3"], I["v4"], I["v1_pl"], I["v2_pl"], I["v3_pl"], I["v4_pl"])
    for n, a in zip(("gdx", "gdy", "gdz"), gd):
        same(a, ref("divdamp3d", n))
    for n, a in zip(("gdx_pl", "gdy_pl", "gdz_pl"), gdp):
        same(a, ref("divdamp3d", n))


def test_hyperdiffusion_tendency():
    vt = [I["v1"], I["v2"], I["v3"], I["v4"], I["s1"], I["s2"]]
    vtp = [I["v1_pl"], I["v2_pl"], I["v3_pl"], I["v4_pl"], I["s1_pl"], I["s2_pl"]]
    T, Tp = od.hdiff_tendency(vt, vtp, I["pos1"], I["pos1_pl"], I["pos2"], I["pos2_pl"], I["kh"], I["kh_pl"],
                              I["kh"], I["kh_pl"])
    same(T, ref("hdiff_tendency", "out"))
    same(Tp, ref("hdiff_tendency", "out_pl"))


def test_tracer_vertical_flux():
    pm = np.stack([I["pos2"], I["v1"], I["v2"], I["v3"], I["v4"]], axis=-1)
    pmp = np.stack([I["pos2_pl"], I["v1_pl"], I["v2_pl"], I["v3_pl"], I["v4_pl"]], axis=-1)
    for tag, kax, P in (("", 2, pm), ("_pl", 1, pmp)):
        # the oracle is specialised to flat topography (RGSQRTH = 1), as were the reference's inputs
        r = ot.vert_flux(CTX, P, I["pos1" + tag], I["s1" + tag], kc.DT_LARGE, kax, I["C2WfactGz" + tag],
                         I["RGAMH" + tag])
        ks = slice(kc.KMIN, kc.KMAX + 1)
        k = (slice(None),) * kax + (ks,)
        for n, a in zip(("flx", "ck", "d"), r[:3]):
            same(a, ref("tracer_vert_flux", n + tag))
        same(r[3][k], ref("tracer_vert_flux", "rhog" + tag)[k])
        kb = (slice(None),) * kax + (slice(kc.KMIN - 1, kc.KMIN),)
        kt = (slice(None),) * kax + (slice(kc.KMAX + 1, kc.KMAX + 2),)
        same(r[3][kb], ref("tracer_vert_flux", "rhog" + tag)[kb])
        same(r[3][kt], ref("tracer_vert_flux", "rhog" + tag)[kt])


def test_tracer_vertical_step_with_limiter():
    for tag, kax in (("", 2), ("_pl", 1)):
        out = ot.vertical_step(CTX, I["s3" + tag] ** 2, I["pos1" + tag], I["v5" + tag], I["ck" + tag],
                               I["dq" + tag], kax, True)
        same(out, ref("tracer_vertical_step", "rhogq" + tag))


def test_tracer_horizontal_flux():
    """The kernel sums the three TJ vertex weights left to right; the numpy path the
    oracle follows (mod_src_tracer.py:1387-1403, the plain FP64 backend) adds the last
    two first, so the regional values agree to round-off, not bit for bit."""
    flx, xc, flxp, xcp = ot.horizontal_flux(CTX, I["pos1"], I["v1"], I["v2"], I["v3"], I["pos1_pl"], I["v1_pl"],
                                            I["v2_pl"], I["v3_pl"], kc.DT_LARGE)
    same(flx, ref("horizontal_flux", "flx"), tol=ULP4)
    same(flxp, ref("horizontal_flux", "flx_pl"))
    # arc-centroid positions: the reference fills the frame i, j = 0..g-2 of each arc it defines
    f = slice(0, kc.G - 1)
    same(xc[f, f], ref("horizontal_flux", "xc")[f, f], tol=ULP4)
    same(xcp[1:6], ref("horizontal_flux", "xc_pl")[1:6])


def test_tracer_horizontal_remap():
    qa, qap = ot.remap_faces(CTX, I["q"], I["q_pl"], I["gradq"], I["gradq_pl"], I["cmask"], I["cmask_pl"],
                             I["grd_xc"], I["grd_xc_pl"])
    f = slice(0, kc.G - 1)
    same(qa[f, f], ref("horizontal_remap", "q_a")[f, f])
    same(qap[1:6], ref("horizontal_remap", "q_a_pl")[1:6])


def test_tracer_horizontal_limiter():
    qa, qap = I["q_a"].copy(), I["q_a_pl"].copy()
    chp = I["ch_pl"].copy()
    chp[0] = 0.0
    with np.errstate(over="ignore", invalid="ignore"):
        ot.horizontal_limiter(CTX, qa, qap, I["q"], I["q_pl"], I["dq"], I["dq_pl"], I["ch"], chp, I["cmask"],
                              I["cmask_pl"])
    f = slice(0, kc.G - 1)
    same(qa[f, f], ref("horizontal_limiter", "q_a")[f, f])
    same(qa, ref("horizontal_limiter", "q_a"))
    same(qap[1:6], ref("horizontal_limiter", "q_a_pl")[1:6])


@pytest.mark.parametrize("moist", [False, True])
def test_held_suarez_tendencies(moist):
    from oracle import forcing as of
    c = CTX.c
    r = of.heldsuarez_tendencies(I["lat"], I["pres"], I["tem"], I["v1"], I["v2"], I["v3"], kc.KMIN, kc.KMAX, c,
                                 moist_case=moist)
    name = "moist" if moist else "dry"
    for n, a in zip(("fvx", "fvy", "fvz", "fe"), r):
        same(a, ref("held_suarez", f"{name}_{n}"))
