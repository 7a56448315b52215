"""Held-Suarez forcing step (nh_forcing_hs) against the oracle restatement.

The oracle (oracle/forcing.py) is pinned bit for bit against the reference's
AF_heldsuarez in test_oracle_kernels.py; here the library's forcing step, run
on the host build of the kernel source and on the GPU, is compared with the
oracle's forcing step on the same state: the forced increments of rhogvx/vy/vz
and rhoge relative to their largest value (the device's log, pow, sin and cos
may differ from numpy's in the last bit), the untouched prognostics and pole
points exactly, halos refilled.
"""
import numpy as np
import pytest

from conftest import mod

hipdyn = mod("hipdyn")

INC_TOL = 1e-12      # forced increment, relative to its largest value (FP64)
INC_TOL_MIXED = 2e-5


def _moist(state, seed=3):
    rng = np.random.default_rng(seed)
    s = {k: v.copy() for k, v in state.items()}
    s["PROGq"][..., 0] = s["PROG"][..., 0] * 1e-3 * (1.0 + 0.5 * rng.random(s["PROG"].shape[:-1]))
    s["PROGq_pl"][..., 0] = s["PROG_pl"][..., 0] * 1e-3
    return s


def check_held_suarez(ctx, case, state, moist, tol):
    from oracle import dynamics as od
    from oracle import forcing as of
    t = case.topo
    octx = od.OracleContext(case.geo, case.const, None)
    dt = float(case.const.DTL)
    P, Pp = of.forcing_step_hs(octx, state["PROG"], state["PROG_pl"], state["PROGq"], state["PROGq_pl"], dt,
                               moist_case=moist)
    dyn = hipdyn.HipDynamics(ctx, state)
    dyn.forcing_step(moist_case=moist)
    got = dyn.state()
    G, Gp = got["PROG"], got["PROG_pl"]
    for v in range(6):
        inc_ref = P[..., v] - state["PROG"][..., v]
        inc_got = G[..., v].astype(np.float64) - state["PROG"][..., v]
        if v in (0, 4):
            assert np.array_equal(G[..., v], state["PROG"][..., v].astype(G.dtype)), v
            continue
        scale = np.abs(inc_ref).max()
        if v != 3:   # rhogvz: the zonal start flow has no z component
            assert scale > 0.0, v
        err = np.abs(inc_got - inc_ref).max() / max(scale, np.abs(P[..., v]).max(), 1e-300)
        assert err < tol, (v, err)
    # the pole points are not forced (the halo fill leaves the pole centres as they were)
    assert np.array_equal(Gp[0], state["PROG_pl"][0].astype(Gp.dtype))
    # the regional halos were refilled from the forced interiors
    g = t.gall_1d
    assert np.abs(G[0, 1:g - 1, t.kmin, :, 1] - P[0, 1:g - 1, t.kmin, :, 1]).max() <= \
        tol * np.abs(P[..., 1] - state["PROG"][..., 1]).max() + 1e-300


@pytest.mark.parametrize("moist", [False, True])
def test_host_build_held_suarez_matches_oracle(case_gl04, state_gl04_tracers, hostemu_lib, moist):
    ctx = hipdyn.HipContext(case_gl04.geo, case_gl04.const, device="cpu", lib_path=hostemu_lib)
    try:
        check_held_suarez(ctx, case_gl04, _moist(state_gl04_tracers), moist, INC_TOL)
    finally:
        ctx.close()


def test_host_build_held_suarez_mixed(case_gl04, state_gl04_tracers, hostemu_lib):
    ctx = hipdyn.HipContext(case_gl04.geo, case_gl04.const, device="cpu", lib_path=hostemu_lib, precision="mixed")
    try:
        check_held_suarez(ctx, case_gl04, _moist(state_gl04_tracers), False, INC_TOL_MIXED)
    finally:
        ctx.close()


def test_forcing_argument_errors(case_gl04, hostemu_lib):
    import ctypes as C
    native = mod("native")
    ctx = hipdyn.HipContext(case_gl04.geo, case_gl04.const, device="cpu", lib_path=hostemu_lib)
    try:
        prog = ctx.empty(6)
        lib = ctx.lib
        assert lib.nh_forcing_hs(ctx.handle, C.c_void_p(prog.data_ptr()), C.c_void_p(None), 600.0, 0,
                                 C.c_void_p(None)) == -1
        assert lib.nh_forcing_hs(ctx.handle, C.c_void_p(prog.data_ptr()), C.c_void_p(prog.data_ptr()), 0.0, 0,
                                 C.c_void_p(None)) == -1
        assert lib.nh_forcing_hs(ctx.handle, C.c_void_p(prog.data_ptr()), C.c_void_p(prog.data_ptr()), 600.0, 2,
                                 C.c_void_p(None)) == -1
        assert lib.nh_forcing_hs_f32(ctx.handle, C.c_void_p(prog.data_ptr()), C.c_void_p(prog.data_ptr()), 600.0,
                                     0, C.c_void_p(None)) == -2
        with pytest.raises(NotImplementedError):
            hipdyn.HipDynamics(ctx, mod("setup").build_case(glevel=4, vlayer=10, ztop=10000.0, dz_bottom=200.0,
                                                            dtl=1200.0).state).forcing_step("DCMIP")
        with pytest.raises(RuntimeError, match="NH_ERR_ARG"):
            native.check(lib, -1, "nh_forcing_hs")
    finally:
        ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,moist,tol", [("float64", False, INC_TOL), ("float64", True, INC_TOL),
                                                  ("mixed", False, INC_TOL_MIXED)])
def test_gpu_held_suarez_matches_oracle(case_gl05, precision, moist, tol):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print(f"  held-suarez gl05 z40 {precision} moist={moist}", flush=True)
    ctx = hipdyn.HipContext(case_gl05.geo, case_gl05.const, device="cuda:0", precision=precision)
    try:
        check_held_suarez(ctx, case_gl05, _moist(case_gl05.state), moist, tol)
    finally:
        ctx.close()
