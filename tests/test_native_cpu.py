"""The C-ABI library: exports, argument checks, and the kernel source run on the host.

No GPU is used: the product library (gfx950) is only loaded and its symbols
listed; the kernel logic is exercised through the host build of the same
sources (tests/hostemu) against the CPU oracle.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, interior, mod, pole_centre, rel_err

native = mod("native")
layout = mod("layout")
hipdyn = mod("hipdyn")

HEADER = os.path.join(ROOT, "include", "nicamhip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(nh_\w+)\s*\(", text, re.M)))


def test_header_declares_the_bound_symbols():
    assert set(declared_symbols()) == set(native.EXPORTS)


def test_gpu_library_exports_every_declared_symbol():
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("gfx950 library not built (run __graft_entry__.build())")
    lib = C.CDLL(native.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="not found"):
        native.load(str(tmp_path / "libnicamhip.so"))


def test_layout_round_trip():
    rng = np.random.default_rng(0)
    g, kall, L, nv = 6, 5, 10, 3
    var = rng.standard_normal((g, g, kall, L, nv))
    var_pl = rng.standard_normal((6, kall, 2, nv))
    flat = layout.pack(var, var_pl)
    assert flat.shape == (nv, layout.field_size(g, kall, L))
    # regional point (i, j, k, l) sits at ((l*kall + k)*g + j)*g + i
    i, j, k, l = 2, 3, 4, 7
    assert flat[1, ((l * kall + k) * g + j) * g + i] == var[i, j, k, l, 1]
    R = L * kall * g * g
    assert flat[2, R + (1 * kall + 3) * 6 + 5] == var_pl[5, 3, 1, 2]
    a, b = layout.unpack(flat, g, kall, L)
    assert np.array_equal(a, var) and np.array_equal(b, var_pl)


def test_create_rejects_bad_dims(hostemu_lib):
    lib = native.load(hostemu_lib)
    h = C.c_void_p()
    cst = native.NhConsts(*([1.0] * 11))
    bad = native.NhDims(10, 12, 0, 10, 10)      # kmin must be 1
    rc = lib.nh_create(C.byref(bad), C.byref(cst), 0, C.byref(h))
    assert rc == -1 and b"dimensions" in lib.nh_last_error()
    big = native.NhDims(10, 100, 1, 98, 10)     # kall above the column buffers
    assert lib.nh_create(C.byref(big), C.byref(cst), 0, C.byref(h)) == -1
    odd = native.NhDims(8, 12, 1, 10, 10)       # a region 6 points wide: no icosahedral grid
    rc = lib.nh_create(C.byref(odd), C.byref(cst), 0, C.byref(h))
    assert rc == -1 and b"power of two" in lib.nh_last_error()


@pytest.fixture(scope="module")
def emu_ctx(case_gl04, hostemu_lib):
    ctx = hipdyn.HipContext(case_gl04.geo, case_gl04.const, device="cpu", lib_path=hostemu_lib)
    yield ctx
    ctx.close()


def test_step_requires_a_tracer_field(emu_ctx):
    lib = emu_ctx.lib
    prog = emu_ctx.empty(6)
    rc = lib.nh_dyn_step(emu_ctx.handle, C.c_void_p(prog.data_ptr()), C.c_void_p(None), 0, 1200.0, 6,
                         C.c_void_p(None))
    assert rc == -1
    with pytest.raises(RuntimeError, match="NH_ERR_ARG"):
        native.check(lib, rc, "nh_dyn_step")


def test_host_build_operators_bit_exact(case_gl04, emu_ctx):
    import torch
    from oracle import dynamics as od
    geo = case_gl04.geo
    t = geo.topo
    g, kall, L = t.gall_1d, t.kall, t.lall
    ctx = od.OracleContext(geo, case_gl04.const, None)
    rng = np.random.default_rng(3)
    f = rng.standard_normal((g, g, kall, L))
    fp = rng.standard_normal((6, kall, 2))
    t.halo_fill_host(f, fp)
    T = torch.from_numpy(layout.pack1(f, fp))
    lo, lop = od.laplacian(ctx, f, fp)
    a, ap = layout.unpack1(emu_ctx.OPRT_laplacian(T).numpy(), g, kall, L)
    s = slice(1, g - 1)
    assert np.array_equal(a[s, s], lo[s, s]) and np.array_equal(ap[0], lop[0])
    go, gop = od.gradient(ctx, f, fp)
    G = emu_ctx.OPRT_gradient(T).numpy()
    for d in range(3):
        a, ap = layout.unpack1(G[d], g, kall, L)
        assert np.array_equal(a[s, s], go[s, s, ..., d]) and np.array_equal(ap[0], gop[0, ..., d])
    vx, vy, vz = (rng.standard_normal((g, g, kall, L)) for _ in range(3))
    vxp, vyp, vzp = (rng.standard_normal((6, kall, 2)) for _ in range(3))
    do, dop = od.divergence(ctx, vx, vy, vz, vxp, vyp, vzp)
    D = emu_ctx.divergence(*(torch.from_numpy(layout.pack1(a, b)) for a, b in ((vx, vxp), (vy, vyp), (vz, vzp))))
    a, ap = layout.unpack1(D.numpy(), g, kall, L)
    assert np.array_equal(a[s, s], do[s, s]) and np.array_equal(ap[0], dop[0])


def test_host_build_step_matches_oracle(case_gl04, state_gl04_tracers, emu_ctx):
    """Two large steps (dynamics + tracer transport) of the kernel source, host build, against the oracle."""
    from oracle import dynamics as od
    c = case_gl04
    t = c.topo
    ref = od.OracleDynamics(od.OracleContext(c.geo, c.const, None), state_gl04_tracers)
    dyn = hipdyn.HipDynamics(emu_ctx, state_gl04_tracers)
    for _ in range(2):
        ref.step()
        dyn.dynamics_step()
    got = dyn.state()
    for v in range(6):
        a, b = interior(ref.PROG[..., v], t), interior(got["PROG"][..., v], t)
        ap, bp = pole_centre(ref.PROG_pl[..., v], t), pole_centre(got["PROG_pl"][..., v], t)
        # rhogw is measured against its own (small) peak; round-off of the terrain terms enters there
        assert rel_err(a, b, ap, bp) < 1e-10, v
    for q in range(1, got["PROGq"].shape[-1]):
        a, b = interior(ref.PROGq[..., q], t), interior(got["PROGq"][..., q], t)
        ap, bp = pole_centre(ref.PROGq_pl[..., q], t), pole_centre(got["PROGq_pl"][..., q], t)
        assert rel_err(a, b, ap, bp) < 1e-12, q


def _moist(state, seed=3):
    """The gl04 state with a positive, varying qv."""
    rng = np.random.default_rng(seed)
    s = {k: v.copy() for k, v in state.items()}
    s["PROGq"][..., 0] = s["PROG"][..., 0] * 1e-3 * (1.0 + 0.5 * rng.random(s["PROG"].shape[:-1]))
    s["PROGq_pl"][..., 0] = s["PROG_pl"][..., 0] * 1e-3
    return s


def test_host_build_budget_matches_oracle(case_gl04, state_gl04_tracers, emu_ctx):
    from oracle import budget as ob
    from oracle import dynamics as od
    c = case_gl04
    s = _moist(state_gl04_tracers)
    want = ob.store_var(od.OracleContext(c.geo, c.const, None), s["PROG"], s["PROGq"], c.const.LHV)
    got = hipdyn.HipDynamics(emu_ctx, s).budget()
    assert set(got) == set(want)
    for key, v in want.items():
        assert abs(got[key] - v) <= 1e-13 * abs(v), (key, got[key], v)


def test_budget_rejects_missing_arguments(emu_ctx):
    lib = emu_ctx.lib
    prog = emu_ctx.empty(6)
    out = (C.c_double * 6)()
    rc = lib.nh_budget(emu_ctx.handle, C.c_void_p(prog.data_ptr()), C.c_void_p(None), 1, 2.5e6, out,
                       C.c_void_p(None))
    assert rc == -1 and b"nh_budget" in lib.nh_last_error()


DEFORM = (12 * 86400.0, 40.0, 0.1, 10000.0)


@pytest.mark.parametrize("wind,params", [("dcmip11", None), ("deform", DEFORM)])
def test_host_build_offline_tracer_matches_oracle(case_gl04, state_gl04_tracers, emu_ctx, wind, params):
    """Prescribed winds and three TRCADV steps of the kernel source, host build, against the oracle."""
    from oracle import dynamics as od
    from oracle import trcadv as ot
    c = case_gl04
    t = c.topo
    ref = ot.OfflineTracer(od.OracleContext(c.geo, c.const, None), state_gl04_tracers, wind, params)
    run = hipdyn.HipOfflineTracer(emu_ctx, state_gl04_tracers, wind, params)
    got = run.state()
    # momentum components against the largest momentum component (a component can vanish identically)
    mom = max(np.abs(ref.PROG[..., 1:4]).max(), np.abs(ref.PROG_pl[..., 1:4]).max())
    for v in range(1, 4):
        e = max(np.abs(ref.PROG[..., v] - got["PROG"][..., v]).max(),
                np.abs(ref.PROG_pl[..., v] - got["PROG_pl"][..., v]).max())
        assert e < 1e-13 * mom, v
    assert rel_err(ref.PROG[..., 4], got["PROG"][..., 4], ref.PROG_pl[..., 4], got["PROG_pl"][..., 4]) < 1e-13
    for _ in range(3):
        ref.step()
        run.dynamics_step()
    got = run.state()
    assert run.time == ref.time
    for q in range(1, got["PROGq"].shape[-1]):
        a, b = interior(ref.PROGq[..., q], t), interior(got["PROGq"][..., q], t)
        ap, bp = pole_centre(ref.PROGq_pl[..., q], t), pole_centre(got["PROGq_pl"][..., q], t)
        assert rel_err(a, b, ap, bp) < 1e-12, q


def test_host_build_offline_tracer_without_limiters_matches_oracle(case_gl04, state_gl04_tracers, hostemu_lib):
    """THUBURN_LIM off: unlimited face values in both directions (mod_src_tracer.py:789,
    the vertical switch likewise), two TRCADV steps against the oracle."""
    import copy
    from oracle import dynamics as od
    from oracle import trcadv as ot
    c = case_gl04
    t = c.topo
    const = copy.copy(c.const)
    const.THUBURN_LIM = False
    ref = ot.OfflineTracer(od.OracleContext(c.geo, const, None), state_gl04_tracers, "deform", DEFORM)
    ctx = hipdyn.HipContext(c.geo, const, device="cpu", lib_path=hostemu_lib)
    try:
        run = hipdyn.HipOfflineTracer(ctx, state_gl04_tracers, "deform", DEFORM)
        for _ in range(2):
            ref.step()
            run.dynamics_step()
        got = run.state()
    finally:
        ctx.close()
    for q in range(1, got["PROGq"].shape[-1]):
        a, b = interior(ref.PROGq[..., q], t), interior(got["PROGq"][..., q], t)
        ap, bp = pole_centre(ref.PROGq_pl[..., q], t), pole_centre(got["PROGq_pl"][..., q], t)
        assert rel_err(a, b, ap, bp) < 1e-12, q


def test_trcadv_winds_argument_errors(emu_ctx):
    lib = emu_ctx.lib
    prog = emu_ctx.empty(6)
    bad = (C.c_double * 4)(0.0, 40.0, 0.1, 1e4)
    assert lib.nh_trcadv_winds(emu_ctx.handle, C.c_void_p(prog.data_ptr()), 0.0, 1, bad, C.c_void_p(None)) == -1
    assert lib.nh_trcadv_winds(emu_ctx.handle, C.c_void_p(prog.data_ptr()), 0.0, 7, None, C.c_void_p(None)) == -1
    with pytest.raises(ValueError):
        hipdyn.HipOfflineTracer(emu_ctx, {"PROG": np.zeros(1)}, "nope")


# Mixed form (float fields, FP64 solve) against the FP64 oracle: largest error
# relative to the largest value of each prognostic.  Measured on the host build
# after two gl04 steps: rhog 2.7e-7, rhogvx/vy 1.0e-6, rhogvz 1.4e-4 (a small
# component: meridional flow only), rhogw 4.1e-5, rhoge 2.1e-7; tracers 2.6e-7.
MIXED_TOL = {0: 1e-5, 1: 3e-5, 2: 3e-5, 3: 3e-3, 4: 1e-3, 5: 1e-5}
MIXED_QTOL = 1e-5


@pytest.fixture(scope="module")
def emu_mixed(case_gl04, hostemu_lib):
    ctx = hipdyn.HipContext(case_gl04.geo, case_gl04.const, device="cpu", lib_path=hostemu_lib, precision="mixed")
    yield ctx
    ctx.close()


def test_host_build_mixed_step_close_to_oracle(case_gl04, state_gl04_tracers, emu_mixed):
    from oracle import dynamics as od
    c = case_gl04
    t = c.topo
    ref = od.OracleDynamics(od.OracleContext(c.geo, c.const, None), state_gl04_tracers)
    dyn = hipdyn.HipDynamics(emu_mixed, state_gl04_tracers)
    assert dyn.prog.dtype == torch_dtype("float32")
    for _ in range(2):
        ref.step()
        dyn.dynamics_step()
    got = dyn.state()
    for v, tol in MIXED_TOL.items():
        a, b = interior(ref.PROG[..., v], t), interior(got["PROG"][..., v].astype(np.float64), t)
        assert np.abs(a - b).max() / np.abs(a).max() < tol, v
    for q in range(1, got["PROGq"].shape[-1]):
        a, b = interior(ref.PROGq[..., q], t), interior(got["PROGq"][..., q].astype(np.float64), t)
        assert np.abs(a - b).max() / np.abs(a).max() < MIXED_QTOL, q


def test_host_build_mixed_operators(case_gl04, emu_mixed):
    """FP32 Laplacian and gradient against the FP64 oracle: relative 1e-5 of the result's range."""
    import torch
    from oracle import dynamics as od
    geo = case_gl04.geo
    t = geo.topo
    g, kall, L = t.gall_1d, t.kall, t.lall
    ctx = od.OracleContext(geo, case_gl04.const, None)
    rng = np.random.default_rng(2)
    f = rng.standard_normal((g, g, kall, L))
    fp = rng.standard_normal((6, kall, 2))
    T = torch.from_numpy(layout.pack1(f, fp)).to(torch.float32)
    s = slice(1, g - 1)
    lo, _ = od.laplacian(ctx, f, fp)
    a, _ = layout.unpack1(emu_mixed.OPRT_laplacian(T).double().numpy(), g, kall, L)
    assert np.abs(a[s, s] - lo[s, s]).max() < 1e-5 * np.abs(lo[s, s]).max()
    go, _ = od.gradient(ctx, f, fp)
    G = emu_mixed.OPRT_gradient(T).double().numpy()
    a, _ = layout.unpack1(G[0], g, kall, L)
    assert np.abs(a[s, s] - go[s, s, ..., 0]).max() < 1e-5 * np.abs(go[s, s, ..., 0]).max()


def test_entry_point_must_match_the_context_precision(emu_ctx, emu_mixed):
    lib = emu_ctx.lib
    p64, p32 = emu_ctx.empty(6), emu_mixed.empty(6)
    rc = lib.nh_dyn_step_f32(emu_ctx.handle, C.c_void_p(p64.data_ptr()), C.c_void_p(p64.data_ptr()), 1, 1200.0, 6,
                             C.c_void_p(None))
    assert rc == -2 and b"_f32" in lib.nh_last_error()
    rc = lib.nh_dyn_step(emu_mixed.handle, C.c_void_p(p32.data_ptr()), C.c_void_p(p32.data_ptr()), 1, 1200.0, 6,
                         C.c_void_p(None))
    assert rc == -2
    assert lib.nh_precision(emu_ctx.handle) == 0 and lib.nh_precision(emu_mixed.handle) == 1
    with pytest.raises(ValueError):
        hipdyn.HipContext(None, None, device="cpu", precision="half")


def torch_dtype(name):
    import torch
    return getattr(torch, name)


def test_backend_selection(hostemu_lib):
    """Backend.configure (mod_backend.py:175-283) with the 'hip' branch: both forms, loud failures."""
    backend = mod("backend")
    b = backend.Backend().configure("hip", "float64", lib_path=hostemu_lib)
    assert b.type == "hip" and b.precision == "float64" and b.ndtype == np.float64
    b = backend.Backend().configure("hip", "float32", lib_path=hostemu_lib)
    assert b.precision == "mixed" and b.ndtype == np.float32
    with pytest.raises(ValueError):
        backend.Backend().configure("cuda", "float64")
    with pytest.raises(ValueError):
        backend.Backend().configure("hip", "float16", lib_path=hostemu_lib)


def test_host_build_gl06_step_multitile_matches_oracle(hostemu_lib):
    """One gl06 z40 FP64 step of the kernel source (host build) against the oracle's step
    (the fingerprint tests/golden/make_step_fixtures.py wrote).  At gl06 a region is 32
    points wide: the column tiles run 32 x 8 tiles, four rows of them per region, so their
    multi-tile paths are exercised (gl04 and gl05 regions fit one tile)."""
    from conftest import GOLDEN
    fx = np.load(os.path.join(GOLDEN, "oracle_gl06_z40_1steps.npz"))
    case = mod("setup").build_case(glevel=int(fx["glevel"]), vlayer=int(fx["vlayer"]), dtl=float(fx["dtl"]),
                                   seed=int(fx["seed"]))
    t = case.topo
    ctx = hipdyn.HipContext(case.geo, case.const, device="cpu", lib_path=hostemu_lib)
    try:
        dyn = hipdyn.HipDynamics(ctx, case.state)
        dyn.dynamics_step()
        st = dyn.state()
    finally:
        ctx.close()
    P, Pp = st["PROG"], st["PROG_pl"]
    i, j, k, l = fx["idx"]
    ks = slice(t.kmin, t.kmax + 1)
    for v in range(6):
        a = fx["sample"][v]
        e = max(np.abs(P[i, j, k, l, v] - a).max(), np.abs(Pp[0, ks, :, v] - fx["pole"][..., v]).max())
        assert e <= 1e-10 * np.abs(a).max(), (v, e / np.abs(a).max())
    g = t.gall_1d
    rms = np.sqrt(np.mean(P[1:g - 1, 1:g - 1, ks][..., :6] ** 2, axis=(0, 1, 3)))
    assert np.abs(rms - fx["level_rms"]).max() <= 1e-12 * fx["level_rms"].max()
