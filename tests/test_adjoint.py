"""Adjoint of the dynamics step: finite-difference checks, the mixed form
against FP64, the surface-pressure diagnostic, argument errors.

The reference has no adjoint and no gradient test: what pins the transposed
kernels is the derivative itself, i.e. central finite differences of the
forward step (the kernel source of the product, host build on the CPU, the
gfx950 build on the GPU) along seeded random directions.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import interior, mod

hipdyn = mod("hipdyn")
adjoint = mod("adjoint")
native = mod("native")


def _scales(prog):
    """Per-variable direction scales: the field peak, the horizontal-momentum peak for the
    (initially zero) vertical components."""
    s = prog.abs().amax(dim=1, keepdim=True).clone()
    s[3] = s[1]
    s[4] = 0.01 * s[1]
    return s


def _fd_check(dyn, nsteps, directions, eps, seed=1, hold_q=False):
    """(fd, adjoint) pairs of J = sum(w ps^2) after nsteps along seeded random directions."""
    import torch
    rng = np.random.default_rng(seed)
    w = torch.from_numpy(rng.random(dyn.ctx.ncol)).to(dyn.ctx.device)
    x0, q0 = dyn.prog.clone(), dyn.progq.clone()

    def J(x):
        dyn.prog.copy_(x)
        dyn.progq.copy_(q0)
        for _ in range(nsteps):
            dyn.dynamics_step()
        ps = dyn.sfc_pressure(progq=q0 if hold_q else None)
        return float((w * ps * ps).sum())

    dyn.prog.copy_(x0)
    dyn.progq.copy_(q0)
    traj = []
    for _ in range(nsteps):
        traj.append((dyn.prog.clone(), dyn.progq.clone()))
        dyn.dynamics_step()
    qJ = q0 if hold_q else dyn.progq
    ps = dyn.sfc_pressure(progq=qJ)
    gb = torch.zeros_like(dyn.prog)
    dyn.sfc_pressure_adj(dyn.prog, qJ, 2 * w * ps, gb)
    for p, q in reversed(traj):
        dyn.dynamics_step_adj(p, q, gb)
    sc = _scales(x0)
    out = []
    for mask in directions:
        v = torch.from_numpy(rng.standard_normal(tuple(x0.shape))).to(x0.dtype).to(x0.device) * sc
        if mask is not None:
            m = torch.zeros((6, 1), dtype=x0.dtype, device=x0.device)
            m[mask] = 1
            v = v * m
        fd = (J(x0 + eps * v) - J(x0 - eps * v)) / (2 * eps)
        out.append((fd, float((gb.double() * v.double()).sum())))
    dyn.prog.copy_(x0)
    dyn.progq.copy_(q0)
    return out


# ----------------------------------------------------------------------------
# host build (CPU)
# ----------------------------------------------------------------------------
@pytest.fixture(scope="module")
def emu64(case_gl04, hostemu_lib):
    ctx = hipdyn.HipContext(case_gl04.geo, case_gl04.const, device="cpu", lib_path=hostemu_lib)
    yield ctx
    ctx.close()


def test_host_adjoint_matches_finite_differences_per_variable(case_gl04, emu64):
    """One step, FP64: each prognostic variable alone, relative agreement 1e-6
    (the finite differences themselves are good to about 1e-9 at this step size)."""
    dyn = hipdyn.HipDynamics(emu64, case_gl04.state)
    for q, (fd, ad) in enumerate(_fd_check(dyn, 1, list(range(6)), eps=1e-4)):
        assert abs(fd) > 0 and abs(fd - ad) <= 1e-6 * abs(fd), (q, fd, ad)


def test_host_adjoint_three_steps_with_strong_damping(case_gl04, hostemu_lib):
    """Three steps with hyperdiffusion and divergence damping 300x stronger, so that their
    transposes carry a visible share of the gradient."""
    import copy
    const = copy.copy(case_gl04.const)
    const.GAMMA_H = 3.0e18
    const.ALPHA_D = 3.0e18
    ctx = hipdyn.HipContext(case_gl04.geo, const, device="cpu", lib_path=hostemu_lib)
    try:
        dyn = hipdyn.HipDynamics(ctx, case_gl04.state)
        for fd, ad in _fd_check(dyn, 3, [None, None], eps=1e-5, seed=2):
            assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
    finally:
        ctx.close()


def test_host_adjoint_moist_thermodynamics(case_gl04, emu64):
    """Nonzero qv enters the step through q = rhogq / rhog (cv, gas constant); with the
    tracer held at its input (as the transpose holds it) the derivative is exact."""
    state = {k: v.copy() for k, v in case_gl04.state.items()}
    rng = np.random.default_rng(5)
    state["PROGq"][..., 0] = state["PROG"][..., 0] * 0.01 * rng.random(state["PROGq"].shape[:-1])
    state["PROGq_pl"][..., 0] = state["PROG_pl"][..., 0] * 0.01
    dyn = hipdyn.HipDynamics(emu64, state)
    for fd, ad in _fd_check(dyn, 1, [None, 0, 5], eps=1e-5, seed=3, hold_q=True):
        assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)


def test_host_autograd_window_mixed_matches_fp64(case_gl04, hostemu_lib):
    """J = sum over the 30-50N, 0-20E box of ps^2 after three steps, through
    torch.autograd: the mixed form's gradient against the FP64 form's."""
    import torch
    grads = {}
    for prec in ("float64", "mixed"):
        ctx = hipdyn.HipContext(case_gl04.geo, case_gl04.const, device="cpu", lib_path=hostemu_lib,
                                precision=prec)
        try:
            dyn = hipdyn.HipDynamics(ctx, case_gl04.state)
            m = adjoint.box_mask(dyn)
            assert m.sum() > 0
            x = dyn.prog.clone().requires_grad_(True)
            ps = adjoint.sfc_pressure(dyn, adjoint.forecast(dyn, x, 3))
            (m * ps * ps).sum().backward()
            grads[prec] = x.grad.double()
        finally:
            ctx.close()
    g64, g32 = grads["float64"], grads["mixed"]
    for q in range(6):
        sc = g64[q].abs().max()
        assert sc > 0 and (g64[q] - g32[q]).abs().max() <= 1e-4 * sc, q


def test_host_taped_step_equals_plain_step(case_gl04, emu64):
    """The taped forward step writes its intermediates to the tape instead of the
    workspace: the state it produces is bit-identical."""
    import torch
    a = hipdyn.HipDynamics(emu64, case_gl04.state)
    b = hipdyn.HipDynamics(emu64, case_gl04.state)
    tape = b.new_tape()
    for _ in range(2):
        a.dynamics_step()
        b.dynamics_step_tape(tape)
    assert torch.equal(a.prog, b.prog) and torch.equal(a.progq, b.progq)


@pytest.mark.parametrize("store", ["tape", "stages"])
def test_host_autograd_window_equals_direct_adjoint(case_gl04, emu64, store):
    import torch
    dyn = hipdyn.HipDynamics(emu64, case_gl04.state)
    m = adjoint.box_mask(dyn)
    x = dyn.prog.clone().requires_grad_(True)
    ps = adjoint.sfc_pressure(dyn, adjoint.forecast(dyn, x, 2, store=store))
    (m * ps * ps).sum().backward()
    # the same by hand
    traj = []
    for _ in range(2):
        traj.append((dyn.prog.clone(), dyn.progq.clone()))
        dyn.dynamics_step()
    ps2 = dyn.sfc_pressure()
    gb = torch.zeros_like(dyn.prog)
    dyn.sfc_pressure_adj(dyn.prog, dyn.progq, 2 * m * ps2, gb)
    for p, q in reversed(traj):
        dyn.dynamics_step_adj(p, q, gb)
    assert torch.equal(ps.detach(), ps2)
    assert torch.allclose(x.grad, gb, rtol=0, atol=1e-12 * float(gb.abs().max()))


def test_host_sfc_pressure_matches_oracle(case_gl04, state_gl04_tracers, emu64):
    from oracle import budget as ob
    from oracle import dynamics as od
    c = case_gl04
    t = c.topo
    dyn = hipdyn.HipDynamics(emu64, state_gl04_tracers)
    ps = dyn.sfc_pressure().numpy()
    octx = od.OracleContext(c.geo, c.const, None)
    _, ref = ob.sfc_pressure(octx, state_gl04_tracers["PROG"], state_gl04_tracers["PROGq"])
    g = t.gall_1d
    got = np.transpose(ps[:t.lall * g * g].reshape(t.lall, g, g), (2, 1, 0))
    s = slice(1, g - 1)
    assert np.max(np.abs(got[s, s] - ref[s, s])) <= 1e-12 * np.max(np.abs(ref))


def test_adjoint_argument_errors(emu64):
    lib = emu64.lib
    prog = emu64.empty(6)
    q = emu64.empty(1)
    p = [C.c_void_p(a.data_ptr()) for a in (prog, q)]
    rc = lib.nh_dyn_step_adj(emu64.handle, p[0], p[1], 1, 1200.0, 6, C.c_void_p(None), C.c_void_p(None))
    assert rc == -1 and b"nh_dyn_step_adj" in lib.nh_last_error()
    rc = lib.nh_dyn_step_adj(emu64.handle, p[0], p[1], 1, 1200.0, 4, p[0], C.c_void_p(None))
    assert rc == -1
    rc = lib.nh_dyn_step_adj_f32(emu64.handle, p[0], p[1], 1, 1200.0, 6, p[0], C.c_void_p(None))
    assert rc == -2 and b"without _f32" in lib.nh_last_error()
    rc = lib.nh_sfc_pressure_adj(emu64.handle, p[0], p[1], C.c_void_p(None), p[0], C.c_void_p(None))
    assert rc == -1


def test_adjoint_wrapper_argument_errors(case_gl04, emu64):
    """The Python entry points refuse tensors of the wrong shape, dtype or size before any
    pointer reaches the library."""
    import torch
    dyn = hipdyn.HipDynamics(emu64, case_gl04.state)
    prog, q = dyn.prog.clone(), dyn.progq.clone()
    five = emu64.empty(5)
    with pytest.raises(ValueError, match="stage_in"):
        dyn.dynamics_step_adj_rec(emu64.empty(17), q, prog.clone())
    with pytest.raises(ValueError, match="6 fields"):
        dyn.dynamics_step_adj_rec(emu64.empty(18), q, five)
    with pytest.raises(ValueError, match="tape"):
        dyn.dynamics_step_adj_tape(torch.empty(16, dtype=torch.uint8), q, prog.clone())
    with pytest.raises(ValueError, match="tape"):
        dyn.dynamics_step_adj_tape(torch.empty(dyn.tape_bytes(), dtype=torch.float64), q, prog.clone())
    with pytest.raises(ValueError, match="6 fields"):
        dyn.dynamics_step_adj(five, q, prog.clone())
    with pytest.raises(ValueError):
        dyn.sfc_pressure(prog.float())
    with pytest.raises(ValueError, match="6 fields"):
        dyn.sfc_pressure(five)
    with pytest.raises(ValueError, match="ps_b"):
        dyn.sfc_pressure_adj(prog, q, torch.zeros(emu64.ncol - 1, dtype=torch.float64), prog.clone())
    with pytest.raises(ValueError):
        dyn.sfc_pressure_adj(prog, q, torch.zeros(emu64.ncol, dtype=torch.float64), prog[:, :-1].clone())


def test_adjoint_refuses_a_halo_map_it_cannot_transpose(case_gl04, hostemu_lib):
    """A destination that is also a source would make the one-pass scatter order-dependent."""
    ctx = hipdyn.HipContext(case_gl04.geo, case_gl04.const, device="cpu", lib_path=hostemu_lib)
    try:
        lib = ctx.lib
        dst = np.array([0, 1], dtype=np.int32)
        src = np.array([1, 2], dtype=np.int32)
        ip = C.POINTER(C.c_int32)
        native.check(lib, lib.nh_upload_halo_map(ctx.handle, dst.ctypes.data_as(ip), src.ctypes.data_as(ip), 2))
        prog, q = ctx.empty(6).zero_(), ctx.empty(1).zero_()
        rc = lib.nh_dyn_step_adj(ctx.handle, C.c_void_p(prog.data_ptr()), C.c_void_p(q.data_ptr()), 1, 1200.0,
                                 6, C.c_void_p(prog.data_ptr()), C.c_void_p(None))
        assert rc == -2 and b"halo map" in lib.nh_last_error()
    finally:
        ctx.close()


# ----------------------------------------------------------------------------
# tracer transport (offline tracer step), host build
# ----------------------------------------------------------------------------
def _rough_tracers(state, nq=2, amp=0.3, seed=0):
    """Tracers with mixing ratios 1 + amp * U(0, 1) per point: no two neighbours tie, so
    every min/max/clip of the limiters has a unique winner."""
    st = {k: v.copy() for k, v in state.items()}
    rng = np.random.default_rng(seed)
    shp, shp_pl = st["PROG"].shape[:-1] + (nq,), st["PROG_pl"].shape[:-1] + (nq,)
    st["PROGq"] = st["PROG"][..., :1] * (1.0 + amp * rng.random(shp))
    st["PROGq_pl"] = st["PROG_pl"][..., :1] * (1.0 + amp * rng.random(shp_pl))
    return st


def _trc_run(ctx, state, w0=3.0):
    """Deformational flow with vertical motion w0 [m/s] at its initial time."""
    return hipdyn.HipOfflineTracer(ctx, _rough_tracers(state), "deform", (12 * 86400.0, 40.0, w0, 10000.0))


def _trc_step(run, prog, q):
    q = q.clone()
    ctx = run.ctx
    native.check(ctx.lib, ctx.fn("nh_tracer_step")(ctx.handle, C.c_void_p(prog.data_ptr()), C.c_void_p(q.data_ptr()),
                                                  q.shape[0], run.dtl, ctx.stream()), "nh_tracer_step")
    return q


def _column_directions(run, ncols, rng):
    """Random perturbations of all levels of ``ncols`` random columns, per tracer at its peak."""
    ctx = run.ctx
    q0 = run.progq
    g, ka = ctx.geo.topo.gall_1d, ctx.geo.topo.kall
    nreg = ctx.ncol - 12
    sc = q0.abs().amax(dim=1, keepdim=True)
    for c in rng.integers(0, ctx.ncol, size=ncols):
        if c < nreg:
            l, r = divmod(int(c), g * g)
            idx = [(l * ka + k) * g * g + r for k in range(ka)]
        else:
            lp, v = divmod(int(c) - nreg, 6)
            idx = [nreg * ka + (lp * ka + k) * 6 + v for k in range(ka)]
        d = q0.new_zeros(q0.shape)
        d[:, idx] = q0.new_tensor(rng.standard_normal((q0.shape[0], ka))) * sc
        yield d


def test_host_tracer_adjoint_exact_without_limiter(case_gl04, hostemu_lib):
    """Limiters off: the step is linear in the tracers and J = sum(w q^2) quadratic, so a
    central difference of any size is exact; agreement to rounding (1e-10)."""
    import copy
    import torch
    const = copy.copy(case_gl04.const)
    const.THUBURN_LIM = False
    ctx = hipdyn.HipContext(case_gl04.geo, const, device="cpu", lib_path=hostemu_lib)
    try:
        run = _trc_run(ctx, case_gl04.state, w0=0.1)
        prog, q0 = run.prog.clone(), run.progq.clone()
        rng = np.random.default_rng(11)
        w = torch.from_numpy(rng.random(tuple(q0.shape)))
        out = _trc_step(run, prog, q0)
        gb = 2 * w * out
        run.tracer_step_adj(prog, q0, gb)
        sc = q0.abs().amax(dim=1, keepdim=True)
        J = lambda q: float((w * _trc_step(run, prog, q) ** 2).sum())
        for _ in range(2):
            v = torch.from_numpy(rng.standard_normal(tuple(q0.shape))) * sc
            fd = (J(q0 + 1e-2 * v) - J(q0 - 1e-2 * v)) / 2e-2
            ad = float((gb * v).sum())
            assert abs(fd - ad) <= 1e-10 * abs(fd), (fd, ad)
    finally:
        ctx.close()


def test_host_tracer_adjoint_with_limiters(case_gl04, hostemu_lib):
    """Thuburn limiters on, horizontal and vertical: the step is piecewise linear, so the
    check perturbs one column at a time (all levels, both tracers) and dots the field
    difference with w, which keeps rounding local; agreement 1e-5 on 24 columns."""
    import torch
    ctx = hipdyn.HipContext(case_gl04.geo, case_gl04.const, device="cpu", lib_path=hostemu_lib)
    try:
        run = _trc_run(ctx, case_gl04.state)
        prog, q0 = run.prog.clone(), run.progq.clone()
        rng = np.random.default_rng(12)
        w = torch.from_numpy(rng.random(tuple(q0.shape)))
        gb = w.clone()
        run.tracer_step_adj(prog, q0, gb)
        eps = 1e-6
        for v in _column_directions(run, 24, rng):
            fd = float((w * (_trc_step(run, prog, q0 + eps * v) - _trc_step(run, prog, q0 - eps * v))).sum()) / (2 * eps)
            ad = float((gb * v).sum())
            assert abs(fd - ad) <= 1e-5 * abs(fd), (fd, ad)
    finally:
        ctx.close()


def test_host_tracer_forecast_autograd_and_mixed(case_gl04, hostemu_lib):
    """Three steps through torch.autograd equal three direct adjoint calls (to summation
    order); limiters off,
    the mixed form's gradient agrees with FP64's to 1e-5 of its peak."""
    import copy
    import torch
    const = copy.copy(case_gl04.const)
    const.THUBURN_LIM = False
    grads = {}
    for prec in ("float64", "mixed"):
        ctx = hipdyn.HipContext(case_gl04.geo, const, device="cpu", lib_path=hostemu_lib, precision=prec)
        try:
            run = _trc_run(ctx, case_gl04.state, w0=0.1)
            w = torch.from_numpy(np.random.default_rng(13).random(tuple(run.progq.shape))).to(ctx.dtype)
            q = run.progq.clone().requires_grad_(True)
            t0, p0 = run.time, run.prog.clone()
            (w * adjoint.tracer_forecast(run, q, 3)).sum().backward()
            assert run.time == t0 and torch.equal(run.prog, p0)
            grads[prec] = q.grad.double()
            if prec == "float64":
                traj = []
                for _ in range(3):
                    traj.append((run.prog.clone(), run.progq.clone()))
                    run.dynamics_step()
                gb = w.clone()
                for p, qq in reversed(traj):
                    run.tracer_step_adj(p, qq, gb)
                # the scatter transposes add atomically: equal up to summation order
                assert torch.allclose(gb, q.grad, rtol=0, atol=1e-12 * float(gb.abs().max()))
        finally:
            ctx.close()
    sc = grads["float64"].abs().max()
    assert (grads["float64"] - grads["mixed"]).abs().max() <= 1e-5 * sc


def test_tracer_adjoint_argument_errors(emu64):
    lib = emu64.lib
    prog, q = emu64.empty(6), emu64.empty(1)
    p = [C.c_void_p(a.data_ptr()) for a in (prog, q)]
    assert lib.nh_tracer_step_adj(emu64.handle, p[0], p[1], 1, 1200.0, C.c_void_p(None), C.c_void_p(None)) == -1
    assert b"nh_tracer_step_adj" in lib.nh_last_error()
    assert lib.nh_tracer_step_adj(emu64.handle, p[0], p[1], 0, 1200.0, p[1], C.c_void_p(None)) == -1
    assert lib.nh_tracer_step_adj(emu64.handle, p[0], p[1], 1, 0.0, p[1], C.c_void_p(None)) == -1
    assert lib.nh_tracer_step_adj_f32(emu64.handle, p[0], p[1], 1, 1200.0, p[1], C.c_void_p(None)) == -2


# ----------------------------------------------------------------------------
# gfx950 build
# ----------------------------------------------------------------------------
@pytest.mark.gpu
def test_gpu_adjoint_matches_finite_differences(case_gl05):
    """gl05 z40, two steps, FP64: J = sum(w ps^2) along three random directions over all
    six prognostic variables, relative agreement 1e-6."""
    print("gpu adjoint fd: gl05 z40 float64, 2 steps", flush=True)
    ctx = hipdyn.HipContext(case_gl05.geo, case_gl05.const, device="cuda:0")
    try:
        dyn = hipdyn.HipDynamics(ctx, case_gl05.state)
        for fd, ad in _fd_check(dyn, 2, [None, None, None], eps=1e-5, seed=4):
            assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
    finally:
        ctx.close()


@pytest.mark.gpu
def test_gpu_adjoint_mixed_matches_fp64(case_gl05):
    """gl05 z40, three steps: the box functional's gradient, mixed form against FP64."""
    import torch
    print("gpu adjoint mixed vs fp64: gl05 z40, 3 steps", flush=True)
    grads = {}
    for prec in ("float64", "mixed"):
        ctx = hipdyn.HipContext(case_gl05.geo, case_gl05.const, device="cuda:0", precision=prec)
        try:
            dyn = hipdyn.HipDynamics(ctx, case_gl05.state)
            m = adjoint.box_mask(dyn)
            x = dyn.prog.clone().requires_grad_(True)
            ps = adjoint.sfc_pressure(dyn, adjoint.forecast(dyn, x, 3))
            (m * ps * ps).sum().backward()
            torch.cuda.synchronize()
            grads[prec] = x.grad.double().cpu()
        finally:
            ctx.close()
    g64, g32 = grads["float64"], grads["mixed"]
    for q in range(6):
        sc = g64[q].abs().max()
        assert sc > 0 and (g64[q] - g32[q]).abs().max() <= 1e-3 * sc, q


@pytest.mark.gpu
def test_gpu_tracer_adjoint_matches_finite_differences(case_gl05):
    """gl05 z40, the offline tracer step on gfx950: limiters off, global directions (exact
    central differences, 1e-10); limiters on, 12 single-column directions (1e-5)."""
    import copy
    import torch
    print("gpu tracer adjoint fd: gl05 z40 float64", flush=True)
    for lim in (False, True):
        const = copy.copy(case_gl05.const)
        const.THUBURN_LIM = lim
        ctx = hipdyn.HipContext(case_gl05.geo, const, device="cuda:0")
        try:
            run = _trc_run(ctx, case_gl05.state, w0=3.0 if lim else 0.1)
            prog, q0 = run.prog.clone(), run.progq.clone()
            rng = np.random.default_rng(21)
            w = torch.from_numpy(rng.random(tuple(q0.shape))).to(q0.device)
            if not lim:
                gb = 2 * w * _trc_step(run, prog, q0)
            else:
                gb = w.clone()
            run.tracer_step_adj(prog, q0, gb)
            torch.cuda.synchronize()
            if not lim:
                sc = q0.abs().amax(dim=1, keepdim=True)
                J = lambda q: float((w * _trc_step(run, prog, q) ** 2).sum())
                v = torch.from_numpy(rng.standard_normal(tuple(q0.shape))).to(q0.device) * sc
                fd = (J(q0 + 1e-2 * v) - J(q0 - 1e-2 * v)) / 2e-2
                ad = float((gb * v).sum())
                assert abs(fd - ad) <= 1e-10 * abs(fd), (fd, ad)
            else:
                eps = 1e-6
                for v in _column_directions(run, 12, rng):
                    d = _trc_step(run, prog, q0 + eps * v) - _trc_step(run, prog, q0 - eps * v)
                    fd = float((w * d).sum()) / (2 * eps)
                    ad = float((gb * v).sum())
                    assert abs(fd - ad) <= 1e-5 * abs(fd), (fd, ad)
        finally:
            ctx.close()
