"""The standalone operator entry points (nh_vi_matrix, nh_vi_solve, nh_prepost,
nh_hdiffusion, nh_divdamp3d, nh_limiter_v, nh_limiter_h) against the oracle.

Each runs on seeded inputs through the C-ABI: on the host build of the kernel
source at gl04 z10 (CPU) and on the MI355X at gl05 z40 (``-m gpu``).  The
oracle routines are pinned against the reference's own kernels
(test_oracle_kernels.py); the FP64 device results must equal them bit for bit
where the two follow the same expression order, and to 1e-13 of the field's
largest value where the device fuses a stencil into tiles (hyperdiffusion,
divergence damping: the same sums, one rounding order).  The reference-signature
wrappers (ops.compute_rhow_matrix_reg, ...) are run on the same data.
"""
import types

import numpy as np
import pytest

from conftest import mod

hipdyn = mod("hipdyn")
layout = mod("layout")
ops_mod = mod("ops")


@pytest.fixture(scope="module", params=["host", pytest.param("gpu", marks=pytest.mark.gpu)])
def env(request, hostemu_lib):
    import torch
    if request.param == "host":
        case = mod("setup").build_case(glevel=4, vlayer=10, ztop=10000.0, dz_bottom=200.0, dtl=1200.0, seed=0)
        ctx = hipdyn.HipContext(case.geo, case.const, device="cpu", lib_path=hostemu_lib)
    else:
        if not torch.cuda.is_available():
            pytest.skip("no GPU")
        case = mod("setup").build_case(glevel=5, vlayer=40, dtl=1200.0, seed=0)
        ctx = hipdyn.HipContext(case.geo, case.const, device="cuda:0")
    print(f"  operators on {request.param}: gl{case.topo.gall_1d} kall {case.topo.kall}", flush=True)
    yield types.SimpleNamespace(case=case, ctx=ctx, ops=ops_mod.HipOperators(ctx), name=request.param)
    ctx.close()


def _oracle(case):
    from oracle import dynamics as od
    octx = od.OracleContext(case.geo, case.const, None)
    return od, octx, od.OracleDynamics(octx, case.state)


def _rand_field(rng, t, lo=-1.0, hi=1.0):
    return (rng.uniform(lo, hi, (t.gall_1d, t.gall_1d, t.kall, t.lall)), rng.uniform(lo, hi, (6, t.kall, 2)))


def _dev1(env, pair):
    return env.ctx.to_device(pair[0][..., None], pair[1][..., None])[0]


def _host1(env, flat):
    t = env.case.topo
    return layout.unpack1(flat.detach().cpu().numpy(), t.gall_1d, t.kall, t.lall)


def test_vi_matrix_and_solve(env):
    """nh_vi_matrix / nh_vi_solve against the oracle's _rhow_matrix / _rhow_solve
    (kernels/vimatrix.py, virhowsolver.py), bit for bit, regional and pole columns."""
    import torch
    case, t = env.case, env.case.topo
    od, octx, dyn = _oracle(case)
    rng = np.random.default_rng(11)
    dt = 200.0
    eth = _rand_field(rng, t, 2.0e5, 3.0e5)
    gt = _rand_field(rng, t, -1.0, 1.0)
    M = env.ops.vi_matrix(_dev1(env, eth), _dev1(env, gt), dt)
    ks = slice(t.kmin + 1, t.kmax + 1)
    ref = [dyn._rhow_matrix(eth[0], gt[0], dt, "", 2), dyn._rhow_matrix(eth[1], gt[1], dt, "_pl", 1)]
    for q in range(3):
        a, ap = _host1(env, M[q])
        assert np.array_equal(a[:, :, ks], ref[0][q]), q
        assert np.array_equal(ap[:, ks], ref[1][q]), q
    # the reference-signature wrappers return the same slabs
    cfg = types.SimpleNamespace(kmin=t.kmin, kmax=t.kmax)
    vg = case.geo.vgrid
    one = np.ones((t.gall_1d, t.gall_1d, t.kall, t.lall))
    got = ops_mod.compute_rhow_matrix_reg(eth[0], gt[0], one, vg.rdgzh, vg.dfact, vg.cfact, None, dt, cfg,
                                          ctx=env.ctx)
    for q in range(3):
        assert np.array_equal(got[q], ref[0][q])
    # the solve, with sources of realistic size and this matrix
    rw = _rand_field(rng, t, -1.0, 1.0)
    rw0 = _rand_field(rng, t, -1.0, 1.0)
    pre0 = _rand_field(rng, t, 1.0e4, 1.0e5)
    rho0 = _rand_field(rng, t, 0.1, 1.2)
    Sr = _rand_field(rng, t, -1e-5, 1e-5)
    Sw = _rand_field(rng, t, -1e-3, 1e-3)
    Sp = _rand_field(rng, t, -1.0, 1.0)
    ins = [rw, rw0, pre0, rho0, Sr, Sw, Sp]
    out = env.ops.vi_solve(*(_dev1(env, p) for p in ins), M, dt)
    a, ap = _host1(env, out)
    for reg, tag, kax, got_arr in ((0, "", 2, a), (1, "_pl", 1, ap)):
        r = dyn._rhow_solve(*(p[reg] for p in ins), *ref[reg], dt, tag, kax)
        assert np.array_equal(got_arr, r), tag
    # in place
    rwt = _dev1(env, rw)
    env.ops.vi_solve(rwt, *(_dev1(env, p) for p in ins[1:]), M, dt, out=rwt)
    assert torch.equal(rwt, out)
    got = ops_mod.compute_rhow_solver_reg(*(p[0] for p in ins), *ref[0], one, one, one, one, one, vg.rdgzh,
                                          vg.afact, vg.bfact, dt, cfg, ctx=env.ctx)
    assert np.array_equal(got, a)


def test_prepost_hdiffusion_divdamp(env):
    """nh_prepost (diagnostics, boundary values, eth), nh_hdiffusion and nh_divdamp3d
    against the oracle's prepost / hdiffusion / divdamp on the seeded state."""
    case, t = env.case, env.case.topo
    od, octx, dyn = _oracle(case)
    st = case.state
    prog = env.ctx.to_device(st["PROG"], st["PROG_pl"])
    progq = env.ctx.to_device(st["PROGq"], st["PROGq_pl"])
    diag, eth = env.ops.prepost(prog, progq)
    pp = dyn.prepost()
    P, Pp = env.ctx.to_host(prog)
    D, Dp = env.ctx.to_host(diag)
    e, ep = _host1(env, eth)
    # bit for bit except the top boundary rhogw (and w), whose terrain-gradient terms
    # are round-off in the oracle and exactly zero on the device: 1e-13 of the momentum
    # (velocity) scale
    mom, vel = np.abs(dyn.PROG[..., 1:4]).max(), np.abs(dyn.DIAG[..., 2:5]).max()
    for v in range(6):
        for got_, ref_, sc in ((P[..., v], dyn.PROG[..., v], mom), (Pp[..., v], dyn.PROG_pl[..., v], mom),
                               (D[..., v], dyn.DIAG[..., v], vel), (Dp[..., v], dyn.DIAG_pl[..., v], vel)):
            sc = max(np.abs(ref_).max(), sc)
            assert np.abs(got_ - ref_).max() <= 1e-13 * sc, v
    ks = slice(0, t.kmax + 1)
    assert np.array_equal(P[:, :, ks, :, :4], dyn.PROG[:, :, ks, :, :4])
    assert np.array_equal(D[..., :5], dyn.DIAG[..., :5])
    assert np.array_equal(e, pp["eth"]) and np.array_equal(ep, pp["eth_pl"])
    # hyperdiffusion tendency, every point
    T, Tp = dyn.hdiffusion(pp)
    got, got_p = env.ctx.to_host(env.ops.hdiffusion(prog, diag))
    momt = np.abs(T[..., 1:5]).max()      # the w tendency (near zero here) on the momentum scale
    for v in range(6):
        sc = max(np.abs(T[..., v]).max(), momt if 1 <= v <= 4 else 1e-300)
        assert np.abs(got[..., v] - T[..., v]).max() <= 1e-13 * sc, v
        assert np.abs(got_p[..., v] - Tp[..., v]).max() <= 1e-13 * sc, v
    # divergence damping of a divergent (seeded, halo-filled) momentum field: the
    # balanced state's own divergence is round-off
    rng = np.random.default_rng(17)
    mv = [_rand_field(rng, t, -30.0, 30.0) for _ in range(4)]
    for f in mv:
        t.halo_fill_host(f[0], f[1])
    gd, gdp = dyn.divdamp(*(f[0] for f in mv), *(f[1] for f in mv))
    G = env.ops.divdamp(*(_dev1(env, f) for f in mv))
    s, ks = slice(1, t.gall_1d - 1), slice(t.kmin, t.kmax + 1)
    for q in range(3):
        a, ap = _host1(env, G[q])
        sc = np.abs(gd[q][s, s, ks]).max()
        assert np.abs(a[s, s, ks] - gd[q][s, s, ks]).max() <= 1e-13 * sc, q
        assert np.abs(ap[0, ks] - gdp[q][0, ks]).max() <= 1e-13 * sc, q


def test_limiters(env):
    """nh_limiter_v and nh_limiter_h against the oracle's vertical_limiter and
    horizontal_limiter (kernels/verticallimiter.py, horizontallimiter.py), bit for
    bit, on seeded face values, Courant numbers and upwind masks."""
    import torch
    from oracle import tracer as ot
    case, t = env.case, env.case.topo
    octx = _oracle(case)[1]
    rng = np.random.default_rng(5)
    g = t.gall_1d
    # vertical
    q = _rand_field(rng, t, 0.0, 1.0)
    qh = _rand_field(rng, t, -0.2, 1.2)
    d = _rand_field(rng, t, -0.1, 0.1)
    ck = [_rand_field(rng, t, -0.6, 0.6) for _ in range(2)]
    ckd = torch.stack([_dev1(env, ck[0]), _dev1(env, ck[1])]).contiguous()
    qhd = _dev1(env, qh)
    env.ops.vertical_limiter(qhd, _dev1(env, q), _dev1(env, d), ckd)
    a, ap = _host1(env, qhd)
    for reg, kax, got in ((0, 2, a), (1, 1, ap)):
        ref = qh[reg].copy()
        ot.vertical_limiter(octx, ref, q[reg], d[reg], np.stack([ck[0][reg], ck[1][reg]], axis=-1), kax)
        assert np.array_equal(got, ref)
    got = ops_mod.compute_vertical_limiter(qh[0], q[0], d[0], np.stack([ck[0][0], ck[1][0]], axis=-1), None,
                                           ctx=env.ctx)
    assert np.array_equal(got, a)
    # horizontal: six faces per cell; the Courant numbers of faces 3..5 are those of
    # the owner's face seen from this cell (sign flipped) and the upwind masks follow
    # the signs, as in the transport (inconsistent masks would multiply the +-BIG
    # bounds into inf/nan, which the reference never produces)
    qa = rng.uniform(-0.2, 1.2, (g, g, t.kall, t.lall, 6))
    qa_pl = rng.uniform(-0.2, 1.2, (6, t.kall, 2))
    ch = rng.uniform(-0.3, 0.3, (g, g, t.kall, t.lall, 6))
    ch[1:, :, ..., 3] = -ch[:-1, :, ..., 0]
    ch[1:, 1:, ..., 4] = -ch[:-1, :-1, ..., 1]
    ch[:, 1:, ..., 5] = -ch[:, :-1, ..., 2]
    ch_pl = rng.uniform(-0.3, 0.3, (6, t.kall, 2))
    eps = np.finfo(np.float64).eps
    cm = 0.5 - np.copysign(0.5, ch - eps)
    cm_pl = 0.5 - np.copysign(0.5, ch_pl - eps)
    ref, ref_pl = qa.copy(), qa_pl.copy()
    with np.errstate(over="ignore", invalid="ignore"):
        ot.horizontal_limiter(octx, ref, ref_pl, q[0], q[1], d[0], d[1], ch, ch_pl, cm, cm_pl)
    got, got_pl = ops_mod.horizontal_limiter(env.ctx, qa, qa_pl, q[0], q[1], d[0], d[1], ch, ch_pl, cm, cm_pl)
    # the reference limits the faces of cells 0..g-2 (and faces 3..5 of their neighbours)
    assert np.array_equal(got, ref)
    assert np.array_equal(got_pl[1:6], ref_pl[1:6])


def test_operator_argument_errors(env):
    """Wrong tensors are refused before any launch."""
    import torch
    ops = env.ops
    f = env.ctx.empty(1)[0]
    with pytest.raises(ValueError):
        ops.vi_matrix(env.ctx.empty(2), f, 100.0)
    with pytest.raises(TypeError):          # the matrix must be FP64
        ops.vi_solve(f, f, f, f, f, f, f, torch.zeros((3, env.ctx.NF), dtype=torch.float32,
                                                      device=env.ctx.device), 100.0)
    with pytest.raises(RuntimeError, match="NH_ERR_ARG"):
        ops.vi_matrix(f, f, -1.0)
    with pytest.raises(ValueError):
        ops.horizontal_limiter(env.ctx.empty(5), f, f, env.ctx.empty(6), env.ctx.empty(6))
