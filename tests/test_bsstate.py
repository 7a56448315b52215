"""The hydrostatic basic state (ref_type 'INPUT', nhm/share/mod_bsstate.py).

The reference file format round trip; the 3-D basic state of _set_basicstate
on a flat grid (physical levels equal to the file's profile, ghost levels from
BNDCND_thermo, the moist gas law); a step with the basic state through the host
build of the kernels and on the MI355X against the oracle (the FP64 form to the
same tolerances as without one); and the mixed form against the FP64 form.
The adjoint refuses a context with a basic state.
"""
import os

import numpy as np
import pytest

from conftest import mod

bsstate = mod("bsstate")
hipdyn = mod("hipdyn")


def test_ref_file_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(0)
    p, t, q = (rng.uniform(size=12) for _ in range(3))
    f = str(tmp_path / "ref.dat")
    bsstate.write_ref(f, p, t, q)
    raw = open(f, "rb").read()
    assert len(raw) == 3 * (8 + 12 * 8)
    assert np.frombuffer(raw[:4], dtype=">i4")[0] == 96 and np.frombuffer(raw[4:12], dtype=">f8")[0] == p[0]
    back = bsstate.read_ref(f, 12)
    for a, b in zip(back, (p, t, q)):
        assert np.array_equal(a, b)
    with pytest.raises(ValueError, match="kall"):
        bsstate.read_ref(f, 13)
    with open(f, "r+b") as fh:
        fh.truncate(len(raw) - 10)
    with pytest.raises(ValueError):
        bsstate.read_ref(f, 12)


def test_basic_state_profiles(case_gl04):
    geo, c, t = case_gl04.geo, case_gl04.const, case_gl04.topo
    pre, tem, qv = bsstate.rest_profile(geo, c)
    bs = bsstate.BasicState(geo, c, pre, tem, qv)
    ks = slice(t.kmin, t.kmax + 1)
    assert np.array_equal(bs.pre[ks], pre[ks]) and np.array_equal(bs.tem[ks], tem[ks])
    assert np.array_equal(bs.rho[ks], pre[ks] / tem[ks] / c.Rdry)
    phi = geo.vmtr["PHI"][1, 1, :, 0]
    k = t.kmin
    assert bs.tem[k - 1] == bs.tem[k] and bs.tem[t.kmax + 1] == bs.tem[t.kmax]
    assert bs.pre[k - 1] == bs.pre[k + 1] - bs.rho[k] * (phi[k - 1] - phi[k + 1])
    # the rest profile is in the model's discrete hydrostatic balance
    vg = geo.vgrid
    for kk in range(t.kmin + 1, t.kmax + 1):
        lhs = (pre[kk] - pre[kk - 1]) / vg.dgzh[kk]
        rhs = -c.GRAV * (vg.afact[kk] * pre[kk] / (c.Rdry * tem[kk]) + vg.bfact[kk] * pre[kk - 1] / (c.Rdry * tem[kk - 1]))
        assert abs(lhs - rhs) <= 1e-12 * abs(rhs)
    with pytest.raises(NotImplementedError):
        bsstate.bsstate_setup(geo, c, "GEN")


@pytest.fixture(scope="module")
def case_input_gl04():
    return mod("setup").build_case(glevel=4, vlayer=10, ztop=10000.0, dz_bottom=200.0, dtl=1200.0, seed=0,
                                   ref_type="INPUT")


def _compare(ref, got, t, tol):
    g = t.gall_1d
    s, ks = slice(1, g - 1), slice(t.kmin, t.kmax + 1)
    mom = np.abs(ref.PROG[s, s, ks, :, 1:4]).max()
    for v in range(6):
        a, b = ref.PROG[s, s, ks, :, v], got["PROG"][s, s, ks, :, v]
        sc = max(np.abs(a).max(), mom if v == 4 else 0.0)
        assert np.abs(a - b).max() <= tol * sc, (v, np.abs(a - b).max() / sc)


def test_host_build_step_with_basic_state(case_input_gl04, hostemu_lib, tmp_path):
    from oracle import dynamics as od
    c = case_input_gl04
    # the same profile through the reference's file format
    f = str(tmp_path / "ref.dat")
    bsstate.write_ref(f, *bsstate.rest_profile(c.geo, c.const))
    bs = bsstate.bsstate_setup(c.geo, c.const, "INPUT", f)
    assert np.array_equal(bs.pre, c.geo.bsstate.pre) and np.array_equal(bs.rho, c.geo.bsstate.rho)
    ref = od.OracleDynamics(od.OracleContext(c.geo, c.const, None), c.state)
    ctx = hipdyn.HipContext(c.geo, c.const, device="cpu", lib_path=hostemu_lib)
    try:
        dyn = hipdyn.HipDynamics(ctx, c.state)
        for _ in range(2):
            ref.step()
            dyn.dynamics_step()
        _compare(ref, dyn.state(), c.topo, 1e-10)
        prog_b = dyn.prog.clone().zero_()
        with pytest.raises(RuntimeError, match="NOBASE"):
            dyn.dynamics_step_adj(dyn.prog, dyn.progq, prog_b)
    finally:
        ctx.close()
    # the basic state changes the step (the rest profile is balanced only to the
    # discretisation of the bumped, rotating state)
    base = mod("setup").build_case(glevel=4, vlayer=10, ztop=10000.0, dz_bottom=200.0, dtl=1200.0, seed=0)
    nob = od.OracleDynamics(od.OracleContext(base.geo, base.const, None), base.state)
    nob.step()
    nob.step()
    assert np.abs(nob.PROG[..., 4] - ref.PROG[..., 4]).max() > 0.0


@pytest.mark.gpu
def test_gpu_step_with_basic_state_fp64_and_mixed():
    """gl05 z40 with the basic state: the FP64 form against the oracle after one
    step (1e-10), the mixed form against the FP64 form (1e-4 of each field's
    largest value; rhogw on the momentum scale)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from oracle import dynamics as od
    c = mod("setup").build_case(glevel=5, vlayer=40, dtl=1200.0, seed=0, ref_type="INPUT")
    print("  gl05 z40 with the INPUT basic state", flush=True)
    ref = od.OracleDynamics(od.OracleContext(c.geo, c.const, None), c.state)
    ref.step()
    out = {}
    for prec in ("float64", "mixed"):
        ctx = hipdyn.HipContext(c.geo, c.const, device="cuda:0", precision=prec)
        try:
            dyn = hipdyn.HipDynamics(ctx, c.state)
            dyn.dynamics_step()
            out[prec] = dyn.state()
        finally:
            ctx.close()
    _compare(ref, out["float64"], c.topo, 1e-10)

    class R:
        PROG = out["float64"]["PROG"]
    _compare(R, {"PROG": out["mixed"]["PROG"].astype(np.float64)}, c.topo, 1e-4)
