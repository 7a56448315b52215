"""nh_budget_noalloc, the allocation-free budget entry point, against nh_budget.

Two calls in a row on the MI355X through the C-ABI, the second on a smaller
grid (and the other field precision), so that the second call reuses the
library's scratch array the first one filled: gl05 z40 FP64, then gl04 z10
mixed.  Each must agree with nh_budget of the same state to 1e-14 of every
sum (both are compensated; only the order of the final additions differs).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import mod

pytestmark = pytest.mark.gpu

hipdyn = mod("hipdyn")
native = mod("native")


def _budgets(ctx, dyn):
    lib = ctx.lib
    ref = (C.c_double * 6)()
    got = (C.c_double * 6)()
    args = (ctx.handle, C.c_void_p(dyn.prog.data_ptr()), C.c_void_p(dyn.progq.data_ptr()),
            int(dyn.progq.shape[0]), float(ctx.const.LHV))
    native.check(lib, ctx.fn("nh_budget")(*args, ref, ctx.stream()), "nh_budget")
    native.check(lib, lib.nh_budget_noalloc(*args, got, ctx.stream()), "nh_budget_noalloc")
    return np.array(ref[:]), np.array(got[:])


def test_budget_noalloc_matches_budget_two_calls():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    cases = [("gl05 z40 float64", dict(glevel=5, vlayer=40, dtl=1200.0, seed=0), "float64"),
             ("gl04 z10 mixed", dict(glevel=4, vlayer=10, ztop=10000.0, dz_bottom=200.0, dtl=1200.0, seed=0),
              "mixed")]
    for name, kw, prec in cases:
        print(f"  nh_budget_noalloc: {name}", flush=True)
        case = mod("setup").build_case(**kw)
        st = mod("initial").blob_tracers(case.geo, case.state, ntracer=2, seed=1)
        ctx = hipdyn.HipContext(case.geo, case.const, device="cuda:0", precision=prec)
        try:
            dyn = hipdyn.HipDynamics(ctx, st)
            dyn.dynamics_step()
            ref, got = _budgets(ctx, dyn)
        finally:
            ctx.close()
        print(f"    nh_budget {ref}\n    noalloc   {got}", flush=True)
        assert np.all(np.abs(got - ref) <= 1e-14 * np.abs(ref)), (got, ref)
