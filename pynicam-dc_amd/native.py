"""ctypes binding of the nicamhip C-ABI (include/nicamhip.h).

The library is built in-tree next to this file (``libnicamhip.so``, gfx950).
Loading fails loudly when it is missing: there is no CPU fallback.  The ABI
style follows the reference's native precedent (share/mod_ncclffi.py:52-79):
``CDLL`` + argtypes/restype, int return codes turned into RuntimeError.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_NAME = "libnicamhip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

NH_ERR = {-1: "NH_ERR_ARG", -2: "NH_ERR_STATE", -3: "NH_ERR_ALLOC", -4: "NH_ERR_DEVICE"}


class NhDims(C.Structure):
    _fields_ = [("gall_1d", C.c_int32), ("kall", C.c_int32), ("kmin", C.c_int32),
                ("kmax", C.c_int32), ("lall", C.c_int32)]


class NhConsts(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "radius", "ohm", "grav", "rdry", "cvdry", "rvap", "cvw_qv", "alpha",
        "gamma_h", "hdiff_fact_rho", "alpha_d")] + [("thuburn_lim", C.c_int32)]


_GEO_FIELDS = [
    "coef_div", "coef_grad", "coef_lap", "coef_intp", "coef_diff", "grd_x",
    "coef_div_pl", "coef_grad_pl", "coef_lap_pl", "coef_intp_pl", "coef_diff_pl", "grd_x_pl",
]
_VERT_FIELDS = ["afact", "bfact", "cfact", "dfact", "rdgz", "rdgzh", "c2w0", "c2w1", "w2c0", "w2c1", "phi"]


_TRC_FIELDS = ["tri_w", "arc_hn", "rarea", "arc_xr", "tri_w_pl", "arc_hn_pl", "rarea_pl", "arc_xr_pl",
               "area", "dgz", "lat", "lon", "lat_pl", "lon_pl", "gz", "gzh", "pre_bs", "tem_bs", "rho_bs"]


class NhGeometry(C.Structure):
    _fields_ = ([(n, C.POINTER(C.c_double)) for n in _GEO_FIELDS]
                + [("sgp", C.POINTER(C.c_int32))]
                + [(n, C.POINTER(C.c_double)) for n in _VERT_FIELDS]
                + [(n, C.POINTER(C.c_double)) for n in _TRC_FIELDS])


EXPORTS = {
    "nh_last_error": (C.c_char_p, []),
    "nh_field_size": (C.c_int64, [C.POINTER(NhDims)]),
    "nh_create": (C.c_int, [C.POINTER(NhDims), C.POINTER(NhConsts), C.c_int, C.POINTER(C.c_void_p)]),
    "nh_destroy": (C.c_int, [C.c_void_p]),
    "nh_upload_geometry": (C.c_int, [C.c_void_p, C.POINTER(NhGeometry)]),
    "nh_upload_halo_map": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int64]),
    "nh_halo_fill": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int32, C.c_void_p]),
    "nh_laplacian": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "nh_divergence": (C.c_int, [C.c_void_p] + [C.c_void_p] * 4 + [C.c_void_p]),
    "nh_gradient": (C.c_int, [C.c_void_p] + [C.c_void_p] * 4 + [C.c_void_p]),
    "nh_dyn_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_int32,
                              C.c_void_p]),
    "nh_tracer_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_void_p]),
    "nh_tracer_step_adj": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_void_p,
                                     C.c_void_p]),
    "nh_trcadv_winds": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_int32, C.POINTER(C.c_double),
                                  C.c_void_p]),
    "nh_budget": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double,
                            C.POINTER(C.c_double), C.c_void_p]),
    "nh_budget_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double,
                                    C.POINTER(C.c_double), C.c_void_p]),
    "nh_create_ex": (C.c_int, [C.POINTER(NhDims), C.POINTER(NhConsts), C.c_int, C.c_int32,
                               C.POINTER(C.c_void_p)]),
    "nh_precision": (C.c_int32, [C.c_void_p]),
    "nh_dyn_step_adj": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_int32,
                                  C.c_void_p, C.c_void_p]),
    "nh_dyn_step_record": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_int32,
                                     C.c_void_p, C.c_void_p]),
    "nh_dyn_step_adj_rec": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_int32,
                                      C.c_void_p, C.c_void_p]),
    "nh_tape_bytes": (C.c_int64, [C.c_void_p, C.c_int32]),
    "nh_dyn_step_tape": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_int32,
                                   C.c_void_p, C.c_void_p]),
    "nh_dyn_step_adj_tape": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_int32,
                                       C.c_void_p, C.c_void_p]),
    "nh_sfc_pressure": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "nh_sfc_pressure_adj": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "nh_forcing_hs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int32, C.c_void_p]),
    "nh_vi_matrix": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double] + [C.c_void_p] * 4),
    "nh_vi_solve": (C.c_int, [C.c_void_p] + [C.c_void_p] * 10 + [C.c_double, C.c_void_p, C.c_void_p]),
    "nh_prepost": (C.c_int, [C.c_void_p] + [C.c_void_p] * 5),
    "nh_hdiffusion": (C.c_int, [C.c_void_p] + [C.c_void_p] * 4),
    "nh_divdamp3d": (C.c_int, [C.c_void_p] + [C.c_void_p] * 8),
    "nh_limiter_v": (C.c_int, [C.c_void_p] + [C.c_void_p] * 5),
    "nh_limiter_h": (C.c_int, [C.c_void_p] + [C.c_void_p] * 6),
    "nh_prof_enable": (C.c_int, [C.c_void_p, C.c_int32]),
    "nh_prof_only": (C.c_int, [C.c_void_p, C.c_char_p]),
    "nh_prof_report": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, C.POINTER(C.c_double),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
}
# the _f32 forms of the field entry points take the same arguments (float fields)
for _name in ("nh_halo_fill", "nh_laplacian", "nh_divergence", "nh_gradient", "nh_dyn_step", "nh_tracer_step",
              "nh_trcadv_winds", "nh_budget", "nh_dyn_step_adj", "nh_sfc_pressure", "nh_sfc_pressure_adj",
              "nh_dyn_step_record", "nh_dyn_step_adj_rec", "nh_dyn_step_tape", "nh_dyn_step_adj_tape",
              "nh_tracer_step_adj", "nh_forcing_hs", "nh_vi_matrix", "nh_vi_solve", "nh_prepost", "nh_hdiffusion",
              "nh_divdamp3d", "nh_limiter_v", "nh_limiter_h"):
    EXPORTS[_name + "_f32"] = EXPORTS[_name]

PREC_FP64, PREC_MIXED = 0, 1

_cache: dict = {}


def load(path: str | None = None):
    """Load the C-ABI library (default: the in-tree gfx950 build) and declare its signatures."""
    path = os.path.abspath(path or LIB_PATH)
    if path in _cache:
        return _cache[path]
    if not os.path.exists(path):
        raise RuntimeError(
            f"nicamhip library not found at {path}; build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    missing = []
    for name, (res, args) in EXPORTS.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            missing.append(name)
            continue
        fn.restype = res
        fn.argtypes = args
    if "nh_create_ex" in missing or "nh_last_error" in missing:
        raise RuntimeError(f"{path} is not a nicamhip library (no nh_create_ex / nh_last_error)")
    for name in missing:
        # a library built from older sources: the entry point fails when called
        setattr(lib, name, _missing_entry(path, name))
    _cache[path] = lib
    return lib


def _missing_entry(path: str, name: str):
    def call(*_args):
        raise RuntimeError(f"{path} has no entry point {name}; rebuild it "
                           "(python -c 'import __graft_entry__ as g; g.build()')")
    return call


def check(lib, rc: int, what: str = ""):
    if rc != 0:
        msg = lib.nh_last_error()
        msg = msg.decode() if msg else ""
        raise RuntimeError(f"{what or 'nicamhip'} failed ({NH_ERR.get(rc, rc)}): {msg}")
