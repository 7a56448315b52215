import importlib
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PKG = "pynicam-dc_amd"
GOLDEN = os.path.join(ROOT, "tests", "golden")
HOSTEMU_SRC = [os.path.join(ROOT, "tests", "hostemu", f) for f in ("nh_hostemu.cpp", "nh_hostemu_f32.cpp")]
HOSTEMU_HDR = os.path.join(ROOT, "tests", "hostemu", "nh_hostexec.h")
HOSTEMU_LIB = os.path.join(ROOT, "tests", "_build", "libnh_hostemu.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def mod(name):
    return importlib.import_module(f"{PKG}.{name}")


@pytest.fixture(scope="session")
def case_gl04():
    """gl04 (2,562 cells), 10 layers up to 10 km: small enough for the pure-numpy oracle."""
    return mod("setup").build_case(glevel=4, vlayer=10, ztop=10000.0, dz_bottom=200.0, dtl=1200.0, seed=0)


@pytest.fixture(scope="session")
def state_gl04_tracers(case_gl04):
    """gl04 state with qv = 0 and three cosine-blob tracers."""
    return mod("initial").blob_tracers(case_gl04.geo, case_gl04.state, ntracer=3, seed=0)


@pytest.fixture(scope="session")
def case_gl05():
    """gl05 z40, 30 km top, 100 m bottom layer, dt 1200 s: the CPU-runnable benchmark case."""
    return mod("setup").build_case(glevel=5, vlayer=40, dtl=1200.0, seed=0)


def _newer(target, *deps):
    return os.path.exists(target) and all(os.path.getmtime(target) >= os.path.getmtime(d) for d in deps)


@pytest.fixture(scope="session")
def hostemu_lib():
    """Host (g++) build of the kernel sources, for CPU tests of the kernel logic."""
    csrc = os.path.join(ROOT, PKG, "csrc")
    deps = HOSTEMU_SRC + [HOSTEMU_HDR, os.path.join(ROOT, "include", "nicamhip.h")] + [
        os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith(".h")]
    if not _newer(HOSTEMU_LIB, *deps):
        os.makedirs(os.path.dirname(HOSTEMU_LIB), exist_ok=True)
        tmp = HOSTEMU_LIB + ".tmp"
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
                        "-o", tmp, *HOSTEMU_SRC], check=True)
        os.replace(tmp, HOSTEMU_LIB)
    return HOSTEMU_LIB


def interior(a, topo):
    """Interior points on physical levels of a regional (i, j, k, l, ...) array."""
    g = topo.gall_1d
    return a[1:g - 1, 1:g - 1, topo.kmin:topo.kmax + 1]


def pole_centre(a_pl, topo):
    return a_pl[0, topo.kmin:topo.kmax + 1]


def rel_err(a, b, a_pl=None, b_pl=None):
    """max |a - b| / max |a| over regional (and pole) values."""
    sc = np.abs(a).max()
    e = np.abs(a - b).max()
    if a_pl is not None:
        sc = max(sc, np.abs(a_pl).max())
        e = max(e, np.abs(a_pl - b_pl).max())
    return e / max(sc, 1e-300)
