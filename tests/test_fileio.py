"""Restart files and history output in the reference's formats (host, CPU).

The restart layout (one (ij, k, l) array per variable, ij = j * gall_1d + i)
is pinned against the reference's own packing (Prgv._pack5d, golden vector
in tests/golden/kernels.npz); the round trip write -> read of the prognostic
state is bit-exact, the poles being refilled by the halo exchange as the
reference's restart_input does.  The zarr store follows the v2 layout that
xarray writes for IO_setup (mod_io.py:132-409), uncompressed.
"""
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, mod

fileio = mod("fileio")
sys.path.insert(0, GOLDEN)
import kernel_cases as kc  # noqa: E402


def test_restart_layout_is_the_references():
    vec = np.load(os.path.join(GOLDEN, "kernels.npz"))
    P = kc.make_inputs(seed=0)["PROG"]
    assert np.array_equal(fileio.pack5d(P[..., 1]), vec["restart_pack/prog_rhogvx"])
    assert np.array_equal(fileio.pack5d(P[..., 5]), vec["restart_pack/prog_rhoge"])
    assert np.array_equal(fileio.unpack5d(fileio.pack5d(P[..., 5]), kc.G), P[..., 5])


def test_restart_round_trip_bit_exact(tmp_path, case_gl04, state_gl04_tracers):
    t = case_gl04.topo
    st = state_gl04_tracers
    names = ("qv", "p1", "p2", "p3")
    base = str(tmp_path / "restart_all_GL04RL01z10.pe")
    path = fileio.restart_output(base, st, tracer_names=names)
    assert os.path.basename(path) == "restart_all_GL04RL01z10.pe00000000.npz"
    with np.load(path) as z:
        want = list(fileio.PRG_NAMES) + ["rhog" + n for n in names]
        assert sorted(z.files) == sorted(want + [w + "_pl" for w in want])
        assert z["rhog"].shape == (t.gall_1d ** 2, t.kall, t.lall)
    back = fileio.restart_input(base, t, tracer_names=names)
    for k in ("PROG", "PROG_pl", "PROGq", "PROGq_pl"):
        assert np.array_equal(back[k], st[k]), k


def test_restart_input_errors(tmp_path, case_gl04):
    t = case_gl04.topo
    st = case_gl04.state
    base = str(tmp_path / "r.pe")
    fileio.restart_output(base, st, tracer_names=("qv",))
    with pytest.raises(KeyError, match="rhogqc"):
        fileio.restart_input(base, t, tracer_names=("qv", "qc"))
    back = fileio.restart_input(base, t, tracer_names=("qv", "qc"), allow_missingq=True)
    assert np.all(back["PROGq"][..., 1] == 0.0)
    other = mod("topology").Topology(3, 1, t.vlayer)
    with pytest.raises(ValueError, match="shape"):
        fileio.restart_input(base, other)


def test_zarr_history_store(tmp_path, case_gl04, state_gl04_tracers):
    t = case_gl04.topo
    st = state_gl04_tracers
    path = str(tmp_path / "testout.zarr")
    h = fileio.ZarrHistory(path, t, case_gl04.geo.GRD_x, nt=2, dtl=1200.0, interval=3, tracer_names=("qv",))
    h.write(st)
    st2 = {k: v * 2.0 for k, v in st.items()}
    h.write(st2)
    with pytest.raises(IndexError):
        h.write(st)
    meta = json.load(open(os.path.join(path, ".zmetadata")))["metadata"]
    assert meta["RHOG/.zattrs"]["_ARRAY_DIMENSIONS"] == ["time", "i", "j", "k", "r"]
    assert meta["RHOG/.zarray"]["shape"] == [2, t.gall_1d, t.gall_1d, t.kall, t.lall]
    assert meta["RHOG/.zarray"]["chunks"] == [1, t.gall_1d, t.gall_1d, t.kall, t.lall]
    assert meta[".zattrs"]["dtl"] == 1200.0
    a = fileio.read_zarr_array(path, "RHOGE")
    assert np.array_equal(a[0], st["PROG"][..., 5]) and np.array_equal(a[1], 2.0 * st["PROG"][..., 5])
    assert np.array_equal(fileio.read_zarr_array(path, "qv")[1], 2.0 * st["PROGq"][..., 0])
    assert np.array_equal(fileio.read_zarr_array(path, "time"), [3600.0, 7200.0])
    assert np.array_equal(fileio.read_zarr_array(path, "GRD_x"), case_gl04.geo.GRD_x[:, :, 0])


def test_restart_without_pole_centres(tmp_path, case_gl04):
    """A file as the reference writes it (regional arrays only): the regions come back
    bit-exact except the halo points the pole centre fills; the centre is its ring's mean."""
    t = case_gl04.topo
    st = case_gl04.state
    base = str(tmp_path / "ref.pe")
    data = {n: fileio.pack5d(st["PROG"][..., v]) for v, n in enumerate(fileio.PRG_NAMES)}
    data["rhogqv"] = fileio.pack5d(st["PROGq"][..., 0])
    np.savez(fileio.restart_path(base), **data)
    back = fileio.restart_input(base, t)
    assert back["pole_from_ring"]
    g = t.gall_1d
    s = slice(1, g - 1)
    assert np.array_equal(back["PROG"][s, s], st["PROG"][s, s])
    assert np.array_equal(back["PROG_pl"][1:6], st["PROG_pl"][1:6])
    assert np.allclose(back["PROG_pl"][0], st["PROG_pl"][1:6].mean(axis=0))
