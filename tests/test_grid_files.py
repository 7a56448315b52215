"""The horizontal grid: spring dynamics and gravity centres against the
reference's own gl05 grid files, and the grid-file reader/writer (CPU).

tests/golden/gl05_grd_x.npz holds the points of the reference's case2 grid
(pynicamdc/test/case2/json_gl05rl01pe08/bboundary_GL05RL01.pe0000000[0-7].json,
its halos included), which the reference made with MKGRD_standard ->
MKGRD_spring (beta 1.15) -> MKGRD_gravcenter.  The same chain here must
reproduce them to round-off.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, mod

grid = mod("grid")
topology = mod("topology")
REF_CASE2 = "/root/reference/pynicamdc/test/case2/json_gl05rl01pe08/bboundary_GL05RL01.pe"


@pytest.fixture(scope="module")
def spring_gl05():
    t = topology.Topology(5, 1, 1)
    hs, n = grid.spring_relax(t, grid.generate_icosahedral(t))
    return t, grid.gravity_centre(t, hs), n


def test_spring_and_gravity_centre_reproduce_the_reference_gl05_grid(spring_gl05):
    t, hg, n = spring_gl05
    assert 1000 < n < 10000
    want = np.load(os.path.join(GOLDEN, "gl05_grd_x.npz"))["grd_x"]       # (l, j, i, 3)
    got = hg.x[:, :, 0].transpose(2, 1, 0, 3)
    assert np.abs(got - want).max() < 1e-12
    # the spring step alone moves the points by about a tenth of a grid length
    std = grid.generate_icosahedral(t).x[:, :, 0].transpose(2, 1, 0, 3)
    assert np.abs(std - want).max() > 1e-3


@pytest.mark.parametrize("io_mode", ["npz", "json"])
def test_grid_file_round_trip(tmp_path, io_mode):
    t = topology.Topology(4, 1, 1)
    hg = grid.generate_icosahedral(t)
    prefix = str(tmp_path / "boundary_GL04RL01.pe")
    paths = grid.write_hgrid(prefix, hg, nfiles=8, io_mode=io_mode)
    assert os.path.basename(paths[7]) == f"boundary_GL04RL01.pe00000007.{io_mode}"
    back = grid.read_hgrid(prefix, t, io_mode=io_mode)
    assert np.array_equal(back.x, hg.x)
    assert np.array_equal(back.xt, hg.xt)
    assert np.array_equal(back.xt_pl, hg.xt_pl)
    novert = grid.read_hgrid(prefix, t, io_mode=io_mode, lrvertex=False)
    assert np.array_equal(novert.xt[1:-1, 1:-1], hg.xt[1:-1, 1:-1])
    with pytest.raises(ValueError, match="regions"):
        grid.read_hgrid(prefix, topology.Topology(5, 1, 1), io_mode=io_mode)


@pytest.mark.skipif(not os.path.exists(REF_CASE2 + "00000000.json"), reason="reference grid files not mounted")
def test_read_the_reference_case2_grid_files(spring_gl05):
    t, hg, _ = spring_gl05
    t40 = topology.Topology(5, 1, 40)
    ref = grid.read_hgrid(REF_CASE2, t40, io_mode="json")
    assert np.abs(ref.x - hg.x).max() < 1e-12
    s = slice(1, t.gall_1d - 1)
    assert np.abs(ref.xt[s, s] - hg.xt[s, s]).max() < 1e-12
    geo = mod("geometry").Geometry(t40, ref, mod("setup").build_case(5, with_state=False).vgrid, 6.371e6, 9.80616)
    total = geo.area[s, s].sum() + geo.area_pl[0].sum()
    assert abs(total / (4 * np.pi * 6.371e6 ** 2) - 1.0) < 1e-12
