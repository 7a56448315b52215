"""Region connectivity and the halo gather map (integer, bit-exact).

Pins: the reference's rlevel-1 region link fixture, the reference's gl05 grid
files (whose halo points the reference filled with its own halo exchange),
and a brute-force geometric matcher that finds, for every halo point, the
interior point with the same coordinates.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, mod

topology = mod("topology")


@pytest.fixture(scope="module")
def topo5():
    return topology.Topology(5, 1, 40)


def test_region_links_match_reference_fixture():
    """edge_tab equals rl01-prc000008.toml RGN_LINK_INFO (prep/mnginfo output)."""
    links = json.load(open(os.path.join(GOLDEN, "rl01_region_links.json")))
    tab = topology.region_link_table(1)
    for e, name in enumerate(("SW", "NW", "NE", "SE")):
        ref = np.array(links[name])
        assert np.array_equal(tab[topology.I_RGNID, e, :], ref[:, 0]), name
        assert np.array_equal(tab[topology.I_DIR, e, :], ref[:, 1]), name


@pytest.mark.parametrize("gl", [2, 3, 5])
def test_gather_map_invariants(gl):
    """One source per halo slot, sources are never halo slots, every ring point is a slot."""
    t = topology.Topology(gl, 1, 40)
    dst = t.halo_dst.astype(np.int64)
    src = t.halo_src.astype(np.int64)
    assert dst.size == np.unique(dst).size
    assert np.intersect1d(dst, src).size == 0
    g = t.gall_1d
    ring = 4 * (g - 1)
    # every regional ring point is a halo slot, plus the 5 spokes of each pole
    assert dst.size == t.lall * ring + 2 * 5
    nreg = t.lall * g * g
    reg = dst[dst < nreg]
    i, j = reg % g, (reg // g) % g
    assert np.all((i == 0) | (j == 0) | (i == g - 1) | (j == g - 1))
    pol = dst[dst >= nreg] - nreg
    assert sorted(pol % 6) == sorted([1, 2, 3, 4, 5] * 2)
    # sources: interior regional points or pole centres
    sreg = src[src < nreg]
    si, sj = sreg % g, (sreg // g) % g
    assert np.all((si >= 1) & (si <= g - 2) & (sj >= 1) & (sj <= g - 2))
    assert np.all((src[src >= nreg] - nreg) % 6 == 0)


@pytest.fixture(scope="module")
def json_grid():
    x = np.load(os.path.join(GOLDEN, "gl05_grd_x.npz"))["grd_x"]        # (l, j, i, 3)
    return np.ascontiguousarray(np.transpose(x, (2, 1, 0, 3)))           # (i, j, l, 3)


def test_halo_fill_reproduces_reference_grid_halos(topo5, json_grid):
    """Filling the reference gl05 grid's halos from its interior gives back its halo values bit for bit."""
    t = topo5
    g = t.gall_1d
    x = json_grid
    var = np.full((g, g, 1, t.lall, 3), np.nan)
    var[1:g - 1, 1:g - 1, 0] = x[1:g - 1, 1:g - 1]
    var_pl = np.full((6, 1, 2, 3), np.nan)
    # pole centres: the value the reference's halos hold where they came from a pole
    nreg = t.lall * g * g
    from_pole = t.halo_src >= nreg
    for lp in range(2):
        sel = from_pole & ((t.halo_src - nreg) // 6 == lp)
        d = t.halo_dst[sel]
        vals = x[d % g, (d // g) % g, d // (g * g)]
        assert np.all(vals == vals[0]), "pole-sourced halo points disagree"
        assert abs(abs(vals[0][2]) - 1.0) < 1e-12 and (vals[0][2] > 0) == (lp == 0)
        var_pl[0, 0, lp] = vals[0]
    t.halo_fill_host(var, var_pl)
    got = var[:, :, 0]
    ring = np.ones((g, g), bool)
    ring[1:g - 1, 1:g - 1] = False
    assert np.array_equal(got[ring], x[ring])


def test_halo_map_matches_geometric_brute_force(topo5, json_grid):
    """Every regional-sourced halo point has the coordinates of exactly its source interior point."""
    t = topo5
    g = t.gall_1d
    x = json_grid
    inner = x[1:g - 1, 1:g - 1].reshape(-1, 3)                               # (i, j, l) order
    ii, jj, ll = np.meshgrid(np.arange(1, g - 1), np.arange(1, g - 1), np.arange(t.lall), indexing="ij")
    cols = (ll * g + jj) * g + ii
    cols = cols.reshape(-1)
    nreg = t.lall * g * g
    dst = t.halo_dst.astype(np.int64)
    src = t.halo_src.astype(np.int64)
    reg = (dst < nreg) & (src < nreg)
    d = dst[reg]
    hx = x[d % g, (d // g) % g, d // (g * g)]                                # (n, 3)
    # nearest interior point by brute force (chunked)
    found = np.empty(d.size, np.int64)
    for s in range(0, d.size, 512):
        dist = np.linalg.norm(hx[s:s + 512, None, :] - inner[None, :, :], axis=-1)
        k = np.argmin(dist, axis=1)
        assert np.all(dist[np.arange(k.size), k] < 1e-12)
        found[s:s + 512] = cols[k]
    assert np.array_equal(found, src[reg])
