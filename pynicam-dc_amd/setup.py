"""One-call construction of a benchmark-family case (host, FP64).

Grid (icosahedral, rlevel 1, flat topography; standard, spring-relaxed or read
from the reference's grid files), stretched vertical grid, constants,
geometry/operator coefficients and the seeded initial state.
"""
from __future__ import annotations

from dataclasses import dataclass

from . import bsstate, config, geometry, grid, initial, topology


@dataclass
class Case:
    topo: topology.Topology
    vgrid: grid.VGrid
    const: config.Const
    geo: geometry.Geometry
    state: dict


def make_hgrid(topo: topology.Topology, kind="standard"):
    """The horizontal grid: "standard" (recursive icosahedron subdivision), "spring"
    (then spring dynamics with beta 1.15 and the gravity-centre move, the reference's
    MKGRD_standard -> MKGRD_spring -> MKGRD_gravcenter), a grid-file prefix
    ("json:<prefix>" or "npz:<prefix>", Grd.GRD_input_hgrid) or an HGrid."""
    if isinstance(kind, grid.HGrid):
        return kind
    if kind == "standard":
        return grid.generate_icosahedral(topo)
    if kind == "spring":
        hs, _ = grid.spring_relax(topo, grid.generate_icosahedral(topo))
        return grid.gravity_centre(topo, hs)
    if isinstance(kind, str) and ":" in kind:
        mode, prefix = kind.split(":", 1)
        return grid.read_hgrid(prefix, topo, io_mode=mode)
    raise ValueError(f"unknown horizontal grid {kind!r}")


def build_case(glevel: int, vlayer: int = 40, ztop: float = 30000.0, dz_bottom: float = 100.0,
               dtl: float = 1200.0, seed: int = 0, rlevel: int = 1, with_state: bool = True,
               hgrid="standard", ref_type: str = "NOBASE", ref_fname: str | None = None) -> Case:
    """``ref_type`` / ``ref_fname``: the basic state ([bsstateparam], mod_bsstate.py:20-90);
    with 'INPUT' and no file the benchmark's resting atmosphere (``bsstate.rest_profile``)."""
    topo = topology.Topology(glevel, rlevel, vlayer)
    hg = make_hgrid(topo, hgrid)
    zh = grid.stretched_interfaces(vlayer, ztop, dz_bottom)
    gz, gzh = grid.vgrid_from_interfaces(zh)
    vg = grid.VGrid(gz, gzh)
    c = config.benchmark_const(glevel, dtl)
    geo = geometry.Geometry(topo, hg, vg, c.RADIUS, c.GRAV)
    if ref_type == "INPUT" and not ref_fname:
        geo.bsstate = bsstate.BasicState(geo, c, *bsstate.rest_profile(geo, c))
    else:
        geo.bsstate = bsstate.bsstate_setup(geo, c, ref_type, ref_fname)
    state = initial.baroclinic_bumps_state(geo, c, seed=seed) if with_state else None
    return Case(topo, vg, c, geo, state)
