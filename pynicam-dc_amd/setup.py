"""One-call construction of a benchmark-family case (host, FP64).

Grid (icosahedral, rlevel 1, flat topography), stretched vertical grid,
constants, geometry/operator coefficients and the seeded initial state.
"""
from __future__ import annotations

from dataclasses import dataclass

from . import config, geometry, grid, initial, topology


@dataclass
class Case:
    topo: topology.Topology
    vgrid: grid.VGrid
    const: config.Const
    geo: geometry.Geometry
    state: dict


def build_case(glevel: int, vlayer: int = 40, ztop: float = 30000.0, dz_bottom: float = 100.0,
               dtl: float = 1200.0, seed: int = 0, rlevel: int = 1, with_state: bool = True) -> Case:
    topo = topology.Topology(glevel, rlevel, vlayer)
    hg = grid.generate_icosahedral(topo)
    zh = grid.stretched_interfaces(vlayer, ztop, dz_bottom)
    gz, gzh = grid.vgrid_from_interfaces(zh)
    vg = grid.VGrid(gz, gzh)
    c = config.benchmark_const(glevel, dtl)
    geo = geometry.Geometry(topo, hg, vg, c.RADIUS, c.GRAV)
    state = initial.baroclinic_bumps_state(geo, c, seed=seed) if with_state else None
    return Case(topo, vg, c, geo, state)
