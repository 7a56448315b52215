"""Hydrostatic basic state (Bsst, nhm/share/mod_bsstate.py).

``ref_type = 'NOBASE'`` (every reference configuration) makes the basic state
zero; ``'INPUT'`` reads a 1-D reference profile (pressure, temperature, qv on
the model's full levels) from a file and forms the 3-D basic state from it
(``_set_basicstate``, mod_bsstate.py:159-200).  The dynamics then works with
perturbations: ``pregd = (pre - pre_bs) GSGAM2`` and ``rhogd = (rho - rho_bs)
GSGAM2`` for the pressure gradient and buoyancy (mod_dynamics.py:1056-1057),
``tem - tem_bs`` and ``rho - rho_bs`` for the hyperdiffusion
(mod_numfilter.py:1068-1076).  With flat topography the basic state depends on
k only; the HIP backend keeps these profiles in FP64 in both forms, so the
mixed form's FP32 fields carry perturbations where the reference state is
subtracted.

File format (``_input_ref``, mod_bsstate.py:96-108): a big-endian Fortran
sequential unformatted file with three records of ``kall`` float64 values
(pre_ref, tem_ref, qv_ref), each record framed by its 4-byte byte count.
"""
from __future__ import annotations

import numpy as np


def write_ref(fname: str, pre_ref, tem_ref, qv_ref) -> None:
    """Write the three profiles in the reference's ``ref.dat`` format."""
    with open(fname, "wb") as f:
        for rec in (pre_ref, tem_ref, qv_ref):
            data = np.asarray(rec, dtype=">f8").tobytes()
            marker = np.array([len(data)], dtype=">i4").tobytes()
            f.write(marker + data + marker)


def read_ref(fname: str, kall: int):
    """(pre_ref, tem_ref, qv_ref), each (kall,) float64, from a ``ref.dat`` file."""
    recs = []
    with open(fname, "rb") as f:
        for _ in range(3):
            head = f.read(4)
            if len(head) != 4:
                raise ValueError(f"{fname}: fewer than three records")
            n = int(np.frombuffer(head, dtype=">i4")[0])
            data = np.frombuffer(f.read(n), dtype=">f8")
            tail = f.read(4)
            if len(tail) != 4 or int(np.frombuffer(tail, dtype=">i4")[0]) != n:
                raise ValueError(f"{fname}: record markers do not match")
            if data.size != kall:
                raise ValueError(f"{fname}: a record holds {data.size} values, expected kall = {kall}")
            recs.append(data.astype(np.float64))
    return recs[0], recs[1], recs[2]


class BasicState:
    """The 3-D basic state of ``_set_basicstate`` on a flat-topography grid.

    With flat topography the height of a level is the same everywhere
    (GRD_vz = gz), so the z -> xi re-sampling (``_vintrpl_z2xi``) returns each
    physical level's own value (the three-point Lagrange weights are exactly
    1, 0, 0) and copies the nearest physical level into the two ghost levels;
    the density follows from the moist gas law, and ``BNDCND_thermo`` (TEM top
    and bottom, hydrostatic ghost pressures, mod_bndcnd.py / kernels/bndcnd.py:54-96)
    resets the ghost levels.  Profiles: ``pre``, ``tem``, ``rho`` (kall,);
    full arrays for the reference layout via ``full()``.
    """

    def __init__(self, geo, const, pre_ref, tem_ref, qv_ref):
        t = geo.topo
        kmin, kmax, kall = t.kmin, t.kmax, t.kall
        pre = np.asarray(pre_ref, dtype=np.float64).copy()
        tem = np.asarray(tem_ref, dtype=np.float64).copy()
        qv = np.asarray(qv_ref, dtype=np.float64).copy()
        if pre.shape != (kall,) or tem.shape != (kall,) or qv.shape != (kall,):
            raise ValueError("reference profiles must have kall entries")
        for a in (pre, tem, qv):
            a[kmin - 1] = a[kmin]
            a[kmax + 1] = a[kmax]
        c = const
        rho = pre / tem / ((1.0 - qv) * c.Rdry + qv * c.Rvap)
        phi = np.asarray(geo.vmtr["PHI"][1, 1, :, 0], dtype=np.float64)
        tem_top, tem_btm = tem[kmax], tem[kmin]
        pre_top = pre[kmax - 1] - rho[kmax] * (phi[kmax + 1] - phi[kmax - 1])
        pre_btm = pre[kmin + 1] - rho[kmin] * (phi[kmin - 1] - phi[kmin + 1])
        tem[kmax + 1], tem[kmin - 1] = tem_top, tem_btm
        pre[kmax + 1], pre[kmin - 1] = pre_top, pre_btm
        rho[kmax + 1] = pre_top / (c.Rdry * tem_top)
        rho[kmin - 1] = pre_btm / (c.Rdry * tem_btm)
        self.pre, self.tem, self.rho, self.qv = pre, tem, rho, qv
        self.topo = t

    def full(self, name: str):
        """(regional (g, g, kall, l), pole (6, kall, 2)) arrays of profile ``name``."""
        t = self.topo
        p = getattr(self, name)
        reg = np.broadcast_to(p[None, None, :, None], (t.gall_1d, t.gall_1d, t.kall, t.lall)).copy()
        pl = np.broadcast_to(p[None, :, None], (6, t.kall, 2)).copy()
        return reg, pl


def bsstate_setup(geo, const, ref_type: str = "NOBASE", ref_fname: str | None = None):
    """Bsst.bsstate_setup (mod_bsstate.py:20-90): None for 'NOBASE', a BasicState
    from ``ref_fname`` for 'INPUT'; other types are refused as the reference does."""
    if ref_type == "NOBASE":
        return None
    if ref_type == "INPUT":
        if not ref_fname:
            raise ValueError("ref_type 'INPUT' needs ref_fname")
        pre, tem, qv = read_ref(ref_fname, geo.topo.kall)
        return BasicState(geo, const, pre, tem, qv)
    raise NotImplementedError(f"ref_type {ref_type!r} (generate) is not implemented, as in the reference")


def rest_profile(geo, const):
    """The benchmark's resting atmosphere on the full levels, in the model's
    discrete hydrostatic balance (the same integration as the initial state):
    (pre_ref, tem_ref, qv_ref) for a ``ref.dat`` file."""
    from .initial import reference_profile
    t, vg, c = geo.topo, geo.vgrid, const
    kmin, kall = t.kmin, t.kall
    tem, pre = reference_profile(vg.gz, grav=c.GRAV, rd=c.Rdry)
    p = pre.copy()
    for k in range(kmin + 1, kall):
        dz = vg.dgzh[k]
        a_, b_ = vg.afact[k], vg.bfact[k]
        rho_km1 = p[k - 1] / (c.Rdry * tem[k - 1])
        p[k] = (p[k - 1] - c.GRAV * dz * b_ * rho_km1) / (1.0 + c.GRAV * dz * a_ / (c.Rdry * tem[k]))
    return p, np.asarray(tem, dtype=np.float64), np.zeros(kall)
