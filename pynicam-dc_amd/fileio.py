"""Restart files and history output in the reference's formats (host side).

Restart (Prgv.restart_output / restart_input, nhm/share/mod_prgvar.py:150-300,
``output_io_mode = "npz"``): one numpy archive per process,
``<basename><rank, 8 digits>.npz``, one array per variable keyed by name --
the prognostics ``rhog rhogvx rhogvy rhogvz rhogw rhoge`` and the tracer
masses ``rhog<tracer>`` (``rhogqv`` ...), each of shape (gall_1d**2, kall,
lall) with the point index ij = j * gall_1d + i (``_pack5d``, :150-154).
Reading fills the halos and the poles' ring points by the halo exchange
(``COMM_var``, :244).  The reference's file holds no pole cell; the pole
centres are written here as well, under ``<name>_pl`` (shape (6, kall, 2)),
which a reader that looks its variables up by name ignores, so that the
prognostic round trip is bit-exact.  A file without them (one the reference
wrote) gets each pole centre as the mean of its five ring points.  With
``diagnostics`` the diagnostic variables are added under the names ``pre tem
vx vy vz w`` and the bare tracer names (:173-179).

History (IO.IO_setup / IO_PRGstep, share/mod_io.py:132-409): a zarr (v2)
directory store with one array per prognostic, dims (time, i, j, k, r),
chunked one time slot by one process's regions, a ``time`` coordinate in
seconds of model time, ``GRD_x`` (i, j, r, xyz), and the consolidated
metadata ``.zmetadata``.  The reference compresses chunks with Blosc; the
chunks here are written uncompressed (``"compressor": null``), which every
zarr v2 reader accepts, and the numcodecs library is not needed.
"""
from __future__ import annotations

import json
import os

import numpy as np

PRG_NAMES = ("rhog", "rhogvx", "rhogvy", "rhogvz", "rhogw", "rhoge")
DIAG_NAMES = ("pre", "tem", "vx", "vy", "vz", "w")
OUT_NAMES = ("RHOG", "RHOGVX", "RHOGVY", "RHOGVZ", "RHOGW", "RHOGE")


def pack5d(a: np.ndarray) -> np.ndarray:
    """(i, j, k, l) -> (ij, k, l) with ij = j * g + i (mod_prgvar.py:150-154)."""
    g = a.shape[0]
    return a.transpose(1, 0, 2, 3).reshape(g * g, a.shape[2], a.shape[3])


def unpack5d(flat: np.ndarray, g: int) -> np.ndarray:
    """(ij, k, l) -> (i, j, k, l) (mod_prgvar.py:221-229)."""
    return np.asarray(flat).reshape(g, g, *flat.shape[1:]).transpose(1, 0, 2, 3)


def restart_path(basename: str, rank: int = 0) -> str:
    return basename + str(rank).zfill(8) + ".npz"


def restart_output(basename: str, state: dict, tracer_names=("qv",), dtype=np.float64, rank: int = 0,
                   diagnostics: dict | None = None) -> str:
    """Write ``state`` (PROG (g,g,kall,l,6), PROGq (g,g,kall,l,nq)) as the reference's npz
    restart.  ``diagnostics``: optional DIAG (g,g,kall,l,6) and q (g,g,kall,l,nq) arrays."""
    P, Q = state["PROG"], state["PROGq"]
    if Q.shape[-1] != len(tracer_names):
        raise ValueError(f"{Q.shape[-1]} tracers but {len(tracer_names)} tracer names")
    data = {n: pack5d(P[..., v]).astype(dtype) for v, n in enumerate(PRG_NAMES)}
    for v, n in enumerate(tracer_names):
        data["rhog" + n] = pack5d(Q[..., v]).astype(dtype)
    if "PROG_pl" in state:
        for v, n in enumerate(PRG_NAMES):
            data[n + "_pl"] = np.asarray(state["PROG_pl"][..., v]).astype(dtype)
        for v, n in enumerate(tracer_names):
            data["rhog" + n + "_pl"] = np.asarray(state["PROGq_pl"][..., v]).astype(dtype)
    if diagnostics is not None:
        for v, n in enumerate(DIAG_NAMES):
            data[n] = pack5d(diagnostics["DIAG"][..., v]).astype(dtype)
        for v, n in enumerate(tracer_names):
            data[n] = pack5d(diagnostics["q"][..., v]).astype(dtype)
    path = restart_path(basename, rank)
    np.savez(path, **data)
    return path


def restart_input(basename: str, topo, tracer_names=("qv",), rank: int = 0, allow_missingq: bool = False,
                  dtype=np.float64) -> dict:
    """Read a prognostic npz restart into a state dict (PROG, PROG_pl, PROGq, PROGq_pl):
    the arrays are unpacked, then the halos and the poles are filled by the halo exchange
    (mod_prgvar.py:241-263).  A missing tracer is an error unless ``allow_missingq``
    (then zero, as the reference's ``allow_missingq``)."""
    g, kall, lall = topo.gall_1d, topo.kall, topo.lall
    with np.load(restart_path(basename, rank), allow_pickle=False) as z:
        missing = [n for n in PRG_NAMES if n not in z.files]
        if missing:
            raise KeyError(f"restart file lacks the prognostic variables {missing}")
        shape = (g * g, kall, lall)
        for n in PRG_NAMES:
            if z[n].shape != shape:
                raise ValueError(f"{n}: shape {z[n].shape}, expected {shape} for this grid")
        P = np.zeros((g, g, kall, lall, 6), dtype=dtype)
        for v, n in enumerate(PRG_NAMES):
            P[..., v] = unpack5d(z[n], g)
        Q = np.zeros((g, g, kall, lall, len(tracer_names)), dtype=dtype)
        for v, n in enumerate(tracer_names):
            key = "rhog" + n
            if key in z.files:
                Q[..., v] = unpack5d(z[key], g)
            elif not allow_missingq:
                raise KeyError(f"prognostic tracer {key!r} not found in the restart file")
        P_pl = np.zeros((6, kall, 2, 6), dtype=dtype)
        Q_pl = np.zeros((6, kall, 2, len(tracer_names)), dtype=dtype)
        with_pole = all(n + "_pl" in z.files for n in PRG_NAMES)
        if with_pole:
            for v, n in enumerate(PRG_NAMES):
                P_pl[..., v] = z[n + "_pl"]
            for v, n in enumerate(tracer_names):
                if "rhog" + n + "_pl" in z.files:
                    Q_pl[..., v] = z["rhog" + n + "_pl"]
    # the ring points of the poles come from the regions; a file without the pole
    # centres gets each centre as the mean of its ring, before the regions' pole
    # halo points are filled from the centre
    topo.halo_fill_host(P, P_pl)
    topo.halo_fill_host(Q, Q_pl)
    if not with_pole:
        P_pl[0] = P_pl[1:6].mean(axis=0)
        Q_pl[0] = Q_pl[1:6].mean(axis=0)
        topo.halo_fill_host(P, P_pl)
        topo.halo_fill_host(Q, Q_pl)
    return {"PROG": P, "PROG_pl": P_pl, "PROGq": Q, "PROGq_pl": Q_pl, "pole_from_ring": not with_pole}


# ----------------------------------------------------------------------------
# zarr v2 history store
# ----------------------------------------------------------------------------
def _zdtype(dt) -> str:
    return np.dtype(dt).newbyteorder("<").str


def _zarray(shape, chunks, dtype, fill=None) -> dict:
    return {"zarr_format": 2, "shape": list(shape), "chunks": list(chunks), "dtype": _zdtype(dtype),
            "compressor": None, "fill_value": fill, "order": "C", "filters": None}


class ZarrHistory:
    """The (time, i, j, k, r) history store of IO_setup / IO_PRGstep (mod_io.py:132-409).

    ``nt`` output slots of ``interval`` large steps of ``dtl`` seconds (plus the initial
    state first when ``step0``); ``write(state)`` fills the next slot with the
    prognostics (and the tracers when ``tracer_names`` is given)."""

    def __init__(self, path: str, topo, grd_x: np.ndarray, nt: int, dtl: float, interval: int,
                 dtype=np.float64, tracer_names=(), step0: bool = False, time_start: float = 0.0):
        self.path = path
        self.g, self.kall, self.lall = topo.gall_1d, topo.kall, topo.lall
        self.nt = int(nt)
        self.dtype = np.dtype(dtype)
        self.names = list(OUT_NAMES) + list(tracer_names)
        self.it = 0
        g, kall, L = self.g, self.kall, self.lall
        os.makedirs(path, exist_ok=True)
        meta = {".zgroup": {"zarr_format": 2},
                ".zattrs": {"title": "pyNICAM-DC output", "dtl": float(dtl), "PRGout_interval": int(interval),
                            "PRGout_interval_2d": int(interval), "PRGout_step0": int(bool(step0))}}
        shape = (self.nt, g, g, kall, L)
        for n in self.names:
            meta[f"{n}/.zarray"] = _zarray(shape, (1, g, g, kall, L), self.dtype, fill="NaN")
            meta[f"{n}/.zattrs"] = {"_ARRAY_DIMENSIONS": ["time", "i", "j", "k", "r"], "coordinates": "GRD_x"}
        cs = ([0] if step0 else []) + [interval * (m + 1) for m in range(self.nt)]
        tsec = np.asarray(cs[:self.nt], dtype="f8") * float(dtl) + float(time_start)
        meta["time/.zarray"] = _zarray((self.nt,), (self.nt,), np.float64, fill="NaN")
        meta["time/.zattrs"] = {"_ARRAY_DIMENSIONS": ["time"], "units": "s", "long_name": "elapsed model time"}
        meta["GRD_x/.zarray"] = _zarray((g, g, L, 3), (g, g, L, 3), self.dtype, fill="NaN")
        meta["GRD_x/.zattrs"] = {"_ARRAY_DIMENSIONS": ["i", "j", "r", "xyz"]}
        for key, v in meta.items():
            self._write_json(key, v)
        self._write_json(".zmetadata", {"zarr_consolidated_format": 1, "metadata": meta})
        self._write_chunk("time", "0", tsec)
        gx = np.asarray(grd_x)
        if gx.ndim == 5:            # (i, j, 1, l, 3) as the reference stores GRD_x
            gx = gx[:, :, 0]
        self._write_chunk("GRD_x", "0.0.0.0", gx.astype(self.dtype))

    def _write_json(self, key, obj):
        p = os.path.join(self.path, key)
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p, "w") as f:
            json.dump(obj, f, indent=2, allow_nan=True)

    def _write_chunk(self, name, key, arr):
        p = os.path.join(self.path, name, key)
        os.makedirs(os.path.dirname(p), exist_ok=True)
        np.ascontiguousarray(arr).astype(self.dtype if name != "time" else np.float64, copy=False).tofile(p)

    def write(self, state: dict):
        """Fill the next time slot from state (PROG (g,g,kall,l,6), PROGq)."""
        if self.it >= self.nt:
            raise IndexError(f"history store holds {self.nt} time slots")
        P = state["PROG"]
        Q = state.get("PROGq")
        fields = [P[..., v] for v in range(6)]
        fields += [Q[..., v] for v in range(len(self.names) - 6)]
        for n, f in zip(self.names, fields):
            self._write_chunk(n, f"{self.it}.0.0.0.0", np.asarray(f)[None].astype(self.dtype))
        self.it += 1


def read_zarr_array(path: str, name: str) -> np.ndarray:
    """Read one array of an uncompressed zarr v2 store (C order, one or more chunks)."""
    meta = json.load(open(os.path.join(path, name, ".zarray")))
    if meta.get("compressor") is not None or meta.get("filters"):
        raise NotImplementedError("compressed zarr chunks")
    shape, chunks, dt = tuple(meta["shape"]), tuple(meta["chunks"]), np.dtype(meta["dtype"])
    out = np.full(shape, np.nan, dtype=dt)
    grid = [-(-s // c) for s, c in zip(shape, chunks)]
    for idx in np.ndindex(*grid):
        p = os.path.join(path, name, ".".join(str(i) for i in idx))
        if not os.path.exists(p):
            continue
        block = np.fromfile(p, dtype=dt).reshape(chunks)
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, shape))
        out[sl] = block[tuple(slice(0, x.stop - x.start) for x in sl)]
    return out
