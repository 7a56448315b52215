"""The ``backend="hip"`` branch of the reference's backend selector.

Mirrors ``Backend.configure(backend_name, precision)``
(pynicamdc/share/mod_backend.py:175-283): ``type``, ``ndtype``/``dtype``,
``xp`` (numpy: setup stays on the host) and ``to_numpy``, plus the handle of
the loaded nicamhip library.  Unlike the reference, an unknown backend name
raises instead of leaving ``type`` unset, and the numpy/jax branches are not
provided here (they are the reference's own).
"""
from __future__ import annotations

import numpy as np

from . import native

KNOWN_BACKENDS = ("numpy", "jax", "hip")


class Backend:
    def __init__(self):
        self.type = None
        self.np = np
        self.xp = None
        self.ndtype = None
        self.dtype = None
        self.precision = None
        self.lib = None

    def configure(self, backend_name: str, precision: str, lib_path: str | None = None):
        if backend_name not in KNOWN_BACKENDS:
            raise ValueError(f"unknown backend {backend_name!r}; expected one of {KNOWN_BACKENDS}")
        if backend_name != "hip":
            raise NotImplementedError(
                f"backend {backend_name!r} is the reference's own; this package provides backend='hip'")
        if precision not in ("float64", "mixed", "float32"):
            raise ValueError(f"backend='hip' precision must be 'float64' or 'mixed', not {precision!r}")
        # 'float32' selects the mixed form: float fields, with the tridiagonal
        # solve and the global sums kept in FP64
        self.precision = "float64" if precision == "float64" else "mixed"
        self.ndtype = np.float64 if self.precision == "float64" else np.float32
        self.dtype = self.ndtype
        self.type = "hip"
        self.xp = np
        self.lib = native.load(lib_path)
        return self

    @staticmethod
    def to_numpy(x):
        if hasattr(x, "detach"):
            return x.detach().cpu().numpy()
        return np.asarray(x)

    def maybe_jit(self, fn, *, static_argnames=None):
        """Kernels of this backend are compiled ahead of time; the wrapper is the identity."""
        return fn
