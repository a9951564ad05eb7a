"""A device-resident batch across chemistry steps (gc_amd_resident_*).

The batch arrays, controls, warm-start step sizes (RSTAT(3) -> RCNTRL(3),
the model's KPPHvalue) and cost keys (ISTAT(3), the next launch order) stay
in HBM between steps; host arrays (numpy, the row-major [nCell, n] layout of
solver.Batch) move through the library's pinned staging.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Resident:
    def __init__(self, ICNTRL, RCNTRL, ATOL, RTOL, mech: str = "fullchem"):
        self.lib = _lib.load(mech)
        self.nvar, self.nfix, self.nreact, _ = _lib.mech_info(self.lib)
        self._keep = [np.ascontiguousarray(ICNTRL, np.int32),
                      np.ascontiguousarray(RCNTRL, np.float64),
                      np.ascontiguousarray(ATOL, np.float64),
                      np.ascontiguousarray(RTOL, np.float64)]
        self.n = self._keep[0].shape[0]
        h = ctypes.c_void_p()
        rc = self.lib.gc_amd_resident_create(self.n, *[_p(a) for a in self._keep], ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"gc_amd_resident_create returned {rc}: {_lib.last_error(self.lib)}")
        self.h = h
        self.t_total = self.t_kernel = 0.0

    def step(self, Tstart, Tend, C=None, RCONST=None, warm=True, outputs=True):
        """Solve one chemistry step; C / RCONST None keep the device's values.
        Returns (C, ISTAT, RSTAT, IERR, Perm used) as numpy arrays, or None
        with outputs=False (nothing copied back)."""
        n = self.n
        Cin = None if C is None else np.ascontiguousarray(C, np.float64)
        Rin = None if RCONST is None else np.ascontiguousarray(RCONST, np.float64)
        if outputs:
            out = (np.empty((n, self.nvar + self.nfix)), np.empty((n, 8), np.int32),
                   np.empty((n, 3)), np.empty(n, np.int32), np.empty(n, np.int32))
        else:
            out = (None,) * 5
        tt, tk = ctypes.c_double(), ctypes.c_double()
        rc = self.lib.gc_amd_resident_step(self.h, float(Tstart), float(Tend), _p(Cin), _p(Rin),
                                           int(bool(warm)), *[_p(a) for a in out],
                                           ctypes.byref(tt), ctypes.byref(tk))
        if rc != 0:
            raise RuntimeError(f"gc_amd_resident_step returned {rc}: {_lib.last_error(self.lib)}")
        self.t_total, self.t_kernel = tt.value, tk.value
        return out if outputs else None

    def close(self):
        if getattr(self, "h", None):
            self.lib.gc_amd_resident_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
