"""Device-resident batch solve on torch tensors (HBM in, HBM out).

Thin binding over ``gc_amd_integrate_batch_device``: torch provides device
memory and the stream, the library does the work.  Arrays use the
reference's Fortran layout, which is row-major ``[nCell, n]`` here:
``C[nCell, NSPEC]`` (VAR then FIX), ``RCONST[nCell, NREACT]``,
``ICNTRL[nCell, 20]`` (int32), ``RCNTRL[nCell, 20]``, ``Perm[nCell]``
(int32, 1-based launch order), ``ATOL``/``RTOL[NVAR]``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import lib as _lib

KERNELS = {"plain": _lib.KERNEL_PLAIN, "persistent": _lib.KERNEL_PERSISTENT}


@dataclass
class Batch:
    C: torch.Tensor
    RCONST: torch.Tensor
    ATOL: torch.Tensor
    RTOL: torch.Tensor
    ICNTRL: torch.Tensor
    RCNTRL: torch.Tensor
    Perm: torch.Tensor
    Tstart: float = 0.0
    Tend: float = 600.0

    @property
    def n(self) -> int:
        return int(self.C.shape[0])

    def to(self, device) -> "Batch":
        f = lambda t: t.to(device) if isinstance(t, torch.Tensor) else t
        return Batch(*(f(getattr(self, k)) for k in
                       ("C", "RCONST", "ATOL", "RTOL", "ICNTRL", "RCNTRL", "Perm")),
                     Tstart=self.Tstart, Tend=self.Tend)

    @staticmethod
    def from_numpy(d: dict, device="cuda") -> "Batch":
        def t(x, dt):
            return torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device=device)
        return Batch(t(d["C"], torch.float64), t(d["RCONST"], torch.float64),
                     t(d["ATOL"], torch.float64), t(d["RTOL"], torch.float64),
                     t(d["ICNTRL"], torch.int32), t(d["RCNTRL"], torch.float64),
                     t(d["Perm"], torch.int32), Tstart=float(d["Tstart"]),
                     Tend=float(d["Tend"]))


@dataclass
class Result:
    C: torch.Tensor
    ISTAT: torch.Tensor
    RSTAT: torch.Tensor
    IERR: torch.Tensor
    t_kernel: float


def _check(b: Batch, mech):
    nvar, nfix, nreact, _ = mech
    n = b.n
    want = {"C": ((n, nvar + nfix), torch.float64), "RCONST": ((n, nreact), torch.float64),
            "ATOL": ((nvar,), torch.float64), "RTOL": ((nvar,), torch.float64),
            "ICNTRL": ((n, 20), torch.int32), "RCNTRL": ((n, 20), torch.float64),
            "Perm": ((n,), torch.int32)}
    for k, (shape, dt) in want.items():
        t = getattr(b, k)
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError(f"{k} must be a device (cuda) tensor")
        if tuple(t.shape) != shape or t.dtype != dt or not t.is_contiguous():
            raise ValueError(f"{k}: expected contiguous {dt} {shape}, got {t.dtype} {tuple(t.shape)}")


def solve(b: Batch, kernel: str = "persistent", chunk: int = 0, mech: str = "fullchem",
          inplace: bool = False, out: Result | None = None) -> Result:
    """Solve every cell of the batch over [Tstart, Tend] on the current device.

    Returns the updated concentrations and the per-cell ISTAT(8), RSTAT(3)
    and IERR exactly as gc_gpu_integrate_batch would."""
    lib = _lib.load(mech)
    _check(b, _lib.mech_info(lib))
    n = b.n
    dev = b.C.device
    if out is None:
        C = b.C if inplace else b.C.clone()
        out = Result(C, torch.empty((n, 8), dtype=torch.int32, device=dev),
                     torch.empty((n, 3), dtype=torch.float64, device=dev),
                     torch.empty((n,), dtype=torch.int32, device=dev), 0.0)
    elif not inplace:
        out.C.copy_(b.C)
    tk = ctypes.c_double(0.0)
    stream = torch.cuda.current_stream(dev).cuda_stream
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    rc = lib.gc_amd_integrate_batch_device(
        KERNELS[kernel], int(chunk), n, float(b.Tstart), float(b.Tend), p(out.C), p(b.RCONST),
        p(b.ATOL), p(b.RTOL), p(b.ICNTRL), p(b.RCNTRL), p(b.Perm), p(out.ISTAT), p(out.RSTAT),
        p(out.IERR), ctypes.c_void_p(stream), ctypes.byref(tk))
    if rc != 0:
        raise RuntimeError(f"gc_amd_integrate_batch_device returned {rc}: {_lib.last_error(lib)}")
    out.t_kernel = tk.value
    return out
