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


def mask_words(n: int) -> int:
    """32-bit words of an auto-reduce mask of n flags."""
    return (n + 31) // 32


def solve(b: Batch, kernel: str = "persistent", chunk: int = 0, mech: str = "fullchem",
          inplace: bool = False, out: Result | None = None, masks=None) -> Result:
    """Solve every cell of the batch over [Tstart, Tend] on the current device.

    Returns the updated concentrations and the per-cell ISTAT(8), RSTAT(3)
    and IERR exactly as gc_gpu_integrate_batch would.  masks = (doFun,
    doJvs, doSlv), int32 device tensors [nCell, ceil(n/32)] of the cells'
    auto-reduce flags (flag k in bit k%32 of word k/32; None: every flag on),
    solves through gc_amd_integrate_batch_masked (no chunking)."""
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
    if masks is not None:
        nvar, _, _, nnz = _lib.mech_info(lib)
        if chunk:
            raise ValueError("masked solves run in one launch (chunk = 0)")
        mk = []
        for t, nflag in zip(masks, (nvar, nnz, nvar)):
            if t is None:
                mk.append(None)
                continue
            want = (n, mask_words(nflag))
            if (not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.int32
                    or tuple(t.shape) != want or not t.is_contiguous()):
                raise ValueError(f"mask: expected a contiguous int32 cuda tensor {want}")
            mk.append(p(t))
        rc = lib.gc_amd_integrate_batch_masked(
            KERNELS[kernel], n, float(b.Tstart), float(b.Tend), p(out.C), p(b.RCONST),
            p(b.ATOL), p(b.RTOL), p(b.ICNTRL), p(b.RCNTRL), p(b.Perm), *mk, p(out.ISTAT),
            p(out.RSTAT), p(out.IERR), ctypes.c_void_p(stream), ctypes.byref(tk))
        if rc != 0:
            raise RuntimeError(f"gc_amd_integrate_batch_masked returned {rc}: {_lib.last_error(lib)}")
        out.t_kernel = tk.value
        return out
    rc = lib.gc_amd_integrate_batch_device(
        KERNELS[kernel], int(chunk), n, float(b.Tstart), float(b.Tend), p(out.C), p(b.RCONST),
        p(b.ATOL), p(b.RTOL), p(b.ICNTRL), p(b.RCNTRL), p(b.Perm), p(out.ISTAT), p(out.RSTAT),
        p(out.IERR), ctypes.c_void_p(stream), ctypes.byref(tk))
    if rc != 0:
        raise RuntimeError(f"gc_amd_integrate_batch_device returned {rc}: {_lib.last_error(lib)}")
    out.t_kernel = tk.value
    return out


def _desc(b: Batch, r: Result):
    p = lambda t: t.data_ptr()
    return _lib.GcAmdBatchDesc(b.n, float(b.Tstart), float(b.Tend), p(r.C), p(b.RCONST),
                               p(b.ATOL), p(b.RTOL), p(b.ICNTRL), p(b.RCNTRL), p(b.Perm),
                               p(r.ISTAT), p(r.RSTAT), p(r.IERR))


def _results(batches):
    out = []
    for b in batches:
        dev = b.C.device
        out.append(Result(b.C.clone(), torch.empty((b.n, 8), dtype=torch.int32, device=dev),
                          torch.empty((b.n, 3), dtype=torch.float64, device=dev),
                          torch.empty((b.n,), dtype=torch.int32, device=dev), 0.0))
    return out


def solve_batches(batches: list[Batch], mech: str = "fullchem",
                  progress=None) -> tuple[list[Result], float]:
    """Solve independent batches (own sizes, launch orders and time spans) in
    one persistent launch (gc_amd_integrate_batches_device); returns their
    results and the kernel seconds.  progress(batches_done, cells_done,
    cells_total) is called as cells finish."""
    lib = _lib.load(mech)
    for b in batches:
        _check(b, _lib.mech_info(lib))
    outs = _results(batches)
    descs = (_lib.GcAmdBatchDesc * len(batches))(*[_desc(b, r) for b, r in zip(batches, outs)])
    tk = ctypes.c_double(0.0)
    stream = torch.cuda.current_stream(batches[0].C.device).cuda_stream
    cb = _lib.PROGRESS_FN(lambda nb, c, t, u: progress(nb, c, t)) if progress else None
    rc = lib.gc_amd_integrate_batches_device(len(batches), descs, cb, None,
                                             ctypes.c_void_p(stream), ctypes.byref(tk))
    if rc != 0:
        raise RuntimeError(f"gc_amd_integrate_batches_device returned {rc}: {_lib.last_error(lib)}")
    for r in outs:
        r.t_kernel = tk.value
    return outs, tk.value
