"""Discrete adjoint of the batched Rodas3 solve, behind torch.autograd.Function.

The forward pass is the persistent-kernel solve in its recording form: every
accepted step's size goes to a record pool in HBM (one record per accepted
step, chained per cell) and every `every`-th step's start state to a
checkpoint pool.  The backward pass walks each cell's records from the last
to the first on the GPU one segment at a time, replays the segment forward
from its checkpoint to regain the start states, recomputes each step's
Jacobian, LU factors and stages, and sweeps the adjoint through the four
stages in reverse with transposed solves on the
same factors (KppSolveTR order, gckpp_LinearAlgebra.F90:2327-3701), the
stoichiometry-weighted adjoint of Fun_SPLIT, the rate-constant and
fixed-species derivatives, and the second-order term of the Jacobian's
dependence on the state.  Step sizes and the accept/reject sequence are held
fixed (the discrete map of the accepted steps is differentiated, as KPP's
Rosenbrock adjoint does).

Gradients are returned for the initial concentrations C (variable and fixed
species) and RCONST.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import inputs as _inputs
from . import lib as _lib
from .solver import Batch, Result, _check

DEFAULT_STEPS_PER_CELL = 96
STATE_BUDGET_BYTES = 32e9   # checkpoint states kept in HBM before thinning them out
STAGE_BUDGET_BYTES = 96e9   # stage vectors kept when every step is a checkpoint


@dataclass
class Records:
    H: torch.Tensor      # [cap] step size of each accepted step
    prev: torch.Tensor   # [cap] int32, previous record of the cell (-1: first)
    yslot: torch.Tensor  # [cap] int32, checkpoint state slot or -1
    head: torch.Tensor   # [nCell] int32, last record of each cell
    Y: torch.Tensor      # [ycap, NVAR] checkpoint states
    K: torch.Tensor | None  # [ycap, 4, NVAR] stage vectors of the checkpoint steps
    every: int           # checkpoint interval
    used: int
    yused: int
    t_kernel: float


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def record_solve(b: Batch, mech: str = "fullchem", every: int = 0, cap: int = 0,
                 state_budget: float = STATE_BUDGET_BYTES,
                 stage_budget: float = STAGE_BUDGET_BYTES) -> tuple[Result, Records]:
    """Forward solve of the batch (persistent kernel) recording every accepted
    step, with the start state of every `every`-th step of each cell as a
    checkpoint.  every = 0 picks the smallest interval whose states fit
    `state_budget` bytes; pools that turn out too small are re-sized to the
    exact need and the solve is repeated.  When every step is a checkpoint
    and their stage vectors fit `stage_budget` bytes, those are kept too, so
    that the sweep recomputes only each step's Jacobian and factors."""
    lib = _lib.load(mech)
    nvar = _lib.mech_info(lib)[0]
    _check(b, _lib.mech_info(lib))
    n, dev = b.n, b.C.device
    state_bytes = nvar * 8
    cap = int(cap or max(1, n * DEFAULT_STEPS_PER_CELL))
    ev = int(every) if every else max(1, -(-cap * state_bytes // int(state_budget)))
    for _ in range(3):
        ycap = max(1, min(cap, -(-cap // ev) + n))
        out = Result(b.C.clone(), torch.empty((n, 8), dtype=torch.int32, device=dev),
                     torch.empty((n, 3), dtype=torch.float64, device=dev),
                     torch.empty((n,), dtype=torch.int32, device=dev), 0.0)
        recH = torch.empty((cap,), dtype=torch.float64, device=dev)
        prev = torch.empty((cap,), dtype=torch.int32, device=dev)
        yslot = torch.empty((cap,), dtype=torch.int32, device=dev)
        head = torch.empty((n,), dtype=torch.int32, device=dev)
        recY = torch.empty((ycap, nvar), dtype=torch.float64, device=dev)
        recK = None
        if ev == 1 and ycap * 4 * state_bytes <= stage_budget:
            recK = torch.empty((ycap, 4, nvar), dtype=torch.float64, device=dev)
        used, yused = ctypes.c_long(0), ctypes.c_long(0)
        tk = ctypes.c_double(0.0)
        stream = torch.cuda.current_stream(dev).cuda_stream
        rc = lib.gc_amd_record_batch_device(
            n, float(b.Tstart), float(b.Tend), _p(out.C), _p(b.RCONST), _p(b.ATOL), _p(b.RTOL),
            _p(b.ICNTRL), _p(b.RCNTRL), _p(b.Perm), _p(out.ISTAT), _p(out.RSTAT), _p(out.IERR),
            _p(recH), _p(prev), _p(yslot), _p(head), cap, _p(recY), ycap, ev,
            None if recK is None else _p(recK),
            ctypes.byref(used), ctypes.byref(yused), ctypes.c_void_p(stream), ctypes.byref(tk))
        if rc == -4:
            del recH, prev, yslot, recY, recK
            cap = max(cap, int(used.value))
            if not every:
                ev = max(ev, -(-int(used.value) * state_bytes // int(state_budget)))
            continue
        if rc != 0:
            raise RuntimeError(f"gc_amd_record_batch_device returned {rc}: {_lib.last_error(lib)}")
        out.t_kernel = tk.value
        return out, Records(recH, prev, yslot, head, recY, recK, ev, int(used.value),
                            int(yused.value), tk.value)
    raise RuntimeError("record pool sizing failed")


def adjoint_sweep(b: Batch, rec: Records, lamC: torch.Tensor, mech: str = "fullchem",
                  perm: torch.Tensor | None = None):
    """dL/dC0 [n, NSPEC], dL/dRCONST [n, NREACT], IERR [n] and the kernel
    time, for dL/dC_final = lamC [n, NSPEC]."""
    lib = _lib.load(mech)
    n, dev = b.n, b.C.device
    lamC = lamC.to(device=dev, dtype=torch.float64).contiguous()
    if tuple(lamC.shape) != tuple(b.C.shape):
        raise ValueError(f"lamC must have shape {tuple(b.C.shape)}")
    gC = torch.empty_like(b.C)
    gR = torch.empty_like(b.RCONST)
    ierr = torch.empty((n,), dtype=torch.int32, device=dev)
    perm = b.Perm if perm is None else perm.to(device=dev, dtype=torch.int32).contiguous()
    tk = ctypes.c_double(0.0)
    stream = torch.cuda.current_stream(dev).cuda_stream
    rc = lib.gc_amd_adjoint_batch_device(
        n, float(b.Tstart), float(b.Tend), _p(b.C), _p(b.RCONST), _p(b.ICNTRL), _p(perm),
        _p(rec.H), _p(rec.prev), _p(rec.yslot), _p(rec.head), int(rec.H.shape[0]), _p(rec.Y),
        int(rec.Y.shape[0]), int(rec.every), None if rec.K is None else _p(rec.K),
        _p(lamC), _p(gC), _p(gR), _p(ierr),
        ctypes.c_void_p(stream), ctypes.byref(tk))
    if rc != 0:
        raise RuntimeError(f"gc_amd_adjoint_batch_device returned {rc}: {_lib.last_error(lib)}")
    return gC, gR, ierr, tk.value


def records_of_cell(rec: Records, cell: int) -> np.ndarray:
    """Step sizes of one cell's accepted steps, first to last (host copy)."""
    head = int(rec.head[cell].item())
    prev = rec.prev.cpu().numpy()
    H = rec.H.cpu().numpy()
    out = []
    s = head
    while s >= 0:
        out.append(H[s])
        s = int(prev[s])
    return np.array(out[::-1])


class Rodas3Solve(torch.autograd.Function):
    """C_final = Rodas3(C0, RCONST) over [Tstart, Tend] for every cell of the
    batch; differentiable in C0 and RCONST through the discrete adjoint."""

    @staticmethod
    def forward(ctx, C, RCONST, template: Batch, mech: str):
        b = Batch(C.detach().contiguous(), RCONST.detach().contiguous(), template.ATOL,
                  template.RTOL, template.ICNTRL, template.RCNTRL, template.Perm,
                  Tstart=template.Tstart, Tend=template.Tend)
        out, rec = record_solve(b, mech)
        if not bool((out.IERR == 1).all()):
            raise RuntimeError("Rodas3Solve: some cells failed in the forward solve")
        ctx.batch, ctx.rec, ctx.mech = b, rec, mech
        # the reverse sweep runs costly cells first
        nacc = out.ISTAT[:, 3].cpu().numpy()
        ctx.perm = torch.as_tensor(_inputs.cost_order(nacc), device=C.device)
        ctx.istat = out.ISTAT
        return out.C

    @staticmethod
    def backward(ctx, gout):
        gC, gR, ierr, _ = adjoint_sweep(ctx.batch, ctx.rec, gout, ctx.mech, perm=ctx.perm)
        if not bool((ierr == 1).all()):
            raise RuntimeError("Rodas3Solve: adjoint sweep failed for some cells")
        return gC, gR, None, None


def solve(C: torch.Tensor, RCONST: torch.Tensor, template: Batch,
          mech: str = "fullchem") -> torch.Tensor:
    """Differentiable batch solve: C [n, NSPEC], RCONST [n, NREACT] (device
    tensors, may require grad); the other inputs come from `template`."""
    return Rodas3Solve.apply(C, RCONST, template, mech)


def record_solve_batches(batches: list[Batch], mech: str = "fullchem", every: int = 0,
                         cap: int = 0, state_budget: float = STATE_BUDGET_BYTES,
                         stage_budget: float = STAGE_BUDGET_BYTES, progress=None):
    """Recording forward solve of several batches in one persistent launch
    (gc_amd_record_batches_device): their results and one Records whose head
    covers the batches' cells one after the other.  every, cap and the
    budgets as in record_solve (pools too small are re-sized to the exact
    need and the launch repeated); progress(batches_done, cells_done,
    cells_total) as in solver.solve_batches."""
    from .solver import _desc, _results
    lib = _lib.load(mech)
    nvar = _lib.mech_info(lib)[0]
    for b in batches:
        _check(b, _lib.mech_info(lib))
    n = sum(b.n for b in batches)
    dev = batches[0].C.device
    state_bytes = nvar * 8
    cap = int(cap or max(1, n * DEFAULT_STEPS_PER_CELL))
    ev = int(every) if every else max(1, -(-cap * state_bytes // int(state_budget)))
    cb = _lib.PROGRESS_FN(lambda nb, c, t, u: progress(nb, c, t)) if progress else None
    for _ in range(3):
        ycap = max(1, min(cap, -(-cap // ev) + n))
        outs = _results(batches)
        recH = torch.empty((cap,), dtype=torch.float64, device=dev)
        prev = torch.empty((cap,), dtype=torch.int32, device=dev)
        yslot = torch.empty((cap,), dtype=torch.int32, device=dev)
        head = torch.empty((n,), dtype=torch.int32, device=dev)
        recY = torch.empty((ycap, nvar), dtype=torch.float64, device=dev)
        recK = None
        if ev == 1 and ycap * 4 * state_bytes <= stage_budget:
            recK = torch.empty((ycap, 4, nvar), dtype=torch.float64, device=dev)
        descs = (_lib.GcAmdBatchDesc * len(batches))(*[_desc(b, r) for b, r in zip(batches, outs)])
        used, yused, tk = ctypes.c_long(0), ctypes.c_long(0), ctypes.c_double(0.0)
        stream = torch.cuda.current_stream(dev).cuda_stream
        rc = lib.gc_amd_record_batches_device(
            len(batches), descs, _p(recH), _p(prev), _p(yslot), _p(head), cap, _p(recY), ycap,
            ev, None if recK is None else _p(recK), ctypes.byref(used), ctypes.byref(yused),
            cb, None, ctypes.c_void_p(stream), ctypes.byref(tk))
        if rc == -4:
            del recH, prev, yslot, recY, recK
            cap = max(cap, int(used.value))
            if not every:
                ev = max(ev, -(-int(used.value) * state_bytes // int(state_budget)))
            continue
        if rc != 0:
            raise RuntimeError(f"gc_amd_record_batches_device returned {rc}: {_lib.last_error(lib)}")
        for r in outs:
            r.t_kernel = tk.value
        return outs, Records(recH, prev, yslot, head, recY, recK, ev, int(used.value),
                             int(yused.value), tk.value)
    raise RuntimeError("record pool sizing failed")


def adjoint_sweep_batches(batches: list[Batch], rec: Records, lams: list[torch.Tensor],
                          mech: str = "fullchem", progress=None, perms=None):
    """Reverse sweep of a record_solve_batches recording in one launch: per
    batch (dL/dC0, dL/dRCONST, IERR), and the kernel seconds; progress as in
    record_solve_batches.  perms (one 1-based launch order per batch, e.g.
    inputs.cost_order of the recording's accepted steps, so that each wave's
    64 cells have alike costs) replaces the batches' own Perm."""
    lib = _lib.load(mech)
    outs, descs = [], []
    if perms is None:
        perms = [b.Perm for b in batches]
    perms = [torch.as_tensor(p, device=b.C.device).to(torch.int32).contiguous()
             for b, p in zip(batches, perms)]
    for b, lam, perm in zip(batches, lams, perms):
        lam = lam.to(device=b.C.device, dtype=torch.float64).contiguous()
        gC, gR = torch.empty_like(b.C), torch.empty_like(b.RCONST)
        ierr = torch.empty((b.n,), dtype=torch.int32, device=b.C.device)
        outs.append((gC, gR, ierr, lam))
        p = lambda t: t.data_ptr()
        descs.append(_lib.GcAmdAdjBatch(b.n, float(b.Tstart), float(b.Tend), p(b.C), p(b.RCONST),
                                        p(b.ICNTRL), p(perm), p(lam), p(gC), p(gR), p(ierr)))
    arr = (_lib.GcAmdAdjBatch * len(descs))(*descs)
    cb = _lib.PROGRESS_FN(lambda nb, c, t, u: progress(nb, c, t)) if progress else None
    tk = ctypes.c_double(0.0)
    stream = torch.cuda.current_stream(batches[0].C.device).cuda_stream
    rc = lib.gc_amd_adjoint_batches_device(
        len(descs), arr, _p(rec.H), _p(rec.prev), _p(rec.yslot), _p(rec.head),
        int(rec.H.shape[0]), _p(rec.Y), int(rec.Y.shape[0]), int(rec.every),
        None if rec.K is None else _p(rec.K), cb, None, ctypes.c_void_p(stream),
        ctypes.byref(tk))
    if rc != 0:
        raise RuntimeError(f"gc_amd_adjoint_batches_device returned {rc}: {_lib.last_error(lib)}")
    return [o[:3] for o in outs], tk.value
