"""ctypes binding of the C ABI in include/gcchemgpu.h.

Loads ``lib/libgcchemgpu_<mechanism>.so`` (built in-tree by ``build.py``).
There is no fallback: if the library is missing, loading raises, so a GPU
run can never silently take a CPU path.
"""
from __future__ import annotations

import ctypes
from functools import lru_cache
from pathlib import Path

from . import PKG_DIR

LIB_DIR = PKG_DIR / "lib"

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
IP = ctypes.POINTER(ctypes.c_int)
DP = ctypes.POINTER(ctypes.c_double)

MAX_DEV = 16


class GcAmdHybrid(ctypes.Structure):
    """struct GcAmdHybrid of include/gcchemgpu.h (the split's state across steps)."""
    _fields_ = [("work_frac", D), ("fixed_frac", I), ("ndev", I), ("rate_gpu", D),
                ("rate_cpu", D), ("dev_rate", D * MAX_DEV), ("n_gpu", I), ("n_dev_use", I),
                ("dev_lo", I * MAX_DEV), ("dev_hi", I * MAX_DEV), ("t_gpu", D),
                ("t_kernel", D), ("t_cpu", D), ("t_region", D)]

    @classmethod
    def initial(cls, work_frac=0.5, fixed_frac=False, ndev=1):
        st = cls()
        st.work_frac, st.fixed_frac, st.ndev = work_frac, int(fixed_frac), ndev
        st.rate_gpu = st.rate_cpu = -1.0
        for d in range(MAX_DEV):
            st.dev_rate[d] = -1.0
        return st


class GcAmdBatchDesc(ctypes.Structure):
    """struct GcAmdBatchDesc of include/gcchemgpu.h (one batch of a multi-batch launch)."""
    _fields_ = [("nCell", I), ("Tstart", D), ("Tend", D), ("C", P), ("RCONST", P), ("ATOL", P),
                ("RTOL", P), ("ICNTRL", P), ("RCNTRL", P), ("Perm", P), ("ISTAT", P),
                ("RSTAT", P), ("IERR", P)]


class GcAmdAdjBatch(ctypes.Structure):
    """struct GcAmdAdjBatch of include/gcchemgpu.h (one batch of a multi-batch sweep)."""
    _fields_ = [("nCell", I), ("Tstart", D), ("Tend", D), ("C", P), ("RCONST", P),
                ("ICNTRL", P), ("Perm", P), ("lamC", P), ("gC", P), ("gRCONST", P), ("IERR", P)]


# gc_amd_progress_fn: void (*)(int batchesDone, long long cellsDone, long long cellsTotal, void*)
PROGRESS_FN = ctypes.CFUNCTYPE(None, I, ctypes.c_longlong, ctypes.c_longlong, P)

# gc_amd_cpu_tail_fn: void (*)(int first, int last, void *user)
CPU_TAIL_FN = ctypes.CFUNCTYPE(None, I, I, P)

# every symbol of include/gcchemgpu.h with its signature (restype, argtypes)
SIGNATURES = {
    "gc_gpu_integrate_batch": (None, [I, D, D, P, P, P, P, P, P, P, P, P, P, DP, DP]),
    "gc_gpu_integrate_batch_dev": (None, [I, I, D, D, P, P, P, P, P, P, P, P, P, P, DP, DP]),
    "gc_gpu_info": (None, [IP, ctypes.c_char_p, I]),
    "gc_gpu_set_device": (I, [I]),
    "gc_gpu_selftest": (None, [IP, IP, DP]),
    "gc_amd_integrate_batch_device": (I, [I, I, I, D, D, P, P, P, P, P, P, P, P, P, P, P, DP]),
    "gc_amd_integrate_batch_noalloc": (I, [I, I, D, D, P, P, P, P, P, P, P, P, P, P, P, DP]),
    "gc_amd_integrate_batch_masked": (I, [I, I, D, D] + [P] * 14 + [DP]),
    "gc_amd_integrate_batches_device": (I, [I, P, P, P, P, DP]),
    "gc_amd_record_batches_device": (I, [I, P, P, P, P, P, ctypes.c_long, P, ctypes.c_long, I, P,
                                         ctypes.POINTER(ctypes.c_long),
                                         ctypes.POINTER(ctypes.c_long), P, P, P, DP]),
    "gc_amd_adjoint_batches_device": (I, [I, P, P, P, P, P, ctypes.c_long, P, ctypes.c_long, I,
                                          P, P, P, P, DP]),
    "gc_amd_resident_create": (I, [I, P, P, P, P, ctypes.POINTER(ctypes.c_void_p)]),
    "gc_amd_resident_step": (I, [P, D, D, P, P, I, P, P, P, P, P, DP, DP]),
    "gc_amd_resident_destroy": (None, [P]),
    "gc_amd_record_batch_device": (I, [I, D, D] + [P] * 14 + [ctypes.c_long, P, ctypes.c_long, I,
                                       P, ctypes.POINTER(ctypes.c_long),
                                       ctypes.POINTER(ctypes.c_long), P, DP]),
    "gc_amd_adjoint_batch_device": (I, [I, D, D] + [P] * 8 + [ctypes.c_long, P, ctypes.c_long, I,
                                        P, P, P, P, P, P, DP]),
    "gc_amd_cost_slots": (I, [I, P, I, P, P]),
    "gc_amd_hybrid_plan": (I, [I, P, I, DP, D, D, I, P, P, IP, IP, P, P]),
    "gc_amd_hybrid_update": (None, [I, I, P, I, P, P, P, D, D, DP, DP, P]),
    "gc_amd_hybrid_step": (I, [I, D, D] + [P] * 10 + [ctypes.POINTER(GcAmdHybrid),
                                                       CPU_TAIL_FN, P]),
    "gc_amd_set_tuning": (I, [I, I, I, I, I]),
    "gc_amd_get_tuning": (None, [IP, IP, IP, IP, IP]),
    "gc_amd_set_gate": (I, [I, I, ctypes.c_ulonglong]),
    "gc_amd_get_gate": (None, [IP, IP, ctypes.POINTER(ctypes.c_ulonglong)]),
    "gc_amd_mech_info": (None, [IP, IP, IP, IP]),
    "gc_amd_workspace_doubles_per_slot": (ctypes.c_long, []),
    "gc_amd_last_error": (ctypes.c_char_p, []),
    "gc_amd_release": (None, []),
}

KERNEL_PLAIN = 0
KERNEL_PERSISTENT = 1


def lib_path(mech: str = "fullchem") -> Path:
    return LIB_DIR / f"libgcchemgpu_{mech}.so"


def _torch_runtime_first() -> None:
    """Import torch, when it is installed, before the library is loaded: torch
    ships its own HIP runtime, and the library's libamdhip64 dependency then
    resolves to that same instance, so device pointers of torch tensors and
    torch's current device are valid inside the library.  (Loaded first, the
    library would bring a second runtime into the process.)"""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


@lru_cache(maxsize=None)
def load(mech: str = "fullchem") -> ctypes.CDLL:
    _torch_runtime_first()
    path = lib_path(mech)
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: build the HIP library first "
            f"(python geos-chem-gpu-hybrid_amd/build.py {mech})")
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def last_error(lib) -> str:
    s = lib.gc_amd_last_error()
    return s.decode() if s else ""


def mech_info(lib):
    v = [ctypes.c_int() for _ in range(4)]
    lib.gc_amd_mech_info(*(ctypes.byref(x) for x in v))
    return tuple(x.value for x in v)


def get_tuning(lib) -> dict:
    v = [ctypes.c_int() for _ in range(5)]
    lib.gc_amd_get_tuning(*(ctypes.byref(x) for x in v))
    keys = ("block", "blocks_per_cu", "cells_per_wave", "grain", "variant")
    return dict(zip(keys, (x.value for x in v)))


def set_tuning(lib, block=0, blocks_per_cu=-1, cells_per_wave=0, grain=0, variant=0):
    """Set launch parameters; 0 (-1 for blocks_per_cu) keeps the current value."""
    rc = lib.gc_amd_set_tuning(block, blocks_per_cu, cells_per_wave, grain, variant)
    if rc != 0:
        raise ValueError(f"gc_amd_set_tuning rejected "
                         f"{block, blocks_per_cu, cells_per_wave, grain, variant}")


def get_gate(lib) -> dict:
    m, c, sp = ctypes.c_int(), ctypes.c_int(), ctypes.c_ulonglong()
    lib.gc_amd_get_gate(ctypes.byref(m), ctypes.byref(c), ctypes.byref(sp))
    return {"mode": m.value, "cells": c.value, "spin": sp.value}


def set_gate(lib, mode=1, cells=0, spin=600000):
    """Overlapped upload of the host-array entry points (gc_amd_set_gate):
    mode 1 on, 0 off, 2 the failure path (chunks never signalled)."""
    rc = lib.gc_amd_set_gate(mode, cells, spin)
    if rc != 0:
        raise ValueError(f"gc_amd_set_gate rejected {mode, cells, spin}")
