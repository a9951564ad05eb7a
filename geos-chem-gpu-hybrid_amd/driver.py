"""Standalone batch driver over the torch binding.

The Python counterpart of the reference's standalone benchmark driver
(standalone/kpp_batch.F90:16-753): it loads a batch of cells, solves it on
the GPU in one of the library's launch forms, and prints the reference's own
throughput and accuracy lines against a CPU result the caller supplies.

    python driver.py <batch.npz> <mode> [ncap] [--ref ref.npz] [--mech NAME]
                     [--chunk N] [--out out.npz]

Batch file (numpy .npz, row-major [nCell, n] arrays as in solver.Batch):
C[nCell, NSPEC], RCONST[nCell, NREACT], ICNTRL[nCell, 20] (int32),
RCNTRL[nCell, 20], ATOL[NVAR], RTOL[NVAR], Tstart, Tend; Perm[nCell] (1-based)
is optional (identity).  The reference driver reads KPP-Standalone sample
files instead (kpp_batch.F90:101-132); the harvested cells are not available,
so the batch format is this package's own.  ``ref.npz`` holds ``Y[nCell,
NVAR]`` (or the whole C) from the caller's CPU solve of the same cells.

Modes (the GPU modes of kpp_batch.F90:143-753 in this library's launch forms):

* ``plain``       one thread per cell, the reference's ``gpu``/``gpul`` launch
* ``persistent``  the work-queue kernel in the batch's launch order
* ``chunked``     persistent in chunks of ``--chunk`` cells (default 32768,
                  clamped to [1024, nCell]: the reference's low-memory mode,
                  gpulib/src/gc_gpu_batch_api.F90:231-256), replayed as a graph
* ``sorted``      plain in identity order, then in ascending step-count order
                  from that run (the ``gpum`` control, kpp_batch.F90:641-688
                  with Build_Perms :847-880), and the bit-identity of the two
* ``host``        the host-array entry point gc_gpu_integrate_batch (the call
                  GEOS-Chem makes, staging included in the time)

Report lines follow Report (kpp_batch.F90:883-900): time, cells/s, nFail
(IERR /= 1) and maxrel = max |Y - Yref| / max(|Yref|, 1e-20) over all cells and
variable species (also the ``gpu`` mode's gate, :262-272); with a reference
the Gate_Report lines (:902-926) add the count of differing values and the
largest relative difference among them (denominator max(|Yref|, 1e-30)).
"""
from __future__ import annotations

import argparse
import ctypes
import sys
import time
from pathlib import Path

import numpy as np

BATCH_KEYS = ("C", "RCONST", "ICNTRL", "RCNTRL", "ATOL", "RTOL")
MODES = ("plain", "persistent", "chunked", "sorted", "host")


def load_batch(path, ncap: int | None = None) -> dict:
    """Read a batch file, keeping the first `ncap` cells when given."""
    with np.load(path, allow_pickle=False) as z:
        missing = [k for k in BATCH_KEYS + ("Tstart", "Tend") if k not in z.files]
        if missing:
            raise ValueError(f"{path}: missing {', '.join(missing)}")
        d = {k: z[k] for k in z.files}
    n = d["C"].shape[0]
    if ncap is not None:
        if ncap < 1:
            raise ValueError("ncap must be >= 1")
        n = min(n, int(ncap))
    out = {k: np.ascontiguousarray(d[k][:n]) for k in ("C", "RCONST", "ICNTRL", "RCNTRL")}
    out["ICNTRL"] = out["ICNTRL"].astype(np.int32)
    out["ATOL"] = np.ascontiguousarray(d["ATOL"], dtype=np.float64)
    out["RTOL"] = np.ascontiguousarray(d["RTOL"], dtype=np.float64)
    out["Tstart"], out["Tend"] = float(d["Tstart"]), float(d["Tend"])
    perm = d.get("Perm")
    if perm is None or ncap is not None:
        perm = np.arange(1, n + 1, dtype=np.int32)
    out["Perm"] = np.ascontiguousarray(perm, dtype=np.int32)
    return out


def save_batch(path, d: dict) -> None:
    np.savez(path, **{k: d[k] for k in BATCH_KEYS + ("Tstart", "Tend", "Perm") if k in d})


def report(tag: str, t: float, Y, Yref, IERR) -> dict:
    """Report (kpp_batch.F90:883-900): nFail = cells with IERR /= 1, maxrel =
    max over all values of |Y - Yref| / max(|Yref|, 1e-20) (no reference:
    maxrel is None)."""
    n = int(np.asarray(IERR).size)
    r = {"tag": tag, "t": t, "cells_per_s": n / t if t > 0 else float("inf"),
         "nFail": int(np.count_nonzero(np.asarray(IERR) != 1)), "maxrel": None}
    if Yref is not None:
        Y, Yref = np.asarray(Y, np.float64), np.asarray(Yref, np.float64)
        r["maxrel"] = float(np.max(np.abs(Y - Yref) / np.maximum(np.abs(Yref), 1.0e-20)))
    return r


def gate_report(tag: str, Y, Yref, IERR) -> dict:
    """Gate_Report (kpp_batch.F90:902-926): values that differ at all, and the
    largest relative difference among them with denominator
    max(|Yref|, 1e-30); cells with IERR /= 1."""
    Y, Yref = np.asarray(Y, np.float64), np.asarray(Yref, np.float64)
    diff = Y != Yref
    rel = np.abs(Y - Yref)[diff] / np.maximum(np.abs(Yref[diff]), 1.0e-30)
    return {"tag": tag, "differing": int(np.count_nonzero(diff)), "of": int(Y.size),
            "maxrel": float(rel.max()) if rel.size else 0.0,
            "nFail": int(np.count_nonzero(np.asarray(IERR) != 1))}


def build_perms(nsteps):
    """Identity order and the ascending stable counting sort by step count
    (Build_Perms, kpp_batch.F90:847-880), both 1-based."""
    s = np.asarray(nsteps, dtype=np.int64)
    if s.size and s.min() < 0:
        raise ValueError("negative step count")
    n = s.size
    cnt = np.bincount(s, minlength=int(s.max()) + 1 if n else 1)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    perm = np.empty(n, dtype=np.int32)
    for i in range(n):          # stable: cells of one count keep their order
        perm[start[s[i]]] = i + 1
        start[s[i]] += 1
    return np.arange(1, n + 1, dtype=np.int32), perm


def _print(r: dict) -> None:
    line = (f"{r['tag']}  t= {r['t']:10.3f} s  {r['cells_per_s']:10.1f} cells/s  "
            f"nFail= {r['nFail']:8d}")
    if r.get("maxrel") is not None:
        line += f"  maxrel= {r['maxrel']:11.3e}"
    print(line, flush=True)


def _print_gate(g: dict) -> None:
    print(f"{g['tag']} differing values: {g['differing']:10d}  of {g['of']:12d}", flush=True)
    print(f"{g['tag']} max relative diff: {g['maxrel']:12.4e}   cells with IERR/=1: "
          f"{g['nFail']:8d}", flush=True)


def _device_solve(solver, torch, d, kernel, mech, chunk=0, perm=None):
    b = solver.Batch.from_numpy(d)
    if perm is not None:
        b.Perm = torch.as_tensor(perm, dtype=torch.int32, device=b.C.device)
    r = solver.solve(b, kernel=kernel, chunk=chunk, mech=mech)
    torch.cuda.synchronize()
    return (r.C.cpu().numpy(), r.ISTAT.cpu().numpy(), r.RSTAT.cpu().numpy(),
            r.IERR.cpu().numpy(), r.t_kernel)


def _host_solve(lib_mod, d, mech):
    lib = lib_mod.load(mech)
    n = d["C"].shape[0]
    C = np.ascontiguousarray(d["C"].copy())
    IS, RS, IE = np.zeros((n, 8), np.int32), np.zeros((n, 3)), np.zeros(n, np.int32)
    tt, tk = ctypes.c_double(), ctypes.c_double()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib.gc_gpu_integrate_batch(n, d["Tstart"], d["Tend"], p(C), p(d["RCONST"]), p(d["ATOL"]),
                               p(d["RTOL"]), p(d["ICNTRL"]), p(d["RCNTRL"]), p(d["Perm"]), p(IS),
                               p(RS), p(IE), ctypes.byref(tt), ctypes.byref(tk))
    return C, IS, RS, IE, tt.value


def run(d: dict, mode: str, mech: str = "fullchem", chunk: int = 32768,
        Yref=None) -> list[dict]:
    """Solve the batch in `mode`, print the report lines and return them;
    the last entry's "result" holds (C, ISTAT, RSTAT, IERR) of the final run."""
    if mode not in MODES:
        raise ValueError(f"unknown mode {mode!r}; one of {', '.join(MODES)}")
    import torch
    from gcchem_amd import lib as lib_mod
    from gcchem_amd import solver
    if not torch.cuda.is_available():
        raise RuntimeError("the driver's modes run on the GPU")
    n = d["C"].shape[0]
    nvar = lib_mod.mech_info(lib_mod.load(mech))[0]
    ref = None if Yref is None else np.asarray(Yref)[:n, :nvar]
    print(f"Loaded cells:  {n:8d}", flush=True)
    out = []

    def finish(tag, res):
        C, IS, RS, IE, t = res
        r = report(tag, t, C[:, :nvar], ref, IE)
        _print(r)
        if ref is not None:
            g = gate_report(tag, C[:, :nvar], ref, IE)
            _print_gate(g)
            r["gate"] = g
        r["result"] = (C, IS, RS, IE)
        out.append(r)

    if mode == "plain":
        finish("GPU plain kernel", _device_solve(solver, torch, d, "plain", mech))
    elif mode == "persistent":
        finish("GPU persistent kernel", _device_solve(solver, torch, d, "persistent", mech))
    elif mode == "chunked":
        c = min(max(int(chunk), 1024), n) if n >= 1024 else n
        finish(f"GPU persistent, chunks of {c}",
               _device_solve(solver, torch, d, "persistent", mech, chunk=c))
    elif mode == "host":
        t0 = time.perf_counter()
        res = _host_solve(lib_mod, d, mech)
        finish("GPU host-array call", res[:4] + (time.perf_counter() - t0,))
    else:  # sorted
        first = _device_solve(solver, torch, d, "plain", mech,
                              perm=np.arange(1, n + 1, dtype=np.int32))
        finish("GPU plain identity", first)
        steps = first[1][:, 2]
        _, perm = build_perms(steps)
        print(f"Cost sort: steps min= {steps.min():6d}  max= {steps.max():6d}  "
              f"mean= {steps.mean():8.2f}", flush=True)
        finish("GPU plain sorted  ", _device_solve(solver, torch, d, "plain", mech, perm=perm))
        same = all(np.array_equal(a, b) for a, b in zip(out[0]["result"], out[1]["result"]))
        print(f"sorted vs identity: {'bit-identical' if same else 'DIFFERENT'}", flush=True)
        out[-1]["orders_identical"] = same
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("batch")
    ap.add_argument("mode", choices=MODES)
    ap.add_argument("ncap", nargs="?", type=int, default=None)
    ap.add_argument("--ref", default=None, help="npz with Y[nCell, NVAR] of a CPU solve")
    ap.add_argument("--mech", default="fullchem")
    ap.add_argument("--chunk", type=int, default=32768)
    ap.add_argument("--out", default=None, help="write C, ISTAT, RSTAT, IERR here")
    a = ap.parse_args(argv)
    d = load_batch(a.batch, a.ncap)
    Yref = None
    if a.ref:
        with np.load(a.ref, allow_pickle=False) as z:
            Yref = z["Y"]
        if Yref.shape[0] < d["C"].shape[0]:
            raise ValueError("the reference holds fewer cells than the batch")
    res = run(d, a.mode, a.mech, a.chunk, Yref)
    if a.out:
        C, IS, RS, IE = res[-1]["result"]
        np.savez(a.out, C=C, ISTAT=IS, RSTAT=RS, IERR=IE)
    return 0 if all(r["nFail"] == 0 for r in res) else 1


def _load_package():
    import importlib.util
    pkg = Path(__file__).resolve().parent
    spec = importlib.util.spec_from_file_location(
        "gcchem_amd", pkg / "__init__.py", submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gcchem_amd"] = mod
    spec.loader.exec_module(mod)


if __name__ == "__main__":
    _load_package()
    sys.exit(main())
