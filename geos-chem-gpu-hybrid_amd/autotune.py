"""Autotuner of the persistent Rodas3 kernel's launch parameters (one MI355X).

Coordinate descent over the five launch parameters of gc_amd_set_tuning:

* ``variant``         LU row-segment size in LDS and occupancy target
                      (1: 32 entries, 8 waves/CU; 2: 16, 16; 3: 24, 12)
* ``block``           work-group size (64-256 lanes)
* ``blocks_per_cu``   resident work-groups per compute unit (0: fill the CU)
* ``cells_per_wave``  lanes of each wave that take cells (16-64)
* ``grain``           queue positions a wave claims per atomic (1-4096)

on the 4x5 grid workload (72 x 46 x 72 = 238,464 cells, fullchem, dt 600 s)
in cost order (the step counts of a first solve give the stable descending
sort of src/geos-chem/GeosCore/fullchem_mod.F90:905-981).  Every candidate is
timed by the library's HIP events around the kernel and its results must be
bit-identical to the first solve's (the launch shape never changes the
numbers); a candidate that differs is reported and never selected.

The choice is written to ``tuning/<mech>.json`` under ``launch`` together with
the sweep; measured HBM traffic already stored there under ``profiles`` is
kept.  Usage::

    python geos-chem-gpu-hybrid_amd/autotune.py [--mech fullchem] [--cells N] [--passes 2]
    python geos-chem-gpu-hybrid_amd/autotune.py --mech stress500 --workload stress --axes grain,variant

``--workload stress`` tunes on the synthetic stress cells in their own order
(how the stress batches run, with no previous step to sort by); ``--axes``
restricts the sweep.  build.py compiles each mechanism's stored launch in as
the library's default, so callers of the C-ABI get it without Python.

``--variants`` instead runs one solve of each representative launch shape in
VARIANTS (after a 1 GiB elementwise copy that calibrates the HBM counters and
the first solve), each checked bit-identical, and prints one line per shape in
launch order: the sequence a profiler's per-dispatch records follow.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent

SPACE = {
    "variant": [1, 2, 3],
    "block": [64, 128, 256],
    "blocks_per_cu": [0, 1, 2, 3, 4, 6, 8],
    "cells_per_wave": [16, 32, 48, 64],
    "grain": [1, 16, 64, 256, 1024, 4096],
}
# representative points of the axes, one solve each under --variants
VARIANTS = [
    ("persistent, 32-entry LU segments, 8 waves/CU, 64 cells/wave (selected)", "persistent",
     {"variant": 1, "block": 256, "blocks_per_cu": 2, "cells_per_wave": 64, "grain": 64}),
    ("persistent, 24-entry LU segments, 12 waves/CU", "persistent",
     {"variant": 3, "block": 256, "blocks_per_cu": 3, "cells_per_wave": 64, "grain": 64}),
    ("persistent, 16-entry LU segments, 16 waves/CU", "persistent",
     {"variant": 2, "block": 256, "blocks_per_cu": 4, "cells_per_wave": 64, "grain": 64}),
    ("persistent, 32-entry segments, 32 cells/wave", "persistent",
     {"variant": 1, "block": 256, "blocks_per_cu": 2, "cells_per_wave": 32, "grain": 64}),
    ("persistent, 32-entry segments, 16 cells/wave", "persistent",
     {"variant": 1, "block": 256, "blocks_per_cu": 2, "cells_per_wave": 16, "grain": 64}),
    ("plain (one cell per lane, no refill), 32-entry segments", "plain",
     {"variant": 1, "block": 256, "blocks_per_cu": 2, "cells_per_wave": 64, "grain": 64}),
]
LDS_PER_CU = 160 * 1024
SEG = {1: 32, 2: 16, 3: 24}
WAVES_PER_SIMD = {1: 2, 2: 4, 3: 3}


def _pkg():
    if "gcchem_amd" not in sys.modules:
        spec = importlib.util.spec_from_file_location(
            "gcchem_amd", HERE / "__init__.py", submodule_search_locations=[str(HERE)])
        mod = importlib.util.module_from_spec(spec)
        sys.modules["gcchem_amd"] = mod
        spec.loader.exec_module(mod)
    return sys.modules["gcchem_amd"]


def feasible(t: dict) -> bool:
    """All requested work-groups of one CU fit its LDS and the variant's
    occupancy target (blocks_per_cu 0 is sized by the library itself)."""
    bpc = t["blocks_per_cu"]
    if bpc == 0:
        return True
    waves = bpc * t["block"] // 64
    lds = waves * (SEG[t["variant"]] + 1) * 64 * 8
    return lds <= LDS_PER_CU and waves <= 4 * WAVES_PER_SIMD[t["variant"]]


def log(*a):
    print(*a, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--mech", default="fullchem")
    ap.add_argument("--cells", type=int, default=0, help="first N cells of the grid (0: all)")
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--out", default="")
    ap.add_argument("--variants", action="store_true",
                    help="one solve per shape in VARIANTS (for a profiler), no sweep")
    ap.add_argument("--workload", default="grid4x5", choices=["grid4x5", "stress"],
                    help="grid4x5: the 4x5 grid generator in cost order; stress: "
                         "inputs.stress_cells (seed 46, 150,000 cells unless --cells) in "
                         "their own order, as the stress batches run")
    ap.add_argument("--axes", default="",
                    help="comma-separated launch parameters to sweep (default: all)")
    ap.add_argument("--only", type=int, default=0,
                    help="with --variants: only the shape of this 1-based index")
    args = ap.parse_args(argv)

    import numpy as np
    import torch
    _pkg()
    from gcchem_amd import inputs, lib as L, mechanism, solver, tuning

    mech = mechanism.load(args.mech)
    lib = L.load(args.mech)
    if args.variants:
        # counter calibration: an elementwise copy reading and writing 1 GiB
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log("dispatch group 0: elementwise copy of 1 GiB (reads 1 GiB, writes 1 GiB)")
    if args.workload == "stress":
        d = inputs.stress_cells(mech, args.cells or 150000, seed=46)
        what = f"inputs.stress_cells seed 46, {d['C'].shape[0]} cells, their own order"
    else:
        d = inputs.grid_4x5_cells(mech, seed=1)
        if args.cells:
            d = inputs.subset(d, np.arange(args.cells))
        what = (f"4x5 grid generator (grid_4x5_cells) seed 1, {d['C'].shape[0]} cells, "
                f"cost order")
    n = d["C"].shape[0]
    b = solver.Batch.from_numpy(d)
    C0 = b.C.clone()
    r = solver.solve(b, kernel="persistent", mech=args.mech)
    torch.cuda.synchronize()
    if args.workload != "stress":
        b.Perm = torch.as_tensor(inputs.cost_order(r.ISTAT[:, 2].cpu().numpy()),
                                 device=b.C.device)
    ref_C = r.C.cpu().numpy().view(np.int64)
    ref_I = r.ISTAT.cpu().numpy()
    log(f"autotune {args.mech}: {what}")

    out = None

    def measure(t: dict) -> float | None:
        nonlocal out
        L.set_tuning(lib, **t)
        b.C.copy_(C0)
        out = solver.solve(b, kernel="persistent", mech=args.mech, out=out)
        torch.cuda.synchronize()
        same = (np.array_equal(out.C.cpu().numpy().view(np.int64), ref_C)
                and np.array_equal(out.ISTAT.cpu().numpy(), ref_I))
        if not same:
            log(f"  {t}: results differ from the first solve -- rejected")
            return None
        return out.t_kernel

    if args.variants:
        rows = []
        for k, (label, kernel, t) in enumerate(VARIANTS, start=1):
            if args.only and k != args.only:
                continue
            L.set_tuning(lib, **t)
            b.C.copy_(C0)
            out = solver.solve(b, kernel=kernel, mech=args.mech, out=out)
            torch.cuda.synchronize()
            same = (np.array_equal(out.C.cpu().numpy().view(np.int64), ref_C)
                    and np.array_equal(out.ISTAT.cpu().numpy(), ref_I))
            rows.append({"label": label, "kernel": kernel, **t, "s": out.t_kernel,
                         "bit_identical": bool(same)})
            log(f"dispatch group {k}: {label}: {kernel} {t}: {out.t_kernel:.3f} s, "
                f"bit-identical {same}")
        print(json.dumps({"workload": f"{args.mech} {what}",
                          "variants": rows}), flush=True)
        return 0 if all(r["bit_identical"] for r in rows) else 1

    cur = dict(L.get_tuning(lib))
    cur["blocks_per_cu"] = 0
    sweep = []
    best_t = measure(cur)
    sweep.append({**cur, "s": best_t})
    log(f"  start {cur}: {best_t:.3f} s")
    axes = [a for a in args.axes.split(",") if a] or list(SPACE)
    unknown = set(axes) - set(SPACE)
    if unknown:
        raise SystemExit(f"unknown axes {sorted(unknown)}; choose from {list(SPACE)}")
    for p in range(args.passes):
        changed = False
        for key, values in SPACE.items():
            if key not in axes:
                continue
            for v in values:
                if v == cur[key]:
                    continue
                cand = {**cur, key: v}
                if not feasible(cand):
                    continue
                t0 = time.perf_counter()
                s = measure(cand)
                sweep.append({**cand, "s": s})
                log(f"  pass {p} {key}={v}: {s if s is None else round(s, 3)} s "
                    f"({time.perf_counter() - t0:.1f} s wall)")
                if s is not None and s < best_t * 0.995:
                    best_t, cur, changed = s, cand, True
            log(f"  best after {key}: {cur} {best_t:.3f} s")
        if not changed:
            break

    path = Path(args.out) if args.out else tuning.TUNE_DIR / f"{args.mech}.json"
    data = json.loads(path.read_text()) if path.exists() else {}
    data["launch"] = cur
    data["autotune"] = {"workload": f"{args.mech} {what}, dt 600 s, persistent kernel",
                        "axes": axes,
                        "best_kernel_s": best_t, "cells_per_s": n / best_t,
                        "device": torch.cuda.get_device_name(0), "sweep": sweep}
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(data, indent=1) + "\n")
    log(f"selected {cur}: {best_t:.3f} s = {n / best_t:.0f} cells/s -> {path}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
