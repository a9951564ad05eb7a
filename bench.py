"""Benchmark: persistent vs plain batched Rodas3 on the 4x5 grid, one MI355X.

Workload: GEOS-Chem fullchem mechanism (353 variable species, 1058
reactions), 72 x 46 columns x 72 levels = 238,464 cells, FP64, dt = 600 s,
synthetic seeded inputs (standard-atmosphere T/P by level with noise, SZA
from a lon/lat grid at a random UTC hour).  The launch order is the model's
cost order: a first (untimed) solve of the same cells stands for the previous
chemistry step, and its per-cell step counts give the stable descending sort
of src/geos-chem/GeosCore/fullchem_mod.F90:905-981.

One "step" is one batch solve of all cells with the inputs resident in HBM.
`value` is the persistent kernel's cells/s; the plain one-thread-per-cell
kernel is timed on the same cells in the same order for the ratio.  The
adjoint/forward cost ratio is measured on 65,536 cells of the same generator
(recording forward + reverse sweep over the forward solve).  The CPU
baseline is the oracle (a plain-C restatement of the reference Rodas3, libm
pow) on a bounded sample of the same cells with OpenMP on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
PKG = REPO / "geos-chem-gpu-hybrid_amd"
METRIC = "cells/s persistent-autotuned vs plain batched Rodas3, 238k cells FP64; adj/fwd"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
HBM_PEAK_GBPS = 8000.0


def roofline(mech, n_cells, t, flops, traffic):
    """The kernel's roofline entry.  The binding resource is HBM (the per-cell
    workspace streams through it), so `achieved` is the algorithmic bytes of
    the launch -- each cell's compulsory inputs and outputs,
    mech.compulsory_bytes_per_cell() (14,440 B for fullchem: the solve writes
    back the NVAR part of C) -- over the kernel time, and `traffic` the bytes
    the HBM counters measured per launch; the FP64 rate of the algorithmic
    flops (solver counters x per-operation flop counts) is the second figure."""
    alg = mech.compulsory_bytes_per_cell() * n_cells
    gbps = alg / t / 1e9
    roof = {"bound": "hbm", "achieved": round(gbps, 3), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(gbps / HBM_PEAK_GBPS, 7), "traffic": traffic,
            "algorithmic_bytes": alg,
            "fp64": {"achieved": round(flops / t / 1e12, 4), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(flops / t / 1e12 / FP64_PEAK_TFLOPS, 5)},
            "note": "no MFMA: per-cell sparse LU is not a dense contraction; the workspace "
                    "of the cells in flight (~102 KB per cell) does not fit on chip, so "
                    "traffic / algorithmic_bytes is the workspace streaming"}
    if traffic:
        roof["measured_traffic_rate"] = {"achieved": round(traffic / t / 1e9, 1),
                                         "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                                         "frac": round(traffic / t / 1e9 / HBM_PEAK_GBPS, 4)}
    return roof


def load_package():
    if "gcchem_amd" in sys.modules:
        return sys.modules["gcchem_amd"]
    spec = importlib.util.spec_from_file_location(
        "gcchem_amd", PKG / "__init__.py", submodule_search_locations=[str(PKG)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gcchem_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(mech, d, perm, target_s=15.0):
    """Oracle (libm pow) on cells evenly spaced in the launch order, sized to
    about target_s seconds of wall time on all host cores: a calibration run
    of 16 cells per thread fixes the per-cell rate, the sample is scaled from
    it."""
    sys.path.insert(0, str(REPO / "oracle"))
    import numpy as np
    from oracle import Oracle
    from gcchem_amd import inputs
    o = Oracle(mech.ir, pow_mode="libm")
    # the host's share of cores for one GPU is 16 threads
    threads = min(len(os.sched_getaffinity(0)) or 1, 16)
    n_all = d["C"].shape[0]

    def run(n):
        idx = perm[np.linspace(0, n_all - 1, n).astype(int)] - 1
        s = inputs.subset(d, idx)
        return o.integrate_batch(s["Tstart"], s["Tend"], s["C"], s["RCONST"], s["ATOL"],
                                 s["RTOL"], s["ICNTRL"], s["RCNTRL"], nthreads=threads)[4]

    n_cal = min(n_all, 16 * threads)
    t_cal = run(n_cal)
    n = int(min(n_all, max(n_cal, n_cal * target_s / max(t_cal, 1e-3))))
    log(f"cpu baseline: calibration {n_cal} cells in {t_cal:.2f} s -> sample of {n} cells")
    secs = run(n)
    if secs < 10.0 and n < n_all:  # the small calibration overestimates the rate
        n = int(min(n_all, n * target_s / max(secs, 1e-3)))
        log(f"cpu baseline: {secs:.1f} s was short -> sample of {n} cells")
        secs = run(n)
    return {"value": round(n / secs, 1), "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"{n} cells evenly spaced in the launch order of the same {n_all}-cell "
                      f"workload, OpenMP schedule(dynamic,16) on {threads} threads, "
                      f"{secs:.1f} s wall",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "s_per_chem_step": round(n_all * secs / n, 1)}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def adjoint_ratio(mech, n=65536):
    """Adjoint workload: n cells of the 4x5 grid generator, loss = weighted
    log final concentrations.  Forward = persistent solve in cost order;
    adjoint = recording forward + reverse sweep (HIP-event kernel times)."""
    import numpy as np
    import torch
    from gcchem_amd import adjoint, inputs, solver
    d = inputs.adjoint_cells(mech, n=n, seed=3)
    b = solver.Batch.from_numpy(d)
    r = solver.solve(b, kernel="persistent")
    torch.cuda.synchronize()
    b.Perm = torch.as_tensor(inputs.cost_order(r.ISTAT[:, 2].cpu().numpy()), device="cuda")
    fwd = []
    for _ in range(2):
        r = solver.solve(b, kernel="persistent")
        torch.cuda.synchronize()
        fwd.append(r.t_kernel)
    out, rec = adjoint.record_solve(b)
    torch.cuda.synchronize()
    w = torch.as_tensor(d["w"], device="cuda")
    at = torch.as_tensor(d["ATOL"], device="cuda")
    x = out.C[:, :mech.nvar]
    lam = torch.zeros_like(out.C)
    lam[:, :mech.nvar] = w * x / (x * x + at * at)
    perm = torch.as_tensor(inputs.cost_order(out.ISTAT[:, 3].cpu().numpy()), device="cuda")
    gC, gR, ierr, t_sweep = adjoint.adjoint_sweep(b, rec, lam, perm=perm)
    torch.cuda.synchronize()
    if not bool((ierr == 1).all()):
        raise RuntimeError("adjoint sweep failed for some cells")
    t_fwd = float(np.mean(fwd))
    log(f"adjoint ({n} cells): forward {t_fwd * 1e3:.1f} ms, record {rec.t_kernel * 1e3:.1f} ms, "
        f"sweep {t_sweep * 1e3:.1f} ms")
    return {"cells": n, "forward_s": round(t_fwd, 4), "record_s": round(rec.t_kernel, 4),
            "sweep_s": round(t_sweep, 4), "records": rec.used,
            "adj_over_fwd": round((rec.t_kernel + t_sweep) / t_fwd, 3)}


def host_api_rate(lib, d, perm):
    """One call of the reference entry point gc_gpu_integrate_batch on host
    arrays (PCIe staging in and out included), made as the model makes it:
    the batch gathered in the launch (cost) order beforehand, untimed, and
    Perm = 1..nCell (fullchem_mod.F90:2359-2360), so the library streams the
    inputs in behind the running solve: (cells/s, tTotal, tKernel)."""
    import ctypes
    import numpy as np
    n = d["C"].shape[0]
    order = np.asarray(perm) - 1
    arr = {k: np.ascontiguousarray(d[k][order]) for k in ("C", "RCONST", "ICNTRL", "RCNTRL")}
    arr.update({k: np.ascontiguousarray(d[k]).copy() for k in ("ATOL", "RTOL")})
    Perm = np.arange(1, n + 1, dtype=np.int32)
    ISTAT = np.zeros((n, 8), np.int32)
    RSTAT = np.zeros((n, 3))
    IERR = np.zeros(n, np.int32)
    tT, tK = ctypes.c_double(), ctypes.c_double()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib.gc_gpu_integrate_batch(n, float(d["Tstart"]), float(d["Tend"]), p(arr["C"]),
                               p(arr["RCONST"]), p(arr["ATOL"]), p(arr["RTOL"]),
                               p(arr["ICNTRL"]), p(arr["RCNTRL"]), p(Perm), p(ISTAT), p(RSTAT),
                               p(IERR), ctypes.byref(tT), ctypes.byref(tK))
    if not (IERR == 1).all():
        raise RuntimeError(f"host entry point: {int((IERR != 1).sum())} cells failed")
    return n / tT.value, tT.value, tK.value


def gate_sweep(lib, d, perm, b, sizes, reps=2):
    """The host-array entry point in launch order at several input-gate chunk
    sizes (gc_amd_set_gate; cells per chunk, 0 = the library's default, -1 =
    upload first), each `reps` times, beside the same solve with its inputs
    already in HBM: one JSON record per case.  The gate setting is restored
    afterwards."""
    from gcchem_amd import lib as L, solver
    import torch
    n = d["C"].shape[0]
    out = None
    tk = []
    for _ in range(3):
        out = solver.solve(b, kernel="persistent", out=out)
        torch.cuda.synchronize()
        tk.append(out.t_kernel)
    ref = float(min(tk))
    log(f"inputs in HBM: kernel {', '.join(f'{t * 1e3:.1f}' for t in tk)} ms")
    host_api_rate(lib, d, perm)  # pinned staging and device buffers set up, untimed
    saved = L.get_gate(lib)
    rows = []
    try:
        for cells in sizes:
            if cells < 0:
                L.set_gate(lib, 0, 0, saved["spin"])
            else:
                L.set_gate(lib, 1, cells, saved["spin"])
            for _ in range(reps):
                _, tt, tkk = host_api_rate(lib, d, perm)
                rows.append({"gate_cells": cells, "t_total_ms": round(tt * 1e3, 1),
                             "t_kernel_ms": round(tkk * 1e3, 1),
                             "kernel_over_resident_ms": round((tkk - ref) * 1e3, 1),
                             "outside_kernel_ms": round((tt - tkk) * 1e3, 1),
                             "frac_of_resident": round(ref / tt, 4)})
                log(f"gate cells {cells}: tTotal {tt * 1e3:.1f} ms, tKernel {tkk * 1e3:.1f} ms, "
                    f"frac of the in-HBM kernel {ref / tt:.4f}")
    finally:
        L.set_gate(lib, saved["mode"], saved["cells"], saved["spin"])
    return {"workload": f"fullchem 4x5 grid, {n} cells, launch order, host arrays",
            "resident_kernel_ms": round(ref * 1e3, 1), "cases": rows}


def resident_rate(d):
    """Two chemistry steps of a device-resident batch (gc_amd_resident_*):
    the first stands for the previous step (identity order, untimed), the
    second is a model step -- C and RCONST uploaded from host arrays through
    the pinned staging, the launch order computed on the device from the
    first step's step counts, C, ISTAT, RSTAT and IERR copied back.  Cold
    start on both (warm = 0), so the solve is the timed persistent one:
    (cells/s, tTotal, tKernel) of the second step."""
    from gcchem_amd import resident
    n = d["C"].shape[0]
    r = resident.Resident(d["ICNTRL"], d["RCNTRL"], d["ATOL"], d["RTOL"])
    try:
        r.step(d["Tstart"], d["Tend"], C=d["C"], RCONST=d["RCONST"], warm=False, outputs=False)
        out = r.step(d["Tstart"], d["Tend"], C=d["C"], RCONST=d["RCONST"], warm=False)
        if not (out[3] == 1).all():
            raise RuntimeError(f"resident step: {int((out[3] != 1).sum())} cells failed")
        return n / r.t_total, r.t_total, r.t_kernel
    finally:
        r.close()


def membench(block_doubles, waves_per_cu, loads_per_store, loads_in_flight=16):
    """Memory ceiling of the workspace access shape on this device
    (csrc/gc_membench.hip, built by build.py): random 512-byte wave accesses
    in per-wave blocks of the solver's size and read/write mix, 16 reads in
    flight per wave (where the shape's rate saturates), and the plain
    streaming read rate of the same buffer."""
    import subprocess
    exe = PKG / "lib" / "gc_membench"
    if not exe.exists():
        log("gc_membench is not built: no measured ceiling")
        return None
    out = subprocess.run([str(exe), str(block_doubles), str(waves_per_cu), str(loads_per_store),
                          str(loads_in_flight)],
                         check=True, capture_output=True, text=True, timeout=120).stdout
    return json.loads(out.strip().splitlines()[-1])


def copy_probe(torch, gib=1):
    """One elementwise copy kernel that reads and writes `gib` GiB: its
    FETCH_SIZE / WRITE_SIZE readings calibrate the counters of a profile."""
    n = gib * (1 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    log(f"copy probe: elementwise copy of {gib} GiB (reads {gib} GiB, writes {gib} GiB)")
    del src, dst


def bench_grid_2x25(args, mech, lib, tune):
    """2x2.5 grid (144 x 91 x 72 = 943,488 cells): the persistent kernel over
    the whole grid in HBM-sized chunks of 262,144 cells, the chunk loop
    captured once as a HIP graph (GC_GPU_WORKSPACE=global mode), against the
    plain kernel over the whole grid in one launch; cost order from a first
    solve."""
    import numpy as np
    import torch
    from gcchem_amd import inputs, solver
    t0 = time.time()
    d = inputs.grid_2x25_cells(mech, seed=2)
    n = d["C"].shape[0]
    log(f"generated {n} cells in {time.time() - t0:.1f} s")
    b = solver.Batch.from_numpy(d)
    del d
    r = solver.solve(b, kernel="persistent")
    torch.cuda.synchronize()
    log(f"previous step (grid order): kernel {r.t_kernel * 1e3:.1f} ms")
    istat = r.ISTAT.cpu().numpy()
    b.Perm = torch.as_tensor(inputs.cost_order(istat[:, 2]), device="cuda")
    chunk = 262144
    out, tk = None, []
    for k in range(args.warmup + args.steps):
        out = solver.solve(b, kernel="persistent", chunk=chunk, out=out)
        torch.cuda.synchronize()
        log(f"persistent chunked step {k}: {out.t_kernel * 1e3:.1f} ms")
        if k >= args.warmup:
            tk.append(out.t_kernel)
    ok = float((out.IERR == 1).float().mean().item())
    # the same solve as one persistent launch (its workspace is one slot per
    # resident lane whatever the cell count, so it needs no chunks)
    c_chunked = out.C.clone()
    ru = solver.solve(b, kernel="persistent", out=out)
    torch.cuda.synchronize()
    same = bool(torch.equal(ru.C, c_chunked))
    del c_chunked
    log(f"persistent, one launch: {ru.t_kernel * 1e3:.1f} ms (C identical to the chunked solve: {same})")
    t_one = ru.t_kernel
    rp = solver.solve(b, kernel="plain", out=out)
    torch.cuda.synchronize()
    log(f"plain: {rp.t_kernel * 1e3:.1f} ms")
    t = float(np.mean(tk))
    flops = mech.algorithmic_flops(istat)
    return {
        "metric": METRIC, "value": round(n / t, 1), "unit": "cells/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "fullchem Rodas3, 2x2.5 grid 144x91x72 = 943,488 cells, dt 600 s, "
                               "cost order, persistent kernel in 262,144-cell chunks replayed "
                               "as one HIP graph", "cells": n, "chunk": chunk, "tuning": tune},
        "plain_cells_per_s": round(n / rp.t_kernel, 1),
        "speedup_vs_plain": round(rp.t_kernel / t, 3),
        "persistent_one_launch_cells_per_s": round(n / t_one, 1),
        "one_launch_equals_chunked": same,
        "ierr_ok_frac": ok,
        "roofline": roofline(mech, n, t, flops, None),
        "cpu_baseline": None,
    }


STRESS_SIZES = (1000, 4000, 16000, 32000, 64000, 100000, 150000, 200000)


def bench_stress(args):
    """Synthetic stress mechanism (500 species, 1,500 reactions, ~3% LU fill)
    in eight ragged batches of 1k-200k cells, dt 600 s, each batch in its
    own launch order (no previous step to sort by).  Legs (--stress-legs):
    fwd, all batches in ONE persistent launch (gc_amd_integrate_batches_device,
    one work queue through the batches, progress lines as cells finish);
    plain, the plain kernel on each batch in turn, for the ratio and an
    equality check; adj, the discrete adjoint (dL/dC_final = 1) of all eight
    batches as ONE recording launch (record pool sized from the forward
    launch's accepted steps) and ONE persistent sweep, against the forward
    launch."""
    import numpy as np
    import torch
    from gcchem_amd import adjoint, inputs, mechanism, solver
    mech = mechanism.load("stress500")
    legs = set(args.stress_legs.split(","))
    sizes = [int(x) for x in args.stress_sizes.split(",")] if args.stress_sizes else STRESS_SIZES
    bs = [solver.Batch.from_numpy(inputs.stress_cells(mech, n, seed=40 + i,
                                                      max_steps=args.stress_max_steps))
          for i, n in enumerate(sizes)]
    ncell = sum(sizes)
    log(f"stress: {len(sizes)} batches {sizes} generated ({ncell} cells), Max_no_steps "
        f"{args.stress_max_steps or 200000}")

    def reporter(what):
        step = [0.05]

        def progress(nb, done, total):
            if done >= step[0] * total or nb == len(sizes):
                log(f"{what}: {done}/{total} cells finished, {nb}/{len(sizes)} batches")
                while step[0] * total <= done and step[0] <= 1.0:
                    step[0] += 0.05
        return progress

    outs, t_one = solver.solve_batches(bs, mech="stress500", progress=reporter("one launch"))
    torch.cuda.synchronize()
    log(f"all {len(sizes)} batches in one persistent launch: {t_one:.2f} s")
    per, total_q = [], 0.0
    for b, r, n in zip(bs, outs, sizes):
        ist = r.ISTAT.cpu().numpy()
        row = {"cells": n, "mean_steps": round(float(ist[:, 2].mean()), 1),
               "max_steps": int(ist[:, 2].max()),
               "ierr_ok_frac": float((r.IERR == 1).float().mean().item())}
        if "plain" in legs:
            rq = solver.solve(b, kernel="plain", mech="stress500")
            torch.cuda.synchronize()
            row["plain_s"] = round(rq.t_kernel, 3)
            row["one_launch_equals_plain"] = all(
                torch.equal(x, y) for x, y in ((r.C, rq.C), (r.ISTAT, rq.ISTAT),
                                               (r.RSTAT, rq.RSTAT), (r.IERR, rq.IERR)))
            total_q += rq.t_kernel
            del rq
        log(f"stress batch {n}: {row}")
        per.append(row)
    adj = None
    if "adj" in legs:
        cap = int(sum(int(r.ISTAT[:, 3].sum().item()) for r in outs))  # accepted steps
        outs2, rec = adjoint.record_solve_batches(bs, mech="stress500", cap=cap,
                                                  progress=reporter("recording launch"))
        torch.cuda.synchronize()
        same = all(torch.equal(a.C, b.C) and torch.equal(a.ISTAT, b.ISTAT)
                   for a, b in zip(outs, outs2))
        log(f"adjoint: recording launch {rec.t_kernel:.2f} s ({rec.used} records, every "
            f"{rec.every}), results == forward launch: {same}")
        # the sweep runs each batch's costly cells first (descending accepted steps)
        perms = [inputs.cost_order(o.ISTAT[:, 3].cpu().numpy()) for o in outs2]
        del outs2
        lams = [torch.ones_like(b.C) for b in bs]
        grads, t_sweep = adjoint.adjoint_sweep_batches(bs, rec, lams, mech="stress500",
                                                       progress=reporter("sweep"), perms=perms)
        torch.cuda.synchronize()
        ratio = (rec.t_kernel + t_sweep) / t_one
        ok = float(np.mean([float((g[2] == 1).float().mean()) for g in grads]))
        log(f"adjoint: sweep {t_sweep:.2f} s, adj/fwd {ratio:.3f}, IERR=1 fraction {ok:.4f}")
        adj = {"batches": list(sizes), "forward_s": round(t_one, 3),
               "record_s": round(rec.t_kernel, 3), "sweep_s": round(t_sweep, 3),
               "records": rec.used, "every": rec.every, "record_equals_forward": same,
               "adj_over_fwd": round(ratio, 3), "ierr_ok_frac": ok, "one_launch_each": True}
    line = {
        "metric": METRIC, "value": round(ncell / t_one, 1), "unit": "cells/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": round(t_one * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"synthetic stress mechanism 500 species x 1500 reactions, "
                               f"{len(sizes)} ragged batches {sizes} in one persistent launch, "
                               f"dt 600 s, Max_no_steps {args.stress_max_steps or 200000}",
                   "cells": ncell, "one_launch": True},
        "batches": per, "adjoint": adj,
        "adj_over_fwd": adj["adj_over_fwd"] if adj else None,
        "roofline": None, "cpu_baseline": None,
    }
    if "plain" in legs:
        line["plain_cells_per_s"] = round(ncell / total_q, 1)
        line["speedup_vs_plain"] = round(total_q / t_one, 3)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cells", type=int, default=0, help="use the first N cells (debug)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement")
    ap.add_argument("--grid", default="4x5", choices=["4x5", "2x2.5", "stress"],
                    help="4x5 (238,464 cells; the headline line), 2x2.5 (943,488 cells, "
                         "chunked graph) or stress (synthetic 500x1500 mechanism, ragged batches)")
    ap.add_argument("--stress-sizes", default="", help="comma-separated batch sizes (stress)")
    ap.add_argument("--stress-legs", default="fwd,plain,adj",
                    help="stress legs: fwd (always), plain, adj")
    ap.add_argument("--stress-max-steps", type=int, default=1000,
                    help="ICNTRL(4) of the stress cells (0: the integrator's 200,000)")
    ap.add_argument("--gate-sweep", default="",
                    help="comma-separated input-gate chunk sizes in cells (0 = default, "
                         "-1 = upload first): time the host-array entry point at each and "
                         "print one record instead of the bench line")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB elementwise copy first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the batched Rodas3 solve has no cross-GPU "
              "exchange and this build runs on one MI355X; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    load_package()
    from gcchem_amd import inputs, lib as L, mechanism, solver, tuning

    torch.cuda.set_device(0)
    if args.copy_probe:
        copy_probe(torch)
    mech = mechanism.load("fullchem")
    lib = L.load("fullchem")
    tune = tuning.apply(lib, "fullchem")
    if args.grid == "2x2.5":
        print(json.dumps(bench_grid_2x25(args, mech, lib, tune)), flush=True)
        return
    if args.grid == "stress":
        print(json.dumps(bench_stress(args)), flush=True)
        return
    t0 = time.time()
    d = inputs.grid_4x5_cells(mech, seed=1)
    if args.cells:
        d = inputs.subset(d, np.arange(args.cells))
    n = d["C"].shape[0]
    log(f"generated {n} cells in {time.time() - t0:.1f} s; tuning {tune}")
    b = solver.Batch.from_numpy(d)

    # previous chemistry step -> cost order (untimed)
    r = solver.solve(b, kernel="persistent")
    torch.cuda.synchronize()
    log(f"previous step (grid order): kernel {r.t_kernel * 1e3:.1f} ms")
    perm = inputs.cost_order(r.ISTAT[:, 2].cpu().numpy())
    b.Perm = torch.as_tensor(perm, device=b.C.device)
    out = None
    if args.gate_sweep:
        sizes = [int(x) for x in args.gate_sweep.split(",") if x.strip()]
        print(json.dumps(gate_sweep(lib, d, perm, b, sizes)), flush=True)
        return

    def timed(kernel, steps, warmup):
        nonlocal out
        for _ in range(warmup):
            out = solver.solve(b, kernel=kernel, out=out)
        torch.cuda.synchronize()
        tk = []
        t0 = time.perf_counter()
        for k in range(steps):
            out = solver.solve(b, kernel=kernel, out=out)
            tk.append(out.t_kernel)
            log(f"{kernel} step {k}: kernel {out.t_kernel * 1e3:.1f} ms")
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, float(np.mean(tk))

    s_pers, tk_pers = timed("persistent", args.steps, args.warmup)
    istat = out.ISTAT.cpu().numpy()
    ierr = out.IERR.cpu().numpy()
    log(f"persistent: {s_pers * 1e3:.1f} ms/step, kernel {tk_pers * 1e3:.1f} ms")
    s_plain, tk_plain = timed("plain", args.steps, 0)
    log(f"plain: {s_plain * 1e3:.1f} ms/step, kernel {tk_plain * 1e3:.1f} ms")
    # the plain kernel in grid order (the order before any cost sort)
    b.Perm = torch.arange(1, n + 1, dtype=torch.int32, device=b.C.device)
    out = solver.solve(b, kernel="plain", out=out)
    torch.cuda.synchronize()
    tk_plain_grid = out.t_kernel
    log(f"plain, grid order: kernel {tk_plain_grid * 1e3:.1f} ms")
    b.Perm = torch.as_tensor(perm, device=b.C.device)

    host_rate, host_tt, host_tk = host_api_rate(lib, d, perm)
    res_rate, res_tt, res_tk = resident_rate(d)
    log(f"resident step: tTotal {res_tt * 1e3:.1f} ms, tKernel {res_tk * 1e3:.1f} ms")
    adj = adjoint_ratio(mech) if not args.no_adjoint else None
    log(f"host-array entry point: tTotal {host_tt * 1e3:.1f} ms, tKernel {host_tk * 1e3:.1f} ms")

    flops = mech.algorithmic_flops(istat)
    prof = tuning.profile_traffic("rodas3_persistent_4x5")
    roof = roofline(mech, n, tk_pers, flops, prof.get("bytes_per_launch") if prof else None)
    if prof and prof.get("bytes_per_launch"):
        gbps = prof["bytes_per_launch"] / tk_pers / 1e9
        # the ceiling this access shape reaches on this device, measured now
        wpc = 4 * {1: 2, 2: 4, 3: 3}[tune.get("variant", 1)] if not tune.get("blocks_per_cu") \
            else tune["blocks_per_cu"] * tune["block"] // 64
        lps = max(1, round(prof["fetch_bytes"] / max(prof["write_bytes"], 1)))
        mb = membench(L.load("fullchem").gc_amd_workspace_doubles_per_slot(), wpc, lps)
        if mb:
            ceil = mb["gather_tbs"] * 1e3
            roof["measured_traffic_rate"].update({"pattern_ceiling": round(ceil, 1),
                                "frac_of_pattern_ceiling": round(gbps / ceil, 4),
                                "pattern": f"random 512-byte wave accesses in per-wave blocks "
                                           f"of {mb['block_doubles']} doubles/lane, "
                                           f"{mb['waves_per_cu']} waves/CU, "
                                           f"{mb['loads_in_flight']} reads in flight per wave, "
                                           f"1 write per {mb['loads_per_store']} reads "
                                           f"(gc_membench)",
                                "stream_read": round(mb["stream_read_tbs"] * 1e3, 1),
                                "frac_of_stream_read": round(gbps / (mb["stream_read_tbs"] * 1e3),
                                                             4)})
            log(f"membench: {mb}")

    res = {
        "metric": METRIC,
        "value": round(n / s_pers, 1),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_pers * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "fullchem Rodas3, 4x5 grid 72x46x72 = 238,464 cells, dt 600 s, "
                               "cost-ordered launch (previous-step Nsteps)",
                   "cells": n, "dt_s": 600.0, "kernel": "persistent", "tuning": tune},
        "plain_cells_per_s": round(n / s_plain, 1),
        "speedup_vs_plain": round(s_plain / s_pers, 3),
        "plain_grid_order_cells_per_s": round(n / tk_plain_grid, 1),
        "speedup_vs_plain_grid_order": round(tk_plain_grid / tk_pers, 3),
        "s_per_chem_step": round(s_pers, 4),
        "host_api_cells_per_s": round(host_rate, 1),
        "host_api_t_total_s": round(host_tt, 4),
        "host_api_frac_of_value": round(host_rate / (n / s_pers), 4),
        "resident_step": {"cells_per_s": round(res_rate, 1), "t_total_s": round(res_tt, 4),
                          "t_kernel_s": round(res_tk, 4),
                          "frac_of_value": round(res_rate / (n / s_pers), 4)},
        "adj_over_fwd": adj["adj_over_fwd"] if adj else None,
        "adjoint": adj,
        "ierr_ok_frac": float((ierr == 1).mean()),
        "mean_steps_per_cell": float(istat[:, 2].mean()),
        "roofline": roof,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(mech, d, perm)
        log(f"cpu baseline: {res['cpu_baseline']}")
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
