"""Build the gfx950 solver libraries (one per mechanism).

For each mechanism IR under ``mechanisms/``:

1. ``codegen/gen_tables.py`` writes ``_build/<mech>/mech_tables.inc`` (the
   operand tables of Fun, Jac_SP + KppDecomp, KppSolve, KppSolveTR and the
   per-wave workspace layout);
2. ``csrc/gc_rodas3.hip`` is compiled once per kernel variant
   (``-DGC_VARIANT=1``: 2 waves per SIMD, 32-entry LU row segments in LDS;
   ``-DGC_VARIANT=2``: 4 waves per SIMD, 16-entry segments) and
   ``csrc/gc_api.hip`` once, and the host-only ``csrc/gc_hybrid.cpp`` (cost
   order and hybrid CPU/GPU split) with g++, in parallel;
3. the objects are linked into ``lib/libgcchemgpu_<mech>.so``; the fullchem
   library is also installed as ``lib/libgcchemgpu.so``, the drop-in name the
   model links (src/GCClassic/CMakeLists.txt:204-228 links ``GC_GPU_LIB``).

Every HIP file is compiled with ``-ffp-contract=off`` (no FMA contraction:
the reference CPU build evaluates every product and sum separately) and with
packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import time
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "lib"
MECH_DIR = PKG / "mechanisms"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
VARIANTS = (1, 2, 3)  # kernel variants (GC_VARIANT)


def lib_path(mech: str = "fullchem") -> Path:
    return LIB / f"libgcchemgpu_{mech}.so"


def _stale(out: Path, deps) -> bool:
    if not out.exists():
        return True
    t = out.stat().st_mtime
    return any(Path(d).stat().st_mtime > t for d in deps)


def _run_parallel(jobs, max_jobs):
    """jobs: list of (argv, label). Raises on the first failure."""
    running = []
    pending = list(jobs)
    while pending or running:
        while pending and len(running) < max_jobs:
            argv, label = pending.pop(0)
            p = subprocess.Popen(argv, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            running.append((p, label, time.time()))
        for item in list(running):
            p, label, t0 = item
            if p.poll() is not None:
                out = p.stdout.read()
                running.remove(item)
                if p.returncode != 0:
                    for q, _, _ in running:
                        q.kill()
                    raise RuntimeError(f"{label} failed:\n{out[-4000:]}")
                print(f"  built {label} ({time.time() - t0:.0f} s)", flush=True)
        time.sleep(0.2)


def build_mechanism(mech: str, force: bool = False, jobs: int | None = None) -> Path:
    todo, link = plan_mechanism(mech, force)
    if todo:
        _run_parallel(todo, jobs or max(1, min(len(todo), os.cpu_count() or 1)))
    return link()


TUNE_MACROS = {"block": "GC_TUNE_BLOCK", "blocks_per_cu": "GC_TUNE_BLOCKS_PER_CU",
               "cells_per_wave": "GC_TUNE_CELLS_PER_WAVE", "grain": "GC_TUNE_GRAIN",
               "variant": "GC_TUNE_VARIANT"}


def tuning_defines(mech: str) -> list[str]:
    """The autotuned launch stored in tuning/<mech>.json as the library's
    compiled-in default (-DGC_TUNE_*); none without a stored launch."""
    p = PKG / "tuning" / f"{mech}.json"
    if not p.exists():
        return []
    launch = json.loads(p.read_text()).get("launch", {})
    return [f"-D{TUNE_MACROS[k]}={int(v)}" for k, v in sorted(launch.items()) if k in TUNE_MACROS]


def plan_mechanism(mech: str, force: bool = False):
    """Generate the mechanism header; return (compile jobs, link function)."""
    sys.path.insert(0, str(PKG / "codegen"))
    import gen_tables  # noqa: E402

    ir_path = MECH_DIR / f"{mech}.json"
    bdir = BUILD / mech
    bdir.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(exist_ok=True)
    inc = bdir / "mech_tables.inc"
    gen_deps = [ir_path, PKG / "codegen" / "gen_tables.py"]
    if force or _stale(inc, gen_deps):
        ir = json.loads(ir_path.read_text())
        text = gen_tables.generate(ir)
        if not inc.exists() or inc.read_text() != text:
            inc.write_text(text)
    common = [CSRC / "gc_math.h", CSRC / "gc_rodas3_kernels.h"]
    todo, objs = [], []
    for v in VARIANTS:
        obj = bdir / f"rodas3_v{v}.o"
        objs.append(obj)
        deps = [CSRC / "gc_rodas3.hip", inc] + common
        if force or _stale(obj, deps):
            todo.append(([HIPCC, *HIP_FLAGS, f"-DGC_VARIANT={v}", f"-I{CSRC}", f"-I{bdir}",
                          "-c", str(CSRC / "gc_rodas3.hip"), "-o", str(obj)],
                         f"{mech}/rodas3_v{v}.o"))
    obj = bdir / "api.o"
    objs.append(obj)
    deps = [CSRC / "gc_api.hip", PKG.parent / "include" / "gcchemgpu.h"] + common
    tune_file = PKG / "tuning" / f"{mech}.json"
    if tune_file.exists():
        deps.append(tune_file)
    if force or _stale(obj, deps):
        todo.append(([HIPCC, *HIP_FLAGS, *tuning_defines(mech), f"-I{CSRC}", "-c",
                      str(CSRC / "gc_api.hip"), "-o", str(obj)], f"{mech}/api.o"))

    obj = bdir / "hybrid.o"
    objs.append(obj)
    deps = [CSRC / "gc_hybrid.cpp", PKG.parent / "include" / "gcchemgpu.h"]
    if force or _stale(obj, deps):
        todo.append((["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall",
                      f"-I{PKG.parent / 'include'}", "-c", str(CSRC / "gc_hybrid.cpp"),
                      "-o", str(obj)], f"{mech}/hybrid.o"))

    def link_one(out: Path):
        if force or _stale(out, objs):
            tmp = out.with_suffix(".so.tmp")
            subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread",
                            f"-Wl,-soname,{out.name}", "-o", str(tmp), *map(str, objs)],
                           check=True)
            os.replace(tmp, out)

    def link() -> Path:
        out = lib_path(mech)
        link_one(out)
        if mech == "fullchem":
            # the drop-in name the model links (-lgcchemgpu), with its own soname
            link_one(LIB / "libgcchemgpu.so")
        return out

    return todo, link


def available_mechanisms():
    return sorted(p.stem for p in MECH_DIR.glob("*.json"))


MEMBENCH = LIB / "gc_membench"


def plan_membench(force: bool = False):
    """The memory-ceiling calibration program (csrc/gc_membench.hip)."""
    src = CSRC / "gc_membench.hip"
    LIB.mkdir(exist_ok=True)
    if force or _stale(MEMBENCH, [src]):
        return [([HIPCC, *HIP_FLAGS, str(src), "-o", str(MEMBENCH)], "gc_membench")]
    return []


LINKTEST = LIB / "gc_link_test"


def build_linktest(force: bool = False) -> Path:
    """The link / OpenMP-coexistence check (csrc/gc_link_test.c): plain C with
    -fopenmp, linked against the drop-in libgcchemgpu.so found next to it."""
    src = CSRC / "gc_link_test.c"
    drop_in = LIB / "libgcchemgpu.so"
    if force or _stale(LINKTEST, [src, drop_in]):
        subprocess.run(["gcc", "-O2", "-fopenmp", f"-I{PKG.parent / 'include'}", str(src),
                        f"-L{LIB}", "-lgcchemgpu", "-Wl,-rpath,$ORIGIN", "-o", str(LINKTEST)],
                       check=True)
    return LINKTEST


FORTRAN_CALLER = LIB / "gc_fortran_caller"
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")


def build_fortran_caller(force: bool = False) -> Path | None:
    """The Fortran caller check (csrc/gc_fortran_caller.F90): flang -fopenmp,
    binding the drop-in library through ISO_C_BINDING as GEOS-Chem does.
    Returns None when no flang is installed."""
    src = CSRC / "gc_fortran_caller.F90"
    drop_in = LIB / "libgcchemgpu.so"
    if not Path(FLANG).exists():
        return None
    if force or _stale(FORTRAN_CALLER, [src, drop_in]):
        subprocess.run([FLANG, "-O2", "-fopenmp", f"-J{BUILD}", str(src), f"-L{LIB}",
                        "-lgcchemgpu", "-Wl,-rpath,$ORIGIN", "-o", str(FORTRAN_CALLER)],
                       check=True, cwd=str(BUILD))
    return FORTRAN_CALLER


def build_all(force: bool = False, mechs=None) -> list[Path]:
    """Build every mechanism library, the calibration program and the link
    check; all compile jobs share one pool."""
    plans = [plan_mechanism(m, force) for m in (mechs or available_mechanisms())]
    todo = [job for jobs, _ in plans for job in jobs] + plan_membench(force)
    if todo:
        _run_parallel(todo, max(1, min(len(todo), os.cpu_count() or 1)))
    libs = [link() for _, link in plans]
    if (LIB / "libgcchemgpu.so").exists():
        build_linktest(force)
        build_fortran_caller(force)
    return libs


if __name__ == "__main__":
    force = "--force" in sys.argv
    mechs = [a for a in sys.argv[1:] if not a.startswith("--")] or None
    for p in build_all(force=force, mechs=mechs):
        print(p)
