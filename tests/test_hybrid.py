"""Cost order and hybrid CPU/GPU split (gc_amd_cost_slots, gc_amd_hybrid_plan,
gc_amd_hybrid_update, gc_amd_hybrid_step) against the CPU restatement of
fullchem_mod.F90's batch mode 3 (oracle/hybrid.py).

The restatement is pinned without retyping it: the slots equal numpy's stable
argsort, the head and the device slices equal brute-force prefix searches,
and hand-computed cases fix the work model, the fraction and the moving
averages.  All of these run on the CPU (no device call: the steps here give
every slot to the CPU); the concurrent device launch is a GPU test.
"""
import ctypes
import time

import numpy as np
import pytest

import hybrid as H
from gcchem_amd import lib as L

p = lambda a: a.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture(scope="module")
def lib():
    return L.load("fullchem")


def lib_slots(lib, prev, sort=1):
    prev = np.ascontiguousarray(prev, dtype=np.int32)
    n = prev.size
    slot, cell = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
    rc = lib.gc_amd_cost_slots(n, p(prev), sort, p(slot), p(cell))
    return rc, slot[:n], cell[:n]


def lib_plan(lib, nsteps, frac, rate_gpu=-1.0, rate_cpu=-1.0, ndev=1, dev_rate=None,
             fixed=False):
    nsteps = np.ascontiguousarray(nsteps, dtype=np.int32)
    n = nsteps.size
    work = np.zeros(n + 1, np.int64)
    lo, hi = np.zeros(max(ndev, 1), np.int32), np.zeros(max(ndev, 1), np.int32)
    f = ctypes.c_double(frac)
    ng, nd = ctypes.c_int(), ctypes.c_int()
    dr = None if dev_rate is None else np.ascontiguousarray(dev_rate, dtype=np.float64)
    rc = lib.gc_amd_hybrid_plan(n, p(nsteps), int(fixed), ctypes.byref(f), rate_gpu, rate_cpu,
                                ndev, None if dr is None else p(dr), p(work), ctypes.byref(ng),
                                ctypes.byref(nd), p(lo), p(hi))
    return rc, work, f.value, ng.value, nd.value, lo, hi


@pytest.mark.parametrize("n,kmax,seed", [(0, 5, 0), (1, 5, 1), (37, 0, 2), (500, 3, 3),
                                         (4000, 300, 4)])
def test_cost_slots_are_a_stable_descending_sort(lib, n, kmax, seed):
    prev = np.random.default_rng(seed).integers(0, kmax + 1, n).astype(np.int32)
    rc, slot, cell = lib_slots(lib, prev)
    assert rc == 0
    order = np.argsort(-prev.astype(np.int64), kind="stable")   # most expensive first
    assert np.array_equal(cell - 1, order)
    assert np.array_equal(slot[cell - 1], np.arange(1, n + 1))
    o_slot, o_cell = H.cost_slots(prev.tolist())
    assert np.array_equal(slot, o_slot) and np.array_equal(cell, o_cell)
    # GC_BATCH_SORT=0: stock order
    rc, slot, cell = lib_slots(lib, prev, sort=0)
    assert rc == 0 and np.array_equal(cell, np.arange(1, n + 1))
    assert H.cost_slots(prev.tolist(), sort=False)[1] == list(range(1, n + 1))


def test_cost_slots_reject_negative_counts(lib):
    assert lib_slots(lib, [3, -1, 2])[0] == -1


def test_plan_known_answers(lib):
    # work prefix [0,5,8,11,12]: half of 12 is 6, only slot 1 fits
    rc, work, frac, ng, nd, lo, hi = lib_plan(lib, [5, 3, 3, 1], 0.5)
    assert rc == 0 and list(work) == [0, 5, 8, 11, 12] and frac == 0.5
    assert (ng, nd, lo[0], hi[0]) == (1, 1, 1, 1)
    # measured rates 3:1 -> fraction 0.75, target 9: slots 1-2
    rc, work, frac, ng, nd, lo, hi = lib_plan(lib, [5, 3, 3, 1], 0.5, 3.0, 1.0)
    assert frac == 0.75 and ng == 2
    # the fraction is clipped to [0.01, 0.99] when it comes from the rates
    assert lib_plan(lib, [5, 3, 3, 1], 0.5, 1000.0, 1.0)[2] == 0.99
    assert lib_plan(lib, [5, 3, 3, 1], 0.5, 1.0, 1000.0)[2] == 0.01
    # ... and kept as given when fixed (clipped to [0, 1] only)
    assert lib_plan(lib, [5, 3, 3, 1], 1.7, 3.0, 1.0, fixed=True)[2:4] == (1.0, 4)
    # a zero step count still costs one unit
    assert list(lib_plan(lib, [0, 0, 0], 0.5)[1]) == [0, 1, 2, 3]
    # two devices, equal weights: work [0,4,8,10,12,14,16], first target 8
    rc, work, frac, ng, nd, lo, hi = lib_plan(lib, [4, 4, 2, 2, 2, 2], 1.0, ndev=2, fixed=True)
    assert (ng, nd) == (6, 2) and list(lo) == [1, 3] and list(hi) == [2, 6]
    # measured device rates 1:3 move the boundary to a quarter of the work
    rc, work, frac, ng, nd, lo, hi = lib_plan(lib, [4, 4, 2, 2, 2, 2], 1.0, ndev=2,
                                              dev_rate=[1.0, 3.0], fixed=True)
    assert list(lo) == [1, 2] and list(hi) == [1, 6]
    # fewer head slots than devices: only as many devices as slots
    rc, work, frac, ng, nd, lo, hi = lib_plan(lib, [4, 4, 2, 2, 2, 2], 0.5, ndev=4)
    assert (ng, nd) == (2, 2) and list(lo[:2]) == [1, 2] and list(hi[:2]) == [1, 2]
    assert list(lo[2:]) == [1, 1] and list(hi[2:]) == [0, 0]
    # nothing for the GPU
    rc, work, frac, ng, nd, lo, hi = lib_plan(lib, [4, 4], 0.0, ndev=3, fixed=True)
    assert (ng, nd, lo[0], hi[0]) == (0, 1, 1, 0)


def test_plan_rejects_bad_arguments(lib):
    assert lib_plan(lib, [1, 2], 0.5, ndev=0)[0] == -1
    assert lib_plan(lib, [1, 2], 0.5, ndev=L.MAX_DEV + 1)[0] == -1


@pytest.mark.parametrize("seed", range(6))
def test_plan_matches_oracle_and_brute_force(lib, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 3000))
    nsteps = rng.integers(0, 250, n).astype(np.int32)
    ndev = int(rng.integers(1, L.MAX_DEV + 1))
    frac = float(rng.uniform(0, 1))
    rg, rc_ = (float(rng.uniform(1, 1e5)), float(rng.uniform(1, 1e5))) if seed % 2 else (-1, -1)
    dev_rate = rng.uniform(0.5, 2.0, ndev) if seed % 3 else None
    rc, work, f, ng, nd, lo, hi = lib_plan(lib, nsteps, frac, rg, rc_, ndev, dev_rate)
    assert rc == 0
    o = H.hybrid_plan(nsteps.tolist(), frac, rg, rc_, ndev,
                      None if dev_rate is None else dev_rate.tolist())
    assert list(work) == o[0] and f == o[1] and (ng, nd) == (o[2], o[3])
    assert list(lo) == o[4] and list(hi) == o[5]
    # brute force: the head is the longest prefix within the GPU's share
    w = np.concatenate([[0], np.cumsum(np.maximum(nsteps.astype(np.int64), 1))])
    assert ng == max(k for k in range(n + 1) if w[k] <= f * w[-1])
    # the device slices tile 1..nGpu in order
    covered = [s for d in range(nd) for s in range(lo[d], hi[d] + 1)]
    assert covered == list(range(1, ng + 1))


def test_update_known_answers():
    lib = L.load("fullchem")
    work = np.array([0, 5, 8, 11, 12], np.int64)
    lo, hi = np.array([1], np.int32), np.array([2], np.int32)
    tw = np.array([2.0])
    dr = np.array([-1.0])
    rg, rc = ctypes.c_double(-1.0), ctypes.c_double(-1.0)
    lib.gc_amd_hybrid_update(4, 2, p(work), 1, p(lo), p(hi), p(tw), 2.0, 4.0,
                             ctypes.byref(rg), ctypes.byref(rc), p(dr))
    assert (rg.value, rc.value, dr[0]) == (4.0, 1.0, 4.0)    # 8/2, (12-8)/4, 8/2
    lib.gc_amd_hybrid_update(4, 2, p(work), 1, p(lo), p(hi), p(np.array([1.0])), 1.0, 1.0,
                             ctypes.byref(rg), ctypes.byref(rc), p(dr))
    assert (rg.value, rc.value, dr[0]) == (6.0, 2.5, 6.0)    # half old, half new
    assert H.hybrid_update(4, 2, list(work), 1, [1], [2], [2.0], 2.0, 4.0,
                           -1.0, -1.0, [-1.0]) == (4.0, 1.0, [4.0])
    # no CPU slots / no time: that rate is left as it was
    lib.gc_amd_hybrid_update(4, 4, p(work), 1, p(lo), p(hi), p(tw), 0.0, 3.0,
                             ctypes.byref(rg), ctypes.byref(rc), p(dr))
    assert (rg.value, rc.value) == (6.0, 2.5)


def test_hybrid_step_all_on_cpu(lib):
    """A step whose split gives the GPU nothing (fixed fraction 0) runs only
    the caller's CPU solve, on the calling thread, over every slot, and
    measures the CPU rate from it."""
    n = 300
    nsteps = np.full(n, 7, np.int32)
    st = L.GcAmdHybrid.initial(work_frac=0.0, fixed_frac=True, ndev=2)
    calls = []

    @L.CPU_TAIL_FN
    def tail(first, last, user):
        calls.append((first, last))
        time.sleep(0.05)

    rc = lib.gc_amd_hybrid_step(n, 0.0, 600.0, None, None, None, None, None, None, p(nsteps),
                                None, None, None, ctypes.byref(st), tail, None)
    assert rc == 0 and calls == [(1, n)]
    assert st.n_gpu == 0 and st.n_dev_use == 1 and st.t_gpu == 0.0
    assert st.rate_gpu == -1.0 and st.t_cpu >= 0.05
    assert st.rate_cpu == pytest.approx(7 * n / st.t_cpu)
    # a CPU share with no CPU solve to run it is refused
    st2 = L.GcAmdHybrid.initial(work_frac=0.0, fixed_frac=True)
    assert lib.gc_amd_hybrid_step(n, 0.0, 600.0, None, None, None, None, None, None,
                                  p(nsteps), None, None, None, ctypes.byref(st2),
                                  L.CPU_TAIL_FN(), None) == -1


def test_python_cost_order_equals_native_slots(lib):
    """The benchmark's launch order (inputs.cost_order) is the library's cost
    slot order: the same stable descending sort."""
    from gcchem_amd import inputs
    prev = np.random.default_rng(9).integers(0, 120, 5000).astype(np.int32)
    rc, slot, cell = lib_slots(lib, prev)
    assert rc == 0 and np.array_equal(inputs.cost_order(prev), cell)


def test_split_converges_to_the_rate_ratio(lib):
    """Plan -> timed solve -> update, repeated: with a GPU that processes
    work units 3x as fast as the CPU, the fraction of work given to the GPU
    settles at 3/4 and both sides then finish together (the reference's runs
    converged to 0.73-0.79 on its hardware, docs/HYBRID_RESULTS.md:156-162).
    Timings are the exact work/rate of each side."""
    rng = np.random.default_rng(11)
    n = 20000
    nsteps = np.sort(rng.integers(1, 300, n))[::-1].astype(np.int32)   # cost order
    r_gpu, r_cpu = 3.0e6, 1.0e6                                         # work units / s
    frac, rg, rc = ctypes.c_double(0.5), ctypes.c_double(-1.0), ctypes.c_double(-1.0)
    dr = np.full(1, -1.0)
    work = np.zeros(n + 1, np.int64)
    lo, hi = np.zeros(1, np.int32), np.zeros(1, np.int32)
    ng, nd = ctypes.c_int(), ctypes.c_int()
    fracs = []
    for _ in range(6):
        assert lib.gc_amd_hybrid_plan(n, p(nsteps), 0, ctypes.byref(frac), rg.value, rc.value,
                                      1, p(dr), p(work), ctypes.byref(ng), ctypes.byref(nd),
                                      p(lo), p(hi)) == 0
        t_gpu = work[ng.value] / r_gpu
        t_cpu = (work[n] - work[ng.value]) / r_cpu
        lib.gc_amd_hybrid_update(n, ng.value, p(work), nd.value, p(lo), p(hi),
                                 p(np.array([t_gpu])), t_gpu, t_cpu, ctypes.byref(rg),
                                 ctypes.byref(rc), p(dr))
        fracs.append(frac.value)
    assert fracs[0] == 0.5 and fracs[1] == pytest.approx(0.75, rel=1e-12)
    assert rg.value == pytest.approx(r_gpu) and rc.value == pytest.approx(r_cpu)
    assert fracs[-1] == pytest.approx(0.75, rel=1e-12)
    t_gpu, t_cpu = work[ng.value] / r_gpu, (work[n] - work[ng.value]) / r_cpu
    assert abs(t_gpu - t_cpu) <= 300 / r_cpu + 300 / r_gpu     # within one cell's work


@pytest.mark.parametrize("sanitize", ["address,undefined", "thread"])
def test_multi_device_step_under_sanitizers(tmp_path, sanitize):
    """gc_amd_hybrid_step's multi-device path, which a one-device box cannot
    run: csrc/gc_hybrid.cpp is linked with a stand-in device solve
    (tests/hybrid_threads_check.cpp) and built with g++ and the host
    sanitizers.  Every head slot is solved once by the device whose slice
    holds it, at the right offsets of every array; the devices run at the
    same time; a device that cannot be selected returns its slice as IERR -99
    and is counted; empty slices are skipped; the tail goes to the callback."""
    import shutil
    import subprocess
    from conftest import REPO
    if not shutil.which("g++"):
        pytest.skip("no g++")
    exe = tmp_path / "hybrid_threads_check"
    flags = [f"-fsanitize={sanitize}", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", *flags, "-pthread",
                    f"-I{REPO / 'include'}", str(REPO / "tests" / "hybrid_threads_check.cpp"),
                    str(REPO / "geos-chem-gpu-hybrid_amd" / "csrc" / "gc_hybrid.cpp"),
                    "-o", str(exe)], check=True)
    cases = [(3000, 4, -1), (3000, 4, 2), (3000, 1, -1), (5, 8, -1), (1, 3, -1), (0, 2, -1),
             (7, 16, 1)] if sanitize != "thread" else [(3000, 4, -1), (3000, 16, 5)]
    for n, ndev, bad in cases:
        r = subprocess.run([str(exe), str(n), str(ndev), str(bad)], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, (n, ndev, bad, r.stdout, r.stderr[-3000:])
        assert r.stdout.strip().splitlines()[-1] == "OK"
