"""Host-side logic that needs no GPU: the benchmark's one-GPU contract, the
seeded input generators, the cost order, the autotuner's feasibility rule,
the operation counts and the binding's argument checks."""
import importlib.util
import subprocess
import sys

import numpy as np
import pytest
import torch

from conftest import PKG, REPO
from gcchem_amd import inputs, mechanism, solver


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "one MI355X" in r.stderr


def test_cost_order_is_a_stable_descending_sort():
    key = np.array([3, 7, 3, 1, 7, 0])
    perm = inputs.cost_order(key)
    assert perm.dtype == np.int32 and sorted(perm) == list(range(1, 7))
    assert list(perm) == [2, 5, 1, 3, 4, 6]          # ties keep their order
    assert list(inputs.cost_order(key, descending=False)) == [6, 4, 1, 3, 2, 5]


def test_generators_are_seeded_and_shaped(mech):
    a = inputs.random_cells(mech, n=16, seed=3)
    b = inputs.random_cells(mech, n=16, seed=3)
    c = inputs.random_cells(mech, n=16, seed=4)
    for k in ("C", "RCONST", "ICNTRL", "RCNTRL"):
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["C"], c["C"])
    assert a["C"].shape == (16, mech.nspec) and a["RCONST"].shape == (16, mech.nreact)
    assert a["ICNTRL"].shape == (16, 20) and a["ICNTRL"].dtype == np.int32
    assert list(a["ICNTRL"][0, [0, 1, 2, 14]]) == [1, 0, 4, -1]
    # literal-rate reactions (A(126), A(735)) never read RCONST: it is 0 there
    assert np.all(a["RCONST"][:, [125, 734]] == 0.0)
    assert np.all(a["C"] > 0.0) and np.all(a["RCONST"] >= 0.0)


def test_grid_and_adjoint_workloads(mech):
    d = inputs.grid_cells(mech, 4, 3, 5, seed=1)
    assert d["C"].shape[0] == 4 * 3 * 5
    d3 = inputs.adjoint_cells(mech, n=100, seed=3)
    assert d3["C"].shape == (100, mech.nspec) and d3["w"].shape == (mech.nvar,)
    s = inputs.subset(d3, np.array([5, 2]))
    assert np.array_equal(s["C"][1], d3["C"][2]) and list(s["Perm"]) == [1, 2]


def test_stress_workload_caps_steps():
    m = mechanism.load("stress500")
    d = inputs.stress_cells(m, 8, seed=4, max_steps=777)
    assert np.all(d["ICNTRL"][:, 3] == 777)
    assert d["C"].shape == (8, m.nspec)


def test_autotune_feasibility():
    spec = importlib.util.spec_from_file_location("autotune", PKG / "autotune.py")
    at = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(at)
    base = {"block": 256, "cells_per_wave": 64, "grain": 64}
    assert at.feasible({**base, "variant": 1, "blocks_per_cu": 2})       # 8 waves, 135 KB LDS
    assert not at.feasible({**base, "variant": 1, "blocks_per_cu": 3})   # 12 waves > target
    assert at.feasible({**base, "variant": 2, "blocks_per_cu": 4})       # 16 waves, 139 KB
    assert at.feasible({**base, "variant": 1, "blocks_per_cu": 0})       # library sizes it


def test_operation_counts_reconcile_with_the_reference_text(mech):
    """The generator's own counts of the reference expressions against counts
    taken independently from the text of the reference's generated Fortran
    (Fun_SPLIT 11,139, Jac_SP 10,607, KppDecomp 44,353 = 2,773 divides +
    2 x 20,790, KppSolve 11,013 flops)."""
    assert abs(mech.flops_fun() - 11139) / 11139 < 0.005
    assert abs(mech.flops_jac() - 10607) / 10607 < 0.005
    assert mech.flops_decomp() == 44353
    assert mech.flops_solve() == 11013
    istat = np.array([[10, 5, 6, 5, 1, 6, 24, 0]])
    f = mech.algorithmic_flops(istat)
    assert f == (mech.flops_fun() * 10 + mech.flops_jac() * 5
                 + (mech.flops_decomp() + mech.flops_prepare()) * 6
                 + mech.flops_solve() * 24 + mech.flops_step_vector() * 6)


def test_solver_rejects_host_tensors_and_bad_shapes(mech):
    d = inputs.random_cells(mech, n=4, seed=0)
    b = solver.Batch.from_numpy(d, device="cpu")
    with pytest.raises(TypeError):
        solver._check(b, (mech.nvar, mech.nfix, mech.nreact, mech.nnz))
    b.C = torch.zeros((4, mech.nspec + 1), dtype=torch.float64)
    with pytest.raises((TypeError, ValueError)):
        solver._check(b, (mech.nvar, mech.nfix, mech.nreact, mech.nnz))


def test_gate_settings_round_trip_and_reject_bad_values():
    """gc_amd_set_gate / gc_amd_get_gate (the overlapped upload of the
    host-array entry points): values round-trip, out-of-range values are
    refused and change nothing.  No device is touched."""
    from gcchem_amd import lib as L
    lib = L.load()
    before = L.get_gate(lib)
    try:
        L.set_gate(lib, mode=0, cells=128, spin=1000)
        assert L.get_gate(lib) == {"mode": 0, "cells": 128, "spin": 1000}
        for bad in ((3, 0, 1), (-1, 0, 1), (1, -64, 1), (1, 0, 0)):
            assert lib.gc_amd_set_gate(*bad) == -1
        assert L.get_gate(lib) == {"mode": 0, "cells": 128, "spin": 1000}
    finally:
        L.set_gate(lib, **before)


def test_gate_starts_off_under_serializing_tools():
    """Under rocprofv3 counter collection or a launch-serializing debug mode
    the overlapped upload starts off (mode 0): the copies could not land while
    the gated kernel runs.  A process without them starts with it on."""
    import os
    import subprocess
    import sys
    from conftest import REPO
    code = ("import sys; sys.path.insert(0, %r); import __graft_entry__ as g; g.load_package(); "
            "from gcchem_amd import lib as L; print(L.get_gate(L.load())['mode'])" % str(REPO))
    env = {k: v for k, v in os.environ.items()
           if k not in ("ROCPROF_COUNTER_COLLECTION", "HIP_LAUNCH_BLOCKING",
                        "AMD_SERIALIZE_KERNEL", "AMD_SERIALIZE_COPY")}
    run = lambda e: subprocess.run([sys.executable, "-c", code], env=e, capture_output=True,
                                   text=True, timeout=300, check=True).stdout.split()[-1]
    assert run(env) == "1"
    assert run({**env, "ROCPROF_COUNTER_COLLECTION": "True"}) == "0"
    assert run({**env, "AMD_SERIALIZE_KERNEL": "3"}) == "0"
    assert run({**env, "HIP_LAUNCH_BLOCKING": "0"}) == "1"


def test_compiled_in_launch_is_the_stored_tuning():
    """build.py compiles each mechanism's autotuned launch (tuning/<mech>.json)
    in as the library's default, so a C-ABI caller that never sets a tuning
    gets it; a fresh process (no GC_AMD_TUNE) reads it back unchanged."""
    import json
    import os
    import subprocess
    import sys
    from conftest import REPO
    tdir = REPO / "geos-chem-gpu-hybrid_amd" / "tuning"
    mechs = sorted(p.stem for p in tdir.glob("*.json"))
    assert "fullchem" in mechs and "stress500" in mechs
    code = ("import sys, json; sys.path.insert(0, %r); import __graft_entry__ as g; "
            "g.load_package(); from gcchem_amd import lib as L; "
            "print(json.dumps({m: L.get_tuning(L.load(m)) for m in %r}))" % (str(REPO), mechs))
    env = {k: v for k, v in os.environ.items() if k != "GC_AMD_TUNE"}
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         timeout=300, check=True).stdout
    got = json.loads([line for line in out.splitlines() if line.strip()][-1])
    for m in mechs:
        want = json.loads((tdir / f"{m}.json").read_text())["launch"]
        assert got[m] == {k: int(v) for k, v in want.items()}, m
