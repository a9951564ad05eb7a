"""GPU: the device-resident batch and the pinned staging of host arrays.

* Two resident chemistry steps equal two calls of the reference host-array
  entry point in which the caller carries the state itself, as
  fullchem_mod.F90 does: the second call gets the first call's C, RCNTRL(3) =
  the first call's RSTAT(3) (KPPHvalue, :1775, 2044) and the launch order of
  the stable descending sort by the first call's ISTAT(3) (:905-981, the
  library's gc_amd_cost_slots); the order the resident step used on the
  device equals that one.  A third step with C changed on the host (as
  transport would) equals the host path too.
* A batch whose arrays span several staging chunks: the host-array entry
  point equals the device entry point bit for bit.
"""
import ctypes

import numpy as np
import pytest
import torch

from gcchem_amd import inputs, lib as L, resident, solver

pytestmark = pytest.mark.gpu


def _host_call(lib, d, C, RCNTRL, Perm):
    n = C.shape[0]
    C = np.ascontiguousarray(C.copy())
    IS, RS, IE = np.zeros((n, 8), np.int32), np.zeros((n, 3)), np.zeros(n, np.int32)
    tt, tk = ctypes.c_double(), ctypes.c_double()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib.gc_gpu_integrate_batch(n, d["Tstart"], d["Tend"], p(C), p(d["RCONST"]), p(d["ATOL"]),
                               p(d["RTOL"]), p(d["ICNTRL"]), p(np.ascontiguousarray(RCNTRL)),
                               p(np.ascontiguousarray(Perm, np.int32)), p(IS), p(RS), p(IE),
                               ctypes.byref(tt), ctypes.byref(tk))
    assert IE.min() > -90, L.last_error(lib)
    return C, IS, RS, IE


def _cost_order(lib, nsteps):
    n = nsteps.size
    slot, cell = np.zeros(n, np.int32), np.zeros(n, np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert lib.gc_amd_cost_slots(n, p(np.ascontiguousarray(nsteps, np.int32)), 1, p(slot),
                                 p(cell)) == 0
    return cell


def test_two_resident_steps_equal_host_path_steps(mech):
    lib = L.load("fullchem")
    d = inputs.random_cells(mech, n=3000, seed=61)
    n = d["C"].shape[0]
    r = resident.Resident(d["ICNTRL"], d["RCNTRL"], d["ATOL"], d["RTOL"])
    try:
        print("case: resident step 1, 3000 cells", flush=True)
        s1 = r.step(d["Tstart"], d["Tend"], C=d["C"], RCONST=d["RCONST"])
        print("case: resident step 2 (C kept on the device, warm start, device cost order)",
              flush=True)
        s2 = r.step(d["Tstart"], d["Tend"])
        C3 = s2[0] * 1.01
        print("case: resident step 3 (C from the host)", flush=True)
        s3 = r.step(d["Tstart"], d["Tend"], C=C3)
    finally:
        r.close()
    h1 = _host_call(lib, d, d["C"], d["RCNTRL"], np.arange(1, n + 1))
    rc2 = d["RCNTRL"].copy()
    rc2[:, 2] = h1[2][:, 2]
    perm2 = _cost_order(lib, h1[1][:, 2])
    assert np.array_equal(perm2, inputs.cost_order(h1[1][:, 2]))
    h2 = _host_call(lib, d, h1[0], rc2, perm2)
    rc3 = d["RCNTRL"].copy()
    rc3[:, 2] = h2[2][:, 2]
    h3 = _host_call(lib, d, C3, rc3, _cost_order(lib, h2[1][:, 2]))
    assert np.array_equal(s1[4], np.arange(1, n + 1))
    assert np.array_equal(s2[4], perm2)
    for s, h in ((s1, h1), (s2, h2), (s3, h3)):
        for a, b in zip(s[:4], h):
            assert np.array_equal(a, b)
    assert np.mean(h2[1][:, 2]) < np.mean(h1[1][:, 2])   # the warm start saves steps


def test_staging_across_chunks_equals_device_path(mech):
    """9,000 cells: RCONST is 76 MB, three 32 MB pinned staging chunks."""
    lib = L.load("fullchem")
    d = inputs.random_cells(mech, n=9000, seed=62)
    n = d["C"].shape[0]
    print("case: host-array call, 9000 cells", flush=True)
    h = _host_call(lib, d, d["C"], d["RCNTRL"], d["Perm"])
    b = solver.Batch.from_numpy(d)
    rr = solver.solve(b, kernel="persistent")
    torch.cuda.synchronize()
    for a, t in zip(h, (rr.C, rr.ISTAT, rr.RSTAT, rr.IERR)):
        assert np.array_equal(a, t.cpu().numpy())
