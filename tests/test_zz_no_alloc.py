"""GPU test of gc_amd_integrate_batch_noalloc, the allocation-free entry
point (its scratch is a static device array of the library).

Called through the C ABI on device tensors, it must give what
gc_amd_integrate_batch_device gives on the same inputs bit for bit, for both
kernels, on two calls in a row, the second one smaller (the scratch keeps the
first call's contents), without changing the device's free memory; the plain
kernel's cell limit is refused with an error and leaves the outputs alone.
"""
import ctypes

import numpy as np
import pytest
import torch

from gcchem_amd import inputs, lib as L, mechanism, solver

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _outputs(n, C):
    return (C.clone(), torch.full((n, 8), -7, dtype=torch.int32, device="cuda"),
            torch.full((n, 3), -7.0, dtype=torch.float64, device="cuda"),
            torch.full((n,), -7, dtype=torch.int32, device="cuda"))


def _call(lib, fn, kernel, b, outs, extra=()):
    C, IS, RS, IE = outs
    tk = ctypes.c_double(0.0)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(lib, fn)(kernel, *extra, b.n, float(b.Tstart), float(b.Tend), _p(C),
                          _p(b.RCONST), _p(b.ATOL), _p(b.RTOL), _p(b.ICNTRL), _p(b.RCNTRL),
                          _p(b.Perm), _p(IS), _p(RS), _p(IE), stream, ctypes.byref(tk))
    torch.cuda.synchronize()
    return rc


@pytest.mark.parametrize("kernel", [L.KERNEL_PERSISTENT, L.KERNEL_PLAIN])
def test_noalloc_equals_device_entry_point(kernel):
    lib = L.load("fullchem")
    mech = mechanism.load("fullchem")
    d = inputs.random_cells(mech, n=2048, seed=31)
    d["Perm"] = (np.random.default_rng(32).permutation(2048) + 1).astype(np.int32)
    first = solver.Batch.from_numpy(d)
    second = solver.Batch.from_numpy(inputs.subset(d, np.arange(700, 1300)))
    second.Perm = torch.flip(second.Perm, [0]).contiguous()
    refs = []
    for b in (first, second):
        ref = _outputs(b.n, b.C)
        print(f"case: device entry point, kernel {kernel}, {b.n} cells", flush=True)
        assert _call(lib, "gc_amd_integrate_batch_device", kernel, b, ref, (0,)) == 0
        refs.append(ref)
    gots = [_outputs(b.n, b.C) for b in (first, second)]   # the test's own allocations
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for b, got in zip((first, second), gots):
        print(f"case: allocation-free entry point, kernel {kernel}, {b.n} cells", flush=True)
        rc = _call(lib, "gc_amd_integrate_batch_noalloc", kernel, b, got)
        assert rc == 0, L.last_error(lib)
    assert torch.cuda.mem_get_info()[0] == free0      # before any comparison allocates
    for got, ref in zip(gots, refs):
        for g, r in zip(got, ref):
            assert torch.equal(g, r)


def test_noalloc_plain_limit_refused():
    lib = L.load("fullchem")
    mech = mechanism.load("fullchem")
    W = lib.gc_amd_workspace_doubles_per_slot()
    limit = ((256 << 20) - 4096) // (512 * W) * 64
    assert limit == 2560
    d = inputs.random_cells(mech, n=limit + 1, seed=33)
    b = solver.Batch.from_numpy(d)
    got = _outputs(b.n, b.C)
    print(f"case: allocation-free plain call over its limit, {b.n} cells", flush=True)
    assert _call(lib, "gc_amd_integrate_batch_noalloc", L.KERNEL_PLAIN, b, got) == -1
    assert "nCell" in L.last_error(lib)
    assert bool((got[3] == -7).all()) and torch.equal(got[0], b.C)
