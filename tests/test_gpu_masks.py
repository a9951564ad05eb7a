"""GPU: the masked kernels (gc_amd_integrate_batch_masked) against the
oracle's guarded statements, bit for bit, for both kernels.

Masks per cell: a random set of species switched off the way auto-reduce
switches off inactive species (DO_FUN and DO_SLV, and DO_JVS over their
Jacobian rows and columns), plus a few cells with independent random DO_JVS
flags; one cell with every flag on.  Switched-off species come back unchanged.
All flags on through the masked kernels equals the unmasked solve."""
import numpy as np
import pytest
import torch

from gcchem_amd import inputs, mechanism, solver
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _masks(ir, n, seed, frac):
    rng = np.random.default_rng(seed)
    nvar = ir["nvar"]
    rows = np.repeat(np.arange(nvar), np.diff(ir["crow"]))
    cols = np.asarray(ir["icol"])
    frozen = rng.uniform(size=(n, nvar)) < frac
    frozen[0] = False                                   # every flag on
    fun = ~frozen
    jvs = ~(frozen[:, rows] | frozen[:, cols])
    jvs[1:6] &= rng.uniform(size=(5, cols.size)) < 0.95  # independent DO_JVS flags
    return frozen, fun, jvs, fun.copy()


def _gpu(d, kernel, mech, words):
    b = solver.Batch.from_numpy(d)
    mk = tuple(None if w is None else torch.as_tensor(w.view(np.int32), device="cuda")
               for w in words)
    r = solver.solve(b, kernel=kernel, mech=mech, masks=mk)
    torch.cuda.synchronize()
    return r.C.cpu().numpy(), r.ISTAT.cpu().numpy(), r.RSTAT.cpu().numpy(), r.IERR.cpu().numpy()


@pytest.mark.parametrize("name,n", [("fullchem", 512), ("stress500", 128)])
def test_masked_kernels_bitexact(name, n):
    m = mechanism.load(name)
    o = Oracle(m.ir, pow_mode="cr")
    if name == "fullchem":
        d = inputs.random_cells(m, n=n, seed=51)
    else:
        d = inputs.stress_cells(m, n=n, seed=52)
        d["Tend"] = 20.0
    frozen, fun, jvs, slv = _masks(m.ir, n, 53, 0.25)
    d["ICNTRL"][1:6, 3] = 1500   # cells with independent DO_JVS flags: a cap on their steps
    words = tuple(Oracle.mask_words(x) for x in (fun, jvs, slv))
    ref = o.integrate_batch(d["Tstart"], d["Tend"], d["C"], d["RCONST"], d["ATOL"], d["RTOL"],
                            d["ICNTRL"], d["RCNTRL"], d["Perm"], masks=words)
    assert np.sum(ref[3] == 1) >= n - 10 and ref[3][0] == 1
    ok = ref[3] == 1
    assert np.array_equal(np.where(frozen[ok], ref[0][ok, :m.nvar], 0.0),
                          np.where(frozen[ok], d["C"][ok, :m.nvar], 0.0))
    for kernel in ("plain", "persistent"):
        print(f"case: masked {name}, {n} cells, {kernel}", flush=True)
        C, IS, RS, IE = _gpu(d, kernel, name, words)
        assert np.array_equal(IE, ref[3]) and np.array_equal(IS, ref[1])
        assert np.array_equal(RS, ref[2])
        assert np.array_equal(C[:, :m.nvar], ref[0][:, :m.nvar])


def test_masked_kernels_all_on_equal_unmasked(mech):
    d = inputs.random_cells(mech, n=256, seed=54)
    ones = tuple(np.full((256, solver.mask_words(k)), 0xFFFFFFFF, np.uint32)
                 for k in (mech.nvar, mech.nnz, mech.nvar))
    b = solver.Batch.from_numpy(d)
    for kernel in ("plain", "persistent"):
        print(f"case: masked kernels, every flag on, {kernel}", flush=True)
        ref = solver.solve(b, kernel=kernel)
        got = _gpu(d, kernel, "fullchem", ones)
        torch.cuda.synchronize()
        for g, r in zip(got, (ref.C, ref.ISTAT, ref.RSTAT, ref.IERR)):
            assert np.array_equal(g, r.cpu().numpy())
