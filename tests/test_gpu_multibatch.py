"""GPU: several independent batches in one persistent launch
(gc_amd_*_batches_device), forward and adjoint.

Ragged batches of the stress mechanism (500 species) with their own launch
orders and their own time spans, and of fullchem, each equal to the oracle
bit for bit, with the launch's finished-cell progress reported up to every
cell; the recording launch over two batches followed by the sweep over
both in one launch gives each batch the gradients of its own single-batch
record and sweep, bit for bit; a sweep whose persistent grid has fewer
lanes than cells (each wave takes several 64-cell groups, one workspace
slot each) gives the gradients of the sweep with one group per wave."""
import numpy as np
import pytest
import torch

from gcchem_amd import adjoint, inputs, mechanism, solver
from gcchem_amd import lib as L
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.int64) if np.asarray(a).dtype == np.float64
                          else np.asarray(a),
                          np.asarray(b).view(np.int64) if np.asarray(b).dtype == np.float64
                          else np.asarray(b))


@pytest.mark.parametrize("name", ["stress500", "fullchem"])
def test_batches_in_one_launch_match_oracle(name):
    m = mechanism.load(name)
    o = Oracle(m.ir, pow_mode="cr")
    rng = np.random.default_rng(71)
    if name == "stress500":
        sizes, spans = (30, 100, 257), ((0.0, 20.0), (0.0, 10.0), (100.0, 115.0))
        mk = lambda n, s: inputs.stress_cells(m, n=n, seed=s)
    else:
        sizes, spans = (500, 77, 1), ((0.0, 600.0), (0.0, 300.0), (0.0, 600.0))
        mk = lambda n, s: inputs.random_cells(m, n=n, seed=s)
    ds = []
    for k, (n, (t0, t1)) in enumerate(zip(sizes, spans)):
        d = mk(n, 72 + k)
        d["Tstart"], d["Tend"] = t0, t1
        d["Perm"] = (rng.permutation(n) + 1).astype(np.int32)
        ds.append(d)
    bs = [solver.Batch.from_numpy(d) for d in ds]
    print(f"case: {name}, batches of {sizes} in one launch", flush=True)
    seen = []
    outs, tk = solver.solve_batches(bs, mech=name, progress=lambda *a: seen.append(a))
    torch.cuda.synchronize()
    total = sum(sizes)
    assert seen and seen[-1] == (len(sizes), total, total)
    assert all(a[1] <= b[1] and a[0] <= b[0] for a, b in zip(seen, seen[1:]))
    for d, r in zip(ds, outs):
        ref = o.integrate_batch(d["Tstart"], d["Tend"], d["C"], d["RCONST"], d["ATOL"], d["RTOL"],
                                d["ICNTRL"], d["RCNTRL"], d["Perm"])
        assert np.array_equal(r.IERR.cpu().numpy(), ref[3])
        assert np.array_equal(r.ISTAT.cpu().numpy(), ref[1])
        assert np.array_equal(r.RSTAT.cpu().numpy(), ref[2])
        assert np.array_equal(r.C.cpu().numpy()[:, :m.nvar], ref[0][:, :m.nvar])


def test_adjoint_of_batches_in_one_launch(mech):
    d1 = inputs.adjoint_cells(mech, n=200, seed=81)
    d2 = inputs.adjoint_cells(mech, n=57, seed=82)
    d2["Tend"] = 300.0
    bs = [solver.Batch.from_numpy(d) for d in (d1, d2)]
    print("case: recording launch over two fullchem batches (200 + 57 cells)", flush=True)
    seen = []
    outs, rec = adjoint.record_solve_batches(bs, progress=lambda *a: seen.append(a))
    torch.cuda.synchronize()
    assert rec.every == 1 and seen[-1] == (2, 257, 257)
    lams = [torch.as_tensor(np.random.default_rng(83 + k).standard_normal(tuple(b.C.shape)),
                            device="cuda") for k, b in enumerate(bs)]
    print("case: sweep over both batches in one launch", flush=True)
    seen = []
    grads, _ = adjoint.adjoint_sweep_batches(bs, rec, lams, progress=lambda *a: seen.append(a))
    torch.cuda.synchronize()
    assert seen[-1] == (2, 257, 257)
    assert all(a[1] <= b[1] and a[0] <= b[0] for a, b in zip(seen, seen[1:]))
    for b, out, lam, (gC, gR, ie) in zip(bs, outs, lams, grads):
        print(f"case: single-batch record + sweep, {b.n} cells", flush=True)
        o1, r1 = adjoint.record_solve(b, every=1, stage_budget=0)
        assert torch.equal(o1.C, out.C) and torch.equal(o1.ISTAT, out.ISTAT)
        g1, k1, e1, _ = adjoint.adjoint_sweep(b, r1, lam)
        torch.cuda.synchronize()
        assert bool((ie == 1).all()) and torch.equal(ie, e1)
        assert _same(gC.cpu().numpy(), g1.cpu().numpy())
        assert _same(gR.cpu().numpy(), k1.cpu().numpy())


def test_sweep_with_fewer_lanes_than_cells():
    m = mechanism.load("fullchem")
    lib = L.load()
    n = 40_000
    d = inputs.adjoint_cells(m, n=n, seed=91)
    d["Tend"] = 120.0
    b = solver.Batch.from_numpy(d)
    print(f"case: recording {n} fullchem cells", flush=True)
    out, rec = adjoint.record_solve(b, every=2)
    assert bool((out.IERR == 1).all())
    lam = torch.as_tensor(np.random.default_rng(92).standard_normal(tuple(b.C.shape)),
                          device="cuda")
    print("case: sweep, one 64-cell group per wave", flush=True)
    g1, k1, e1, _ = adjoint.adjoint_sweep(b, rec, lam)
    torch.cuda.synchronize()
    old = L.get_tuning(lib)
    try:
        # 1 wave per CU: CUs * 64 lanes (16,384 on 256 CUs) for 40,000 cells,
        # so every wave takes two or more 64-cell groups
        L.set_tuning(lib, block=64, blocks_per_cu=1)
        print("case: sweep, 1 wave per CU", flush=True)
        g2, k2, e2, _ = adjoint.adjoint_sweep(b, rec, lam)
        torch.cuda.synchronize()
    finally:
        L.set_tuning(lib, **old)
    assert bool((e1 == 1).all()) and torch.equal(e1, e2)
    assert _same(g1.cpu().numpy(), g2.cpu().numpy())
    assert _same(k1.cpu().numpy(), k2.cpu().numpy())
