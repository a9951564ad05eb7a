"""GPU checks of the recording forward solve and the Rodas3 discrete adjoint.

* Recording changes nothing in the forward results, keeps exactly one record
  per accepted step, and the recorded step sizes replayed by the oracle from
  the initial state reproduce the GPU's final concentrations bit for bit.
* The GPU adjoint equals the NumPy reference adjoint (tests/adjoint_ref.py,
  stoichiometry from the reference's CalcStoichNum) to summation-order
  rounding, and central finite differences of the oracle replay (relative
  step 1e-6) on 256 (cell, parameter) pairs to 1e-6 relative above the
  rounding noise of the loss.
* The torch.autograd.Function returns the same gradients.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from adjoint_ref import Monomials, adjoint_cell, log_loss, stoich_matrix
from gcchem_amd import adjoint, inputs, solver

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def adj_case(mech):
    d = inputs.adjoint_cells(mech, n=1024, seed=3)
    b = solver.Batch.from_numpy(d)
    _say("record solve: adjoint-workload generator, 1024 cells")
    out, rec = adjoint.record_solve(b)
    torch.cuda.synchronize()
    return d, b, out, rec


def test_record_solve_matches_forward(mech, oracle_cr, adj_case):
    d, b, out, rec = adj_case
    ref = solver.solve(b, kernel="persistent")
    torch.cuda.synchronize()
    for k in ("C", "ISTAT", "RSTAT", "IERR"):
        assert torch.equal(getattr(out, k), getattr(ref, k)), k
    nacc = out.ISTAT[:, 3].cpu().numpy()
    assert rec.used == int(nacc.sum())
    rng = np.random.default_rng(1)
    for c in rng.choice(b.n, 16, replace=False):
        Hs = adjoint.records_of_cell(rec, int(c))
        assert Hs.size == nacc[c]
        Y = oracle_cr.replay(d["C"][c][:mech.nvar], d["C"][c][mech.nvar:], d["RCONST"][c], Hs)
        g = out.C[c, :mech.nvar].cpu().numpy()
        assert np.array_equal(Y.view(np.int64), g.view(np.int64)), int(c)


def test_record_pool_regrows(mech):
    d = inputs.random_cells(mech, n=300, seed=8)
    b = solver.Batch.from_numpy(d)
    _say("record solve with a 10-record pool, 300 cells")
    out, rec = adjoint.record_solve(b, cap=10)
    ref = solver.solve(b, kernel="persistent")
    torch.cuda.synchronize()
    assert torch.equal(out.C, ref.C)
    assert rec.used == int(out.ISTAT[:, 3].sum().item())
    assert rec.H.shape[0] >= rec.used and rec.Y.shape[0] >= rec.yused


@pytest.fixture(scope="module")
def adj_grads(mech, adj_case):
    d, b, out, rec = adj_case
    C = out.C.cpu().numpy()
    _, lamV = log_loss(C[:, :mech.nvar], d["w"][None, :], d["ATOL"][None, :])
    lam = np.zeros_like(C)
    lam[:, :mech.nvar] = lamV
    _say("adjoint sweep: 1024 cells")
    gC, gR, ierr, tk = adjoint.adjoint_sweep(b, rec, torch.as_tensor(lam, device="cuda"))
    torch.cuda.synchronize()
    assert bool((ierr == 1).all())
    return lam, gC.cpu().numpy(), gR.cpu().numpy()


def test_adjoint_matches_reference(mech, oracle_cr, checkforms, adj_case, adj_grads):
    d, b, out, rec = adj_case
    lam, gC, gR = adj_grads
    mono = Monomials(mech.ir)
    S = stoich_matrix(checkforms, mech.nvar, mech.nreact)
    nv = mech.nvar
    for c in (0, 100, 517, 1023):
        Hs = adjoint.records_of_cell(rec, c)
        rV, rF, rK = adjoint_cell(oracle_cr, mono, S, d["C"][c][:nv], d["C"][c][nv:],
                                  d["RCONST"][c], Hs, lam[c][:nv], lam[c][nv:])
        for got, ref in ((gC[c][:nv], rV), (gC[c][nv:], rF), (gR[c], rK)):
            scale = np.max(np.abs(ref)) + 1e-300
            assert np.max(np.abs(got - ref)) <= 1e-8 * scale, c
            rel = np.abs(got - ref) / (np.abs(ref) + 1e-300 * scale)
            assert np.mean(rel <= 1e-9) >= 0.95 or ref.size < 10, c


def test_adjoint_finite_differences_256_pairs(mech, oracle_libm, adj_case, adj_grads):
    """Central differences of the frozen-step replay, relative step 1e-6, on
    256 (cell, parameter) pairs over initial VAR concentrations, fixed species
    and RCONST: |g - fd| <= 1e-6 |fd| + the rounding noise of the loss."""
    d, b, out, rec = adj_case
    lam, gC, gR = adj_grads
    nv = mech.nvar
    rng = np.random.default_rng(256)
    cells = rng.choice(b.n, 64, replace=False)
    pairs = []
    for c in cells:
        V, F, R = d["C"][c][:nv], d["C"][c][nv:], d["RCONST"][c]
        g = {"V": gC[c][:nv], "F": gC[c][nv:], "RCT": gR[c]}
        vals = {"V": V, "F": F, "RCT": R}
        for which, n in (("V", 2), ("RCT", 1), ("F", 1)):
            ok = np.nonzero(vals[which] != 0.0)[0]
            sens = np.abs(g[which][ok] * vals[which][ok])
            j = ok[np.argmax(sens)] if rng.uniform() < 0.5 else rng.choice(ok)
            pairs.append((int(c), which, int(j)))
            if n == 2:
                pairs.append((int(c), which, int(rng.choice(ok))))
    pairs = pairs[:256]
    assert len(pairs) == 256
    Hs = {c: adjoint.records_of_cell(rec, c) for c in set(p[0] for p in pairs)}

    def one(p):
        c, which, j = p
        args = {"V": d["C"][c][:nv].copy(), "F": d["C"][c][nv:].copy(), "RCT": d["RCONST"][c].copy()}
        x = args[which][j]
        h = 1e-6 * abs(x)

        def L(v):
            a = {k: v_.copy() for k, v_ in args.items()}
            a[which][j] = v
            Y = oracle_libm.replay(a["V"], a["F"], a["RCT"], Hs[c])
            return log_loss(Y, d["w"], d["ATOL"])[0]
        fd = (L(x + h) - L(x - h)) / (2 * h)
        L0 = L(x)
        # rounding noise of the loss after ~100 stiff steps: 256 ulps of its
        # term magnitude (FD at larger steps converges to the adjoint where this
        # noise dominates, i.e. for gradients far below the cell's largest)
        noise = 2.0 ** -44 * (abs(L0) + np.sum(np.abs(d["w"])) * 40.0) / (2 * h)
        g = {"V": gC[c][:nv], "F": gC[c][nv:], "RCT": gR[c]}[which][j]
        return abs(g - fd) <= 1e-6 * abs(fd) + noise, (c, which, j, g, fd, noise)

    _say("finite differences: 256 pairs on the oracle replay")
    with ThreadPoolExecutor(8) as ex:
        res = list(ex.map(one, pairs))
    bad = [info for ok, info in res if not ok]
    assert not bad, bad[:5]


@pytest.mark.parametrize("every", [3, 8])
def test_sparse_checkpoints_identical(mech, adj_case, adj_grads, every):
    """Checkpointing every `every`-th step and replaying the segments gives
    bit-identical gradients (the replay is the recorded trajectory)."""
    d, b, out, rec = adj_case
    lam, gC, gR = adj_grads
    _say(f"record + adjoint with a checkpoint every {every} steps, 1024 cells")
    out2, rec2 = adjoint.record_solve(b, every=every)
    assert torch.equal(out2.C, out.C)
    nacc = out.ISTAT[:, 3].cpu().numpy()
    assert rec2.yused == int(np.sum((nacc + every - 1) // every))
    g2, r2, ierr, _ = adjoint.adjoint_sweep(b, rec2, torch.as_tensor(lam, device="cuda"))
    torch.cuda.synchronize()
    assert bool((ierr == 1).all())
    assert np.array_equal(g2.cpu().numpy().view(np.int64), gC.view(np.int64))
    assert np.array_equal(r2.cpu().numpy().view(np.int64), gR.view(np.int64))


def test_kept_stage_vectors_identical(mech, adj_case, adj_grads):
    """The recording keeps the checkpoint steps' stage vectors when every step
    is a checkpoint; recomputing them instead gives the same gradients bit
    for bit."""
    d, b, out, rec = adj_case
    lam, gC, gR = adj_grads
    assert rec.every == 1 and rec.K is not None
    _say("record + adjoint, every step a checkpoint, stage vectors recomputed, 1024 cells")
    out2, rec2 = adjoint.record_solve(b, every=1, stage_budget=0)
    assert rec2.K is None and torch.equal(out2.C, out.C)
    g2, r2, ierr, _ = adjoint.adjoint_sweep(b, rec2, torch.as_tensor(lam, device="cuda"))
    torch.cuda.synchronize()
    assert bool((ierr == 1).all())
    assert np.array_equal(g2.cpu().numpy().view(np.int64), gC.view(np.int64))
    assert np.array_equal(r2.cpu().numpy().view(np.int64), gR.view(np.int64))


def test_autograd_function(mech, adj_case, adj_grads):
    d, b, out, rec = adj_case
    lam, gC, gR = adj_grads
    C = b.C.clone().requires_grad_(True)
    R = b.RCONST.clone().requires_grad_(True)
    _say("autograd: Rodas3Solve forward + backward, 1024 cells")
    Cf = adjoint.solve(C, R, b)
    assert torch.equal(Cf.detach(), out.C)
    w = torch.as_tensor(d["w"], device="cuda")
    at = torch.as_tensor(d["ATOL"], device="cuda")
    x = Cf[:, :mech.nvar]
    loss = (w * 0.5 * torch.log(x * x + at * at)).sum()
    loss.backward()
    torch.cuda.synchronize()
    g = C.grad.cpu().numpy()
    scale = np.abs(gC).max()
    assert np.max(np.abs(g - gC)) <= 1e-12 * scale
    assert np.max(np.abs(R.grad.cpu().numpy() - gR)) <= 1e-12 * np.abs(gR).max()


def test_adjoint_non_autonomous_flagged(mech):
    d = inputs.random_cells(mech, n=70, seed=2)
    d["ICNTRL"][:, 0] = 0
    d["ICNTRL"][:5, 0] = 1
    b = solver.Batch.from_numpy(d)
    _say("adjoint of a partly non-autonomous batch, 70 cells")
    out, rec = adjoint.record_solve(b)
    lam = torch.ones_like(b.C)
    gC, gR, ierr, _ = adjoint.adjoint_sweep(b, rec, lam)
    torch.cuda.synchronize()
    ie = ierr.cpu().numpy()
    assert (ie[:5] == 1).all() and (ie[5:] == -3).all()
    assert float(gC[5:].abs().sum()) == 0.0


def test_stress_mechanism_adjoint(oracle_cr):
    """The adjoint leg on the stress mechanism (500 species, 1,500 reactions):
    256 cells over a 20 s step recorded and swept on the GPU; for 4 sampled
    cells the gradients equal the NumPy reference adjoint, whose
    stoichiometry is derived here from the mechanism's reaction list
    (products - reactants), not from the generator; central finite
    differences of the oracle replay (relative step 1e-6) agree on 48
    (cell, parameter) pairs."""
    from gcchem_amd import mechanism, synthetic
    from oracle import Oracle
    m = mechanism.load("stress500")
    o = Oracle(m.ir, pow_mode="cr")
    ol = Oracle(m.ir, pow_mode="libm")
    nv = m.nvar
    d = inputs.stress_cells(m, n=256, seed=44)
    d["Tend"] = 20.0
    w = np.random.default_rng(45).standard_normal(nv)
    b = solver.Batch.from_numpy(d)
    _say("record solve: stress mechanism, 256 cells, 20 s")
    out, rec = adjoint.record_solve(b, mech="stress500")
    torch.cuda.synchronize()
    IE = out.IERR.cpu().numpy()
    assert np.all(IE == 1)
    C = out.C.cpu().numpy()
    _, lamV = log_loss(C[:, :nv], w[None, :], d["ATOL"][None, :])
    lam = np.zeros_like(C)
    lam[:, :nv] = lamV
    _say("adjoint sweep: stress mechanism, 256 cells")
    gC, gR, ierr, _ = adjoint.adjoint_sweep(b, rec, torch.as_tensor(lam, device="cuda"),
                                            mech="stress500")
    torch.cuda.synchronize()
    assert bool((ierr == 1).all())
    gC, gR = gC.cpu().numpy(), gR.cpu().numpy()
    rx, _ = synthetic.stress_reactions()
    S = synthetic.stoichiometry(rx, nv, m.ir["species_order"])
    mono = Monomials(m.ir)
    for c in (0, 77, 150, 255):
        Hs = adjoint.records_of_cell(rec, c)
        Y = o.replay(d["C"][c][:nv], d["C"][c][nv:], d["RCONST"][c], Hs)
        assert np.array_equal(Y.view(np.int64), C[c, :nv].view(np.int64)), c
        rV, rF, rK = adjoint_cell(o, mono, S, d["C"][c][:nv], d["C"][c][nv:], d["RCONST"][c], Hs,
                                  lam[c][:nv], lam[c][nv:])
        _say(f"stress adjoint: cell {c} ({Hs.size} steps) against the NumPy reference")
        for got, ref in ((gC[c][:nv], rV), (gC[c][nv:], rF), (gR[c], rK)):
            scale = np.max(np.abs(ref)) + 1e-300
            assert np.max(np.abs(got - ref)) <= 1e-8 * scale, c
    rng = np.random.default_rng(48)
    pairs = []
    for c in rng.choice(256, 16, replace=False):
        for which in ("V", "V", "RCT"):
            vals = d["C"][c][:nv] if which == "V" else d["RCONST"][c]
            g = gC[c][:nv] if which == "V" else gR[c]
            sens = np.abs(g * vals)
            j = int(np.argmax(sens)) if rng.uniform() < 0.5 else int(rng.integers(vals.size))
            pairs.append((int(c), which, j))
    Hs = {c: adjoint.records_of_cell(rec, c) for c in set(p[0] for p in pairs)}

    def one(p):
        c, which, j = p
        V, F, R = d["C"][c][:nv].copy(), d["C"][c][nv:].copy(), d["RCONST"][c].copy()
        x = (V if which == "V" else R)[j]
        h = 1e-6 * abs(x)

        def L(v):
            V2, R2 = V.copy(), R.copy()
            (V2 if which == "V" else R2)[j] = v
            return log_loss(ol.replay(V2, F, R2, Hs[c]), w, d["ATOL"])[0]
        fd = (L(x + h) - L(x - h)) / (2 * h)
        noise = 2.0 ** -44 * (abs(L(x)) + np.sum(np.abs(w)) * 40.0) / (2 * h)
        g = (gC[c][:nv] if which == "V" else gR[c])[j]
        return abs(g - fd) <= 1e-6 * abs(fd) + noise, (c, which, j, g, fd, noise)

    _say("finite differences: 48 pairs on the stress mechanism's oracle replay")
    with ThreadPoolExecutor(8) as ex:
        res = list(ex.map(one, pairs))
    bad = [info for ok, info in res if not ok]
    assert not bad, bad[:5]
