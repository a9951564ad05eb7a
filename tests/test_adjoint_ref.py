"""CPU checks of the frozen-step replay and of the NumPy discrete adjoint.

* The oracle's replay of the recorded accepted step sizes reproduces the
  oracle's adaptive solve bit for bit (so the replay is the discrete map the
  adaptive solver follows).
* The NumPy adjoint's gradients with respect to initial VAR concentrations,
  fixed species and RCONST equal central finite differences of the replay
  (relative step 1e-6) to 1e-6 relative, on single parameters and on random
  directions through all parameters at once.
"""
import numpy as np
import pytest

from adjoint_ref import Monomials, adjoint_cell, log_loss, stoich_matrix
from gcchem_amd import inputs


@pytest.fixture(scope="module")
def cells(mech, oracle_cr):
    d = inputs.random_cells(mech, n=3, seed=3)
    out = []
    for c in range(3):
        C, IS, RS, IE, Hs = oracle_cr.solve_cell_rec(
            d["Tstart"], d["Tend"], d["C"][c], d["RCONST"][c], d["ATOL"], d["RTOL"],
            d["ICNTRL"][c], d["RCNTRL"][c])
        assert IE == 1 and len(Hs) == IS[3]
        out.append(dict(V=d["C"][c][:mech.nvar].copy(), F=d["C"][c][mech.nvar:].copy(),
                        RCT=d["RCONST"][c].copy(), Hs=Hs, C=C, atol=d["ATOL"]))
    return out


def test_replay_reproduces_adaptive_solve(mech, oracle_cr, cells):
    for c in cells:
        Y = oracle_cr.replay(c["V"], c["F"], c["RCT"], c["Hs"])
        assert np.array_equal(Y.view(np.int64), c["C"][:mech.nvar].view(np.int64))


def _loss(orc, c, w, V, F, RCT):
    Y = orc.replay(V, F, RCT, c["Hs"])
    return log_loss(Y, w, c["atol"])[0]


def _fd(orc, c, w, which, j, rel=1e-6):
    args = {"V": c["V"].copy(), "F": c["F"].copy(), "RCT": c["RCT"].copy()}
    p = args[which][j]
    h = rel * abs(p)
    args[which][j] = p + h
    Lp = _loss(orc, c, w, **args)
    args[which][j] = p - h
    Lm = _loss(orc, c, w, **args)
    # rounding of the loss itself bounds what FD can resolve
    L0 = _loss(orc, c, w, c["V"], c["F"], c["RCT"])
    # rounding noise of the loss after ~100 stiff steps: 256 ulps of its
    # term magnitude
    noise = 2.0 ** -44 * (abs(L0) + np.sum(np.abs(w) * 40.0)) / (2 * h)
    return (Lp - Lm) / (2 * h), noise


@pytest.fixture(scope="module")
def grads(mech, oracle_cr, checkforms, cells):
    mono = Monomials(mech.ir)
    S = stoich_matrix(checkforms, mech.nvar, mech.nreact)
    rng = np.random.default_rng(11)
    out = []
    for c in cells:
        w = rng.standard_normal(mech.nvar)
        _, lam = log_loss(c["C"][:mech.nvar], w, c["atol"])
        gV, gF, gk = adjoint_cell(oracle_cr, mono, S, c["V"], c["F"], c["RCT"], c["Hs"],
                                  lam, np.zeros(mech.nfix))
        out.append((w, gV, gF, gk))
    return out


def _pick(rng, g, vals, n):
    """Parameters with non-zero values; half of them the largest |g*p|."""
    ok = np.nonzero(vals != 0.0)[0]
    sens = np.abs(g[ok] * vals[ok])
    top = ok[np.argsort(sens)[::-1][: n // 2]]
    rest = rng.choice(ok, size=n - top.size, replace=False)
    return np.unique(np.concatenate([top, rest]))


def test_adjoint_ref_matches_finite_differences(mech, oracle_cr, cells, grads):
    rng = np.random.default_rng(5)
    checked = 0
    for c, (w, gV, gF, gk) in zip(cells, grads):
        for which, g, vals, n in (("V", gV, c["V"], 8), ("RCT", gk, c["RCT"], 8),
                                  ("F", gF, c["F"], 3)):
            for j in _pick(rng, g, vals, n):
                fd, noise = _fd(oracle_cr, c, w, which, j)
                assert abs(g[j] - fd) <= 1e-6 * abs(fd) + noise, (which, j, g[j], fd, noise)
                checked += 1
    assert checked >= 40


def test_adjoint_ref_directional_derivative(mech, oracle_cr, cells, grads):
    """d/de L(p + e * p * u) at 0 equals sum_j g_j p_j u_j for random u."""
    rng = np.random.default_rng(9)
    for c, (w, gV, gF, gk) in zip(cells, grads):
        u = {k: rng.standard_normal(c[k].size) for k in ("V", "F", "RCT")}
        pred = (np.sum(gV * c["V"] * u["V"]) + np.sum(gF * c["F"] * u["F"])
                + np.sum(gk * c["RCT"] * u["RCT"]))
        h = 1e-7

        def at(e):
            return _loss(oracle_cr, c, w, *(c[k] * (1.0 + e * u[k]) for k in ("V", "F", "RCT")))
        fd = (at(h) - at(-h)) / (2 * h)
        assert abs(pred - fd) <= 1e-6 * abs(fd) + 1e-8 * np.sum(np.abs(w)), (pred, fd)
