"""Pin the fullchem mechanism IR (and so the oracle and the generated kernels)
against the reference's *independent* generated forms and against calculus.

The IR comes from Fun_SPLIT / Jac_SP / KppSolve (kpp_ingest.py).  The
fixture tests/golden/fullchem_checkforms.json.gz holds other routines of the
same reference files, parsed separately: Fun (aggregate), CalcStoichNum,
Jac_SP_Vec, JacTR_SP_Vec and KppSolveTR.  A dropped term, a wrong sign, a
wrong index or a transposed entry in the IR fails one of these checks.
"""
import numpy as np
import pytest


def _prod(p, vals):
    v = 1.0
    for kind, idx in p:
        v = v * (float(idx) if kind == "lit" else vals[kind][idx])
    return v


def _sum(terms, vals):
    return sum((s * _prod(p, vals) for s, p in terms), 0.0)


def _abs_sum(terms, vals):
    return sum((abs(_prod(p, vals)) for s, p in terms), 0.0)


@pytest.fixture(scope="module")
def state(mech):
    rng = np.random.default_rng(11)
    V = np.exp(rng.uniform(np.log(1e2), np.log(1e9), mech.nvar))
    F = np.array([3e17, 1.9e19, 5e18])
    RCT = np.exp(rng.uniform(np.log(1e-16), np.log(1e-8), mech.nreact))
    return V, F, RCT


def test_dimensions(mech):
    # NVAR/NFIX/NREACT/LU_NONZERO of the generated fullchem mechanism
    # (gckpp_Function.F90:29 DO_FUN(353), gckpp_Jacobian.F90:31 DO_JVS(5683))
    assert (mech.nvar, mech.nfix, mech.nreact, mech.nnz) == (353, 3, 1058, 5683)


def test_lu_pattern(mech):
    crow, icol, diag = mech.ir["crow"], mech.ir["icol"], mech.ir["diag"]
    nL = sum(diag[i] - crow[i] for i in range(mech.nvar))
    nU = sum(crow[i + 1] - diag[i] - 1 for i in range(mech.nvar))
    assert (nL, nU) == (2773, 2557)
    for i in range(mech.nvar):
        cols = icol[crow[i]:crow[i + 1]]
        assert cols == sorted(cols) and icol[diag[i]] == i
    # closed under the IKJ elimination of KppDecomp (no lost fill)
    for k in range(mech.nvar):
        row = set(icol[crow[k]:crow[k + 1]])
        for kk in range(crow[k], diag[k]):
            j = icol[kk]
            assert set(icol[diag[j] + 1:crow[j + 1]]) <= row


def test_rates_match_aggregate_fun(mech, checkforms):
    # Fun (aggregate form) repeats the rate monomials; they must be identical
    assert checkforms["fun_rate"] == mech.ir["rate"]


def test_fun_split_vs_aggregate_and_stoichiometry(mech, oracle_libm, checkforms, state):
    V, F, RCT = state
    vdot, P, D, A = oracle_libm.fun_split(V, F, RCT)
    vals = {"V": V, "F": F, "RCT": RCT, "A": A}
    # A(r) evaluated independently
    A_ref = np.array([_prod(f, vals) for f in mech.ir["rate"]])
    assert np.array_equal(A, A_ref)
    # aggregate Vdot sums (a different grouping of the same terms)
    agg = np.array([_sum(s, vals) for s in checkforms["fun_vdot"]])
    scale = np.array([_abs_sum(s, vals) for s in checkforms["fun_vdot"]]) + 1e-300
    assert np.all(np.abs(vdot - agg) <= 1e-12 * scale)
    # stoichiometric matrix from CalcStoichNum: Vdot = S @ A
    S = np.zeros((mech.nvar, mech.nreact))
    for i, r, v in checkforms["stoich"]:
        S[i, r] = float(v)
    svdot = S @ A
    assert np.all(np.abs(vdot - svdot) <= 1e-11 * (np.abs(S) @ np.abs(A) + 1e-300))


def test_jacobian_pattern_matches_jac_vec_forms(mech, checkforms):
    pat = [(i, j) for i, j, _ in mech.ir["jvs"]]
    for i, terms in enumerate(checkforms["jac_vec"]):
        for k, j in terms:
            assert pat[k] == (i, j)
    for j, terms in enumerate(checkforms["jactr_vec"]):
        for k, i in terms:
            assert pat[k] == (i, j)
    # Jac_SP_Vec skips most structural fill entries (JVS = 0) but covers every
    # entry that has a formula
    covered = {k for t in checkforms["jac_vec"] for k, _ in t}
    formula = {k for k, (_, _, s) in enumerate(mech.ir["jvs"]) if s}
    assert formula <= covered


def test_jacobian_analytic(mech, oracle_libm, checkforms, state):
    """JVS(i,j) = sum_r S(i,r) dA_r/dV_j, from the monomials and CalcStoichNum."""
    V, F, RCT = state
    jvs = oracle_libm.jac_sp(V, F, RCT)
    S = {}
    for i, r, v in checkforms["stoich"]:
        S.setdefault(r, []).append((i, float(v)))
    vals = {"V": V, "F": F, "RCT": RCT}
    J = {}
    Jabs = {}
    for r, f in enumerate(mech.ir["rate"]):
        vs = [idx for kind, idx in f if kind == "V"]
        for s in set(vs):
            # d/dV_s of c * prod V: multiplicity * product without one V_s
            m = vs.count(s)
            rest = list(f)
            rest.remove(["V", s])
            d = m * _prod(rest, vals)
            for i, c in S.get(r, []):
                J[(i, s)] = J.get((i, s), 0.0) + c * d
                Jabs[(i, s)] = Jabs.get((i, s), 0.0) + abs(c * d)
    for k, (i, j, terms) in enumerate(mech.ir["jvs"]):
        ref = J.get((i, j), 0.0)
        if not terms:
            assert jvs[k] == 0.0  # structural fill entry
        else:
            assert abs(jvs[k] - ref) <= 1e-11 * Jabs[(i, j)] + 1e-300, (k, i, j)
    # every analytic nonzero is in the pattern
    pat = {(i, j) for i, j, _ in mech.ir["jvs"]}
    assert all(key in pat for key, v in J.items() if Jabs[key] > 0)


def test_jacobian_finite_differences(mech, oracle_libm, state):
    V, F, RCT = state
    jvs = oracle_libm.jac_sp(V, F, RCT)
    pat = [(i, j) for i, j, _ in mech.ir["jvs"]]
    rng = np.random.default_rng(5)
    for j in rng.choice(mech.nvar, 12, replace=False):
        h = 1e-6 * V[j]
        Vp, Vm = V.copy(), V.copy()
        Vp[j] += h
        Vm[j] -= h
        fp, Pp, Dp, _ = oracle_libm.fun_split(Vp, F, RCT)
        fm, _, _, _ = oracle_libm.fun_split(Vm, F, RCT)
        fd = (fp - fm) / (2 * h)
        # roundoff of the difference quotient: Vdot = P - D*V cancels
        noise = 50 * np.finfo(float).eps * (np.abs(Pp) + np.abs(Dp * Vp)) / h
        col = {i: jvs[k] for k, (i, jj) in enumerate(pat) if jj == j}
        for i in range(mech.nvar):
            a = col.get(i, 0.0)
            assert abs(fd[i] - a) <= 1e-6 * abs(a) + noise[i] + 1e-300, (i, j)


def test_kpp_solve_tr_order(mech, checkforms):
    """KppSolveTR's statements follow the rule the oracle and the generator
    use: U^T rows ascending with U(j,i) for ascending j, then L^T rows
    descending with L(j,i) for ascending j."""
    crow, icol, diag = mech.ir["crow"], mech.ir["icol"], mech.ir["diag"]
    n = mech.nvar
    ucol = [[] for _ in range(n)]
    lcol = [[] for _ in range(n)]
    for j in range(n):
        for k in range(diag[j] + 1, crow[j + 1]):
            ucol[icol[k]].append([k, j])
        for k in range(crow[j], diag[j]):
            lcol[icol[k]].append([k, j])
    st = checkforms["solve_tr"]
    phase1 = [s for s in st if s[2] == "X"]
    phase2 = [s for s in st if s[2] == "XX"]
    assert [s[1] for s in phase1] == list(range(n))
    for lhs, i, src, si, terms, d in phase1:
        assert si == i and d == diag[i] and terms == ucol[i]
    assert [s[1] for s in phase2] == list(reversed(range(n)))
    for lhs, i, src, si, terms, d in phase2:
        assert si == i and d is None and terms == lcol[i]
