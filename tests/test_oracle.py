"""Pin the CPU oracle (the checker of every GPU parity test).

* sparse LU / solves: against dense linear algebra on the same matrices;
* Integrate_Cell_S: the reference's known answers (the boundary self-test,
  gpulib/src/gc_gpu_batch_api.F90:349-376, and the multi-device dummy
  chemistry, gpulib/src/mgpu_test.F90:45-53), the closed-form step sizes of
  the self-test, the failure codes of gckpp_BatchIntegrator.F90:468-592, and
  what mathematics fixes for Rodas3: order 3 and L-stability of its
  stability function, and the Robertson problem against an independent stiff
  solver (SciPy Radau).
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

from gcchem_amd import inputs, mechanism, synthetic
from oracle import Oracle


def _dense(mech, vals):
    crow, icol = mech.ir["crow"], mech.ir["icol"]
    A = np.zeros((mech.nvar, mech.nvar))
    for i in range(mech.nvar):
        for k in range(crow[i], crow[i + 1]):
            A[i, icol[k]] = vals[k]
    return A


@pytest.fixture(scope="module")
def prepared(mech, oracle_libm):
    """Ghimj = -J + (1/(H*0.5)) I at a random state, H = 60 s."""
    rng = np.random.default_rng(3)
    V = np.exp(rng.uniform(np.log(1e2), np.log(1e10), mech.nvar))
    F = np.array([3e17, 1.9e19, 5e18])
    RCT = np.exp(rng.uniform(np.log(1e-15), np.log(1e-10), mech.nreact))
    jvs = oracle_libm.jac_sp(V, F, RCT)
    g = -jvs
    g[mech.ir["diag"]] += 1.0 / (60.0 * 0.5)
    return g


def test_decomp_equals_dense_lu(mech, oracle_libm, prepared):
    lu, ier = oracle_libm.decomp(prepared)
    assert ier == 0
    A = _dense(mech, prepared)
    LU = _dense(mech, lu)
    L = np.tril(LU, -1) + np.eye(mech.nvar)
    U = np.triu(LU)
    assert np.allclose(L @ U, A, rtol=1e-11, atol=1e-11 * np.abs(A).max())
    # entries outside the pattern stay zero: the pattern is the LU pattern
    Ld, Ud = _doolittle(A)
    assert np.allclose(L, Ld, rtol=1e-9, atol=1e-12)
    assert np.allclose(U, Ud, rtol=1e-9, atol=1e-12 * np.abs(Ud).max())


def _doolittle(A):
    n = A.shape[0]
    L, U = np.eye(n), A.copy()
    for k in range(n):
        L[k + 1:, k] = U[k + 1:, k] / U[k, k]
        U[k + 1:, :] -= np.outer(L[k + 1:, k], U[k, :])
        U[k + 1:, k] = 0.0
    return L, U


def test_solve_and_transpose_solve(mech, oracle_libm, prepared):
    lu, _ = oracle_libm.decomp(prepared)
    LU = _dense(mech, lu)
    M = (np.tril(LU, -1) + np.eye(mech.nvar)) @ np.triu(LU)
    b = np.random.default_rng(9).normal(size=mech.nvar)
    x = oracle_libm.solve(lu, b)
    ref = np.linalg.solve(M, b)
    assert np.allclose(x, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    xt = oracle_libm.solve_tr(lu, b)
    ref = np.linalg.solve(M.T, b)
    assert np.allclose(xt, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def _selftest_inputs(mech, n=1):
    C = np.full((n, mech.nspec), 1.0e6)
    R = np.zeros((n, mech.nreact))
    I = np.zeros((n, 20), np.int32)
    I[:, 0], I[:, 1], I[:, 2], I[:, 6], I[:, 14] = 1, 0, 4, 1, -1
    return C, R, np.full(mech.nvar, 1e-2), np.full(mech.nvar, 0.5e-2), I, np.zeros((n, 20))


@pytest.mark.parametrize("mode", ["libm", "cr"])
def test_selftest_known_answer(mech, mode):
    o = Oracle(mech.ir, pow_mode=mode)
    C, R, A, Rt, I, Rc = _selftest_inputs(mech)
    _, IS, RS, IE, _ = o.integrate_batch(0.0, 1200.0, C, R, A, Rt, I, Rc)
    # reference answer rc = 1, nsteps = 12 (docs/GPU_IN_MODEL_RESULTS.md:182-184)
    assert IE[0] == 1 and IS[0, 2] == 12
    # closed form: the error is negligible, so every step grows by FacMax = 6
    # from Hstart = 1e-5 until the last step is clamped to the end time
    H = 1e-5 * 6.0 ** np.arange(11)
    assert RS[0, 0] == 1200.0 and RS[0, 2] == 1200.0
    assert abs(RS[0, 1] - (1200.0 - H.sum())) < 1e-9
    assert list(IS[0]) == [36, 12, 12, 12, 0, 12, 48, 0]


def test_mgpu_dummy_chemistry(mech, oracle_libm):
    """mgpu_test.F90:45-53: C = 1e6 + mod(i,97), RCONST(1:50) = 1e-4*(1+mod(i,13)/13),
    [0,1200] s; the reference expects no failed cell."""
    n = 300
    C, R, A, Rt, I, Rc = _selftest_inputs(mech, n)
    i = np.arange(1, n + 1)
    C[:] = (1.0e6 + np.mod(i, 97))[:, None]
    R[:, :50] = (1.0e-4 * (1.0 + np.mod(i, 13) / 13.0))[:, None]
    _, IS, _, IE, _ = oracle_libm.integrate_batch(0.0, 1200.0, C, R, A, Rt, I, Rc)
    assert np.all(IE == 1)


def test_failure_codes(mech, oracle_libm):
    d = inputs.random_cells(mech, n=8, seed=21)
    I = d["ICNTRL"].copy()
    Rc = d["RCNTRL"].copy()
    I[0, 11] = 1      # auto-reduce requested -> -90
    I[1, 2] = 2       # not Rodas3 -> -2
    I[2, 3] = 1       # Max_no_steps = 1 -> -6 after two steps
    C, IS, RS, IE, _ = oracle_libm.integrate_batch(0.0, 600.0, d["C"], d["RCONST"], d["ATOL"],
                                                   d["RTOL"], I, Rc)
    assert IE[0] == -90 and IE[1] == -2 and IE[2] == -6
    assert np.all(IE[3:] == 1)
    assert np.all(IS[0] == 0) and np.all(IS[1] == 0) and IS[2, 2] == 2
    assert np.array_equal(C[0], d["C"][0]) and np.array_equal(C[1], d["C"][1])
    # step too small -> -7: T + 0.1*H == T at T = 1e20
    C, IS, RS, IE, _ = oracle_libm.integrate_batch(1e20, 1e20 + 1e7, d["C"][:1], d["RCONST"][:1],
                                                   d["ATOL"], d["RTOL"], d["ICNTRL"][:1],
                                                   d["RCNTRL"][:1])
    assert IE[0] == -7


def _one_species(literal_k=None):
    rx = [synthetic.Reaction({0: 1}, {})]
    return synthetic.build_ir("linear1", 1, 1, rx, reorder=False)


def _one_step(o, k, h):
    C = np.array([[1.0, 0.0]])
    R = np.array([[k]])
    I = np.zeros((1, 20), np.int32)
    I[0, 0], I[0, 2] = 1, 4
    Rc = np.zeros((1, 20))
    Rc[0, 0] = h      # Hmin = h: the step is accepted whatever its error
    Rc[0, 2] = h      # Hstart = h
    C1, IS, RS, IE, _ = o.integrate_batch(0.0, h, C, R, np.array([1e-30]), np.array([1e-12]), I, Rc)
    assert IE[0] == 1 and IS[0, 2] == 1
    return C1[0, 0]


def test_rodas3_order_and_l_stability():
    """y' = -k y: one step multiplies y by R(z), z = -k h.  Rodas3 is of order
    3 (local error ~ z^4) and L-stable (R(z) -> 0 as z -> -inf)."""
    o = Oracle(_one_species(), pow_mode="cr")
    errs = []
    for z in (0.02, 0.04, 0.08):
        errs.append(abs(_one_step(o, z, 1.0) - np.exp(-z)))
    r1, r2 = errs[1] / errs[0], errs[2] / errs[1]
    assert 13.0 < r1 < 19.0 and 13.0 < r2 < 19.0  # 2^4 = 16
    assert abs(_one_step(o, 1e8, 1.0)) < 1e-6
    for z in np.linspace(0.1, 100.0, 40):
        assert abs(_one_step(o, z, 1.0)) < 1.0


def test_robertson_against_radau():
    """Robertson's stiff kinetics (A -> B, 2B -> B + C, B + C -> A + C) over
    40 s at tight tolerance against SciPy's Radau IIA."""
    rx = [synthetic.Reaction({0: 1}, {1: 1.0}),
          synthetic.Reaction({1: 2}, {1: 1.0, 2: 1.0}),
          synthetic.Reaction({1: 1, 2: 1}, {0: 1.0, 2: 1.0})]
    ir = synthetic.build_ir("robertson", 3, 1, rx, reorder=False)
    o = Oracle(ir, pow_mode="cr")
    k = np.array([[0.04, 3e7, 1e4]])
    C = np.array([[1.0, 0.0, 0.0, 0.0]])
    I = np.zeros((1, 20), np.int32)
    I[0, 0], I[0, 2] = 1, 4
    C1, IS, RS, IE, _ = o.integrate_batch(0.0, 40.0, C, k, np.full(3, 1e-14), np.full(3, 1e-7),
                                          I, np.zeros((1, 20)))
    assert IE[0] == 1

    def f(t, y):
        a, b, c = y
        return [-0.04 * a + 1e4 * b * c, 0.04 * a - 1e4 * b * c - 3e7 * b * b, 3e7 * b * b]

    ref = solve_ivp(f, (0.0, 40.0), [1.0, 0.0, 0.0], method="Radau", rtol=1e-11, atol=1e-16).y[:, -1]
    assert np.allclose(C1[0, :3], ref, rtol=2e-5, atol=0)


def test_pow_modes_agree_to_tolerance(mech):
    """libm pow and the correctly rounded Err**(1/3) differ in the last ulp for
    a small fraction of arguments; the solutions stay within 1e-9."""
    d = inputs.random_cells(mech, n=64, seed=7)
    args = (d["Tstart"], d["Tend"], d["C"], d["RCONST"], d["ATOL"], d["RTOL"], d["ICNTRL"], d["RCNTRL"])
    a = Oracle(mech.ir, pow_mode="libm").integrate_batch(*args)
    b = Oracle(mech.ir, pow_mode="cr").integrate_batch(*args)
    nv = mech.nvar
    assert np.array_equal(a[3], b[3])
    assert (a[1] == b[1]).all(axis=1).mean() >= 0.999 or (a[1] != b[1]).any(axis=1).sum() <= 1
    big = np.abs(a[0][:, :nv]) > 1e-2
    rel = np.abs(a[0][:, :nv] - b[0][:, :nv]) / np.maximum(np.abs(a[0][:, :nv]), 1e-300)
    assert rel[big].max() <= 1e-9


# ---- step control: rejections before and after the first acceptance -------
QUAD1 = dict(k=1e-3, p=0.01, y0=100.0, Tend=100.0, Hstart=300.0, atol=1e-8, rtol=1e-4)


def rodas3_scalar_by_hand(f, J, y0, Tend, Hstart, atol, rtol):
    """Rodas3 for one species (NVAR = 1) written out from the tableau and the
    control flow of gckpp_BatchIntegrator.F90:464-707, with RCNTRL = 0 except
    Hstart: the 1x1 "LU" is the prepared diagonal -J + 1/(0.5 H), each stage
    is one division.  Returns (y, ISTAT, RSTAT, attempts) with attempts =
    [(accepted, H, Err, FacRej applied / Hnew clipped to H, after first acc)]."""
    ros_A = (0.0, 2.0, 0.0, 2.0, 0.0, 1.0)            # :443-452
    ros_C = (4.0, 1.0, -1.0, 1.0, -1.0, -(8.0 / 3.0))
    ros_M = (2.0, 0.0, 1.0, 1.0)
    ros_E = (0.0, 0.0, 0.0, 1.0)
    FacMin, FacMax, FacRej, FacSafe = 0.2, 6.0, 0.1, 0.9
    ROUNDOFF = 2.0 ** -52
    Hmin, Hmax = 0.0, Tend                            # :504-512, span = Tend - 0
    H = min(abs(Hstart), Tend)
    H = min(max(abs(Hmin), abs(H)), abs(Hmax))
    T, y = 0.0, y0
    ist = [0] * 8
    rst = [0.0] * 3
    rej_last = rej_more = False
    attempts = []
    while (T - Tend) + ROUNDOFF <= 0.0:               # :535
        H = min(H, abs(Tend - T))
        f0 = f(y); ist[0] += 1
        j0 = J(y); ist[1] += 1
        while True:
            g = -j0 + 1.0 / (H * 0.5); ist[5] += 1
            K = []
            for s in range(4):
                fs = f0 if s < 2 else None
                if fs is None:                        # ros_NewF(3), ros_NewF(4)
                    p = y
                    for j in range(s):
                        p = p + ros_A[s * (s - 1) // 2 + j] * K[j]
                    fs = f(p); ist[0] += 1
                r = fs
                for j in range(s):
                    r = r + (ros_C[s * (s - 1) // 2 + j] / H) * K[j]
                K.append(r / g); ist[6] += 1
            ynew = y
            for j in range(4):
                ynew = ynew + ros_M[j] * K[j]
            yerr = 0.0
            for j in range(4):
                yerr = yerr + ros_E[j] * K[j]
            q = yerr / (atol + rtol * max(abs(y), abs(ynew)))
            err = max(np.sqrt(q * q / 1.0), 1.0e-10)
            hnew = H * min(FacMax, max(FacMin, FacSafe / err ** (1.0 / 3.0)))
            ist[2] += 1
            if err <= 1.0:                            # accept, :676-692
                ist[3] += 1
                y = ynew
                T = T + H
                hnew = max(Hmin, min(hnew, Hmax))
                clipped = rej_last and hnew > H
                if rej_last:
                    hnew = min(hnew, H)               # :685
                rst = [T, H, hnew]
                rej_last = rej_more = False
                attempts.append((True, H, err, clipped, ist[3] > 1))
                H = hnew
                break
            facrej = rej_more                         # reject, :693-698
            if rej_more:
                hnew = H * FacRej
            rej_more = rej_last
            rej_last = True
            attempts.append((False, H, err, facrej, ist[3] >= 1))
            H = hnew
            if ist[3] >= 1:
                ist[4] += 1                           # Nrej only after the first acceptance
    return y, ist, rst, attempts


def _quad1_case(c=QUAD1):
    ir = synthetic.build_ir("quad1", 1, 1, [synthetic.Reaction({0: 2}, {}),
                                            synthetic.Reaction({}, {0: 1.0}, {0: 1})],
                            reorder=False)
    C = np.array([[c["y0"], 1.0]])
    R = np.array([[c["k"], c["p"]]])
    I = np.zeros((1, 20), np.int32)
    I[0, 0], I[0, 2] = 1, 4
    Rc = np.zeros((1, 20))
    Rc[0, 2] = c["Hstart"]
    return ir, C, R, I, Rc


def singular_exit_cells(n):
    """Cells of the auto1 mechanism (X -> 2X, so J = +k) with k = 0 and
    Hstart = inf over [0, inf): H = min(max(Hmin, Hstart), Hmax) = inf, so
    ghinv = 1/(0.5 H) = 0 and the prepared diagonal -J + ghinv is 0 on every
    try, and halving H leaves it inf."""
    rng = np.random.default_rng(7)
    C = np.stack([np.exp(rng.uniform(0.0, 5.0, n)), np.ones(n)], axis=1)
    I = np.zeros((n, 20), np.int32)
    I[:, 0], I[:, 2] = 1, 4
    Rc = np.zeros((n, 20))
    Rc[:, 2] = np.inf
    return dict(C=C, RCONST=np.zeros((n, 1)), ICNTRL=I, RCNTRL=Rc, ATOL=np.array([1.0]),
                RTOL=np.array([1e-3]), Perm=np.arange(1, n + 1, dtype=np.int32),
                Tstart=0.0, Tend=np.inf)


def test_singular_six_times_known_answer():
    """Six singular factorisations in a row end a cell with IERR -8: after
    each of the first five H is halved and the matrix prepared again, the
    sixth returns (gckpp_BatchIntegrator.F90:572-592).  Counters by hand: one
    Fun and one Jac_SP for the step, six decompositions, all six singular,
    no stage solve, no step taken; C comes back unchanged."""
    d = singular_exit_cells(4)
    for pm in ("cr", "libm"):
        o = Oracle(mechanism.load("auto1").ir, pow_mode=pm)
        C, IS, RS, IE = o.integrate_batch(d["Tstart"], d["Tend"], d["C"], d["RCONST"], d["ATOL"],
                                          d["RTOL"], d["ICNTRL"], d["RCNTRL"])[:4]
        assert np.all(IE == -8)
        assert np.all(IS == np.array([1, 1, 0, 0, 0, 6, 0, 6]))
        assert np.array_equal(C[:, 0], d["C"][:, 0])


def test_step_control_known_answer():
    """2X -> (nothing) at k and F -> X at p (F = 1): Vdot = p - ((k*2)*y)*y,
    J = -(2*((k*2)*y)) in KPP's generated forms.  With Hstart = 300 s
    (clamped to the 100 s span) and rtol 1e-4 the sequence of attempts holds:
    three rejections before the first acceptance, which Nrej does not count,
    the third with H*FacRej; an acceptance right after rejections whose Hnew
    is clipped to H; three counted rejections later, again ending in
    H*FacRej and a clipped acceptance.  The oracle (libm pow, as the hand
    version) gives the same ISTAT bit for bit and the same y and RSTAT."""
    c = QUAD1
    k, p = c["k"], c["p"]
    y, ist, rst, att = rodas3_scalar_by_hand(
        lambda v: p * 1.0 - ((k * 2.0) * v) * v, lambda v: -(2.0 * ((k * 2.0) * v)),
        c["y0"], c["Tend"], c["Hstart"], c["atol"], c["rtol"])
    first = next(n for n, a in enumerate(att) if a[0])
    assert first == 3 and not any(a[0] for a in att[:3])
    assert att[2][3] and not att[1][3]               # FacRej on the third rejection
    assert att[first][3]                              # Hnew clipped to H
    later = [a for a in att[first:] if not a[0]]
    assert len(later) == 3 and later[-1][3]
    assert ist[4] == 3 and ist[2] - ist[3] == 6
    assert sum(1 for a in att if a[0] and a[3]) == 2
    ir, C, R, I, Rc = _quad1_case()
    for mode in ("libm", "cr"):
        o = Oracle(ir, pow_mode=mode)
        C1, IS, RS, IE, _ = o.integrate_batch(0.0, c["Tend"], C, R, np.array([c["atol"]]),
                                              np.array([c["rtol"]]), I, Rc)
        assert IE[0] == 1
        assert list(IS[0]) == ist, mode
        assert C1[0, 0] == y and list(RS[0]) == rst, mode


# ---- auto-reduce masks (DO_FUN / DO_JVS / DO_SLV) --------------------------
def test_masked_functions_guard_their_statements(mech, oracle_cr):
    """Each guard skips exactly its statement: a species whose DO_FUN flag is
    off keeps the P/D passed in (zero: Vdot = 0), an entry whose DO_JVS flag
    is off keeps the JVS passed in, every other output equals the unmasked
    routine's bit for bit; all flags on is the unmasked routine."""
    o, n, nnz = oracle_cr, mech.nvar, mech.nnz
    rng = np.random.default_rng(4)
    V = np.exp(rng.uniform(np.log(1e2), np.log(1e10), n))
    F = np.array([3e17, 1.9e19, 5e18])
    R = np.exp(rng.uniform(np.log(1e-15), np.log(1e-10), mech.nreact))
    fon = rng.uniform(size=n) < 0.6
    vd, P, D, _ = o.fun_split(V, F, R)
    vm, Pm, Dm, _ = o.fun_split(V, F, R, do_fun=o.mask_words(fon))
    assert np.array_equal(vm[fon], vd[fon]) and np.all(vm[~fon] == 0.0)
    P0, D0 = np.full(n, 7.0), np.full(n, 0.5)
    vs, Ps, Ds, _ = o.fun_split(V, F, R, do_fun=o.mask_words(fon), P0=P0, D0=D0)
    assert np.all(Ps[~fon] == 7.0) and np.all(Ds[~fon] == 0.5)
    assert np.array_equal(vs[~fon], 7.0 - 0.5 * V[~fon])
    assert np.array_equal(o.fun_split(V, F, R, do_fun=o.mask_words(np.ones(n, bool)))[0], vd)
    jon = rng.uniform(size=nnz) < 0.6
    j = o.jac_sp(V, F, R)
    jm = o.jac_sp(V, F, R, do_jvs=o.mask_words(jon), JVS0=np.full(nnz, -3.25))
    assert np.array_equal(jm[jon], j[jon]) and np.all(jm[~jon] == -3.25)


def test_masked_solve_equals_dense_substitution(mech, oracle_cr, prepared):
    """KppSolve with DO_SLV flags off for some rows: dense forward and backward
    substitution (unit L, then U) that leaves exactly those rows' x alone."""
    lu, _ = oracle_cr.decomp(prepared)
    LU = _dense(mech, lu)
    L, U = np.tril(LU, -1), np.triu(LU)
    rng = np.random.default_rng(6)
    on = rng.uniform(size=mech.nvar) < 0.7
    b = rng.normal(size=mech.nvar)
    x = b.copy()
    for i in range(mech.nvar):
        if on[i]:
            x[i] = x[i] - L[i, :i] @ x[:i]
    for i in range(mech.nvar - 1, -1, -1):
        if on[i]:
            x[i] = (x[i] - U[i, i + 1:] @ x[i + 1:]) / U[i, i]
    got = oracle_cr.solve(lu, b, do_slv=oracle_cr.mask_words(on))
    assert np.allclose(got, x, rtol=1e-10, atol=1e-12 * np.abs(x).max())
    assert np.array_equal(got[~on], b[~on])


def test_masked_species_frozen_robertson_against_radau():
    """Robertson's kinetics with species C switched off the way auto-reduce
    switches off an inactive species -- its DO_FUN and DO_SLV flags and the
    DO_JVS flags of its Jacobian row and column -- integrate the reduced
    system with C held at its initial value: C comes back unchanged and A, B
    agree with SciPy's Radau on A' = -k1 A + k3 B C0, B' = k1 A - k3 B C0 - k2 B^2.
    A batch with every flag on equals the unmasked solve bit for bit."""
    rx = [synthetic.Reaction({0: 1}, {1: 1.0}),
          synthetic.Reaction({1: 2}, {1: 1.0, 2: 1.0}),
          synthetic.Reaction({1: 1, 2: 1}, {0: 1.0, 2: 1.0})]
    ir = synthetic.build_ir("robertson", 3, 1, rx, reorder=False)
    o = Oracle(ir, pow_mode="cr")
    k = np.array([[0.04, 3e7, 1e4]])
    C0 = 0.05
    C = np.array([[1.0, 0.0, C0, 0.0]])
    I = np.zeros((1, 20), np.int32)
    I[0, 0], I[0, 2] = 1, 4
    args = (0.0, 40.0, C, k, np.full(3, 1e-14), np.full(3, 1e-7), I, np.zeros((1, 20)))
    rows = np.repeat(np.arange(3), np.diff(ir["crow"]))
    cols = np.asarray(ir["icol"])
    frozen = np.array([False, False, True])
    masks = (o.mask_words(~frozen)[None], o.mask_words(~(frozen[rows] | frozen[cols]))[None],
             o.mask_words(~frozen)[None])
    C1, IS, RS, IE, _ = o.integrate_batch(*args, masks=masks)
    assert IE[0] == 1 and C1[0, 2] == C0

    def f(t, y):
        a, b = y
        return [-0.04 * a + 1e4 * b * C0, 0.04 * a - 1e4 * b * C0 - 3e7 * b * b]

    ref = solve_ivp(f, (0.0, 40.0), [1.0, 0.0], method="Radau", rtol=1e-11, atol=1e-16).y[:, -1]
    assert np.allclose(C1[0, :2], ref, rtol=2e-5, atol=0)
    full = o.integrate_batch(*args)
    on = tuple(np.full_like(m_, 0xFFFFFFFF) for m_ in masks)
    alln = o.integrate_batch(*args, masks=on)
    assert all(np.array_equal(a, b) for a, b in zip(full[:4], alln[:4]))
