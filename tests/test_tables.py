"""The generated operand tables, interpreted on the CPU, against the oracle.

``TableMachine`` walks the int32 tables of codegen/gen_tables.py with exactly
the loops of the device routines in csrc/gc_rodas3.hip (fixed-shape product
groups, folded-sign sums, the fused Jacobian/LU ops with an LDS row segment,
the padded solve rows), vectorised over lanes with numpy (no fused multiply-add,
IEEE double, same order of operations).  Every result must equal the oracle's
bit for bit, so a table generator mistake (a wrong offset, a dropped term, a
lost sign, a padding entry that is not a no-op) fails here without a GPU.
"""
import importlib.util
import struct

import numpy as np
import pytest

from conftest import PKG


def _gen_tables():
    spec = importlib.util.spec_from_file_location("gen_tables", PKG / "codegen" / "gen_tables.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


GT = _gen_tables()


def _coef(lo, hi):
    return struct.unpack("<d", struct.pack("<ii", lo, hi))[0]


class TableMachine:
    """CPU interpreter of one mechanism's tables over ``lanes`` cells."""

    def __init__(self, ir, lanes, seg):
        self.t = GT.Tables(ir)
        self.E = self.t.E
        self.ir = ir
        self.seg = seg
        self.prod = {"rate_y": self.t.rates("Y"), "rate_yn": self.t.rates("YNEW"),
                     "bjac": self.t.bjac()}
        self.pd = {"y": self.t.pd("Y"), "yn": self.t.pd("YNEW")}
        self.lu_ops = self.t.lu(seg)
        self.fwd = [self.t.fwd("K", s) for s in range(4)]
        self.bwd = [self.t.bwd_lds("K", s, seg) for s in range(4)]
        self.tr = self.t.tr("TMP")
        self.nu_tab = self.t.adj_nu()
        self.nu_stream = self.t.adj_nu_stream()
        self.mono = {"Y": self.t.adj_mono("Y"), "YNEW": self.t.adj_mono("YNEW")}
        self.wk = {"Y": self.t.adj_wk("Y"), "YNEW": self.t.adj_wk("YNEW")}
        self.gk2 = self.t.adj_gk2()
        self.lp = {(False, "Y"): self.t.adj_lp(False, "Y"),
                   (False, "YNEW"): self.t.adj_lp(False, "YNEW"),
                   (True, "Y"): self.t.adj_lp(True, "Y")}
        self.lits = np.array(self.t.lits or [0.0])
        self.ws = np.full((self.E["ADJ_TOTAL"], lanes), np.nan)
        self.ws[self.E["ZERO"]] = 0.0
        self.ws[self.E["ONE"]] = 1.0
        self.ws[self.E["NZERO"]] = -0.0
        self.lds = np.full((seg + 1, lanes), np.nan)
        # auto-reduce flags per lane ([lanes, n] booleans) for the masked
        # kernels' guards; None: unmasked kernels
        self.do_fun = self.do_jvs = self.do_slv = None

    # workspace access by byte offset e*512 (a lane's column)
    def ld(self, soff):
        assert soff % 512 == 0
        return self.ws[soff // 512]

    def st(self, soff, v):
        self.ws[soff // 512] = v

    def products(self, name):
        words, counts = self.prod[name]
        r = 0
        for shape, n in enumerate(counts):
            nf, lit = shape // 2 + 1, shape % 2
            for _ in range(n):
                rec = words[8 * r:8 * r + 8]
                v = self.ld(rec[2]).copy()
                if lit:
                    v = self.lits[rec[1]] * v
                for f in range(1, nf):
                    v = v * self.ld(rec[2 + f])
                self.st(rec[0], v)
                r += 1
        assert 8 * r == len(words)

    def sum_terms(self, w, p, n):
        acc = _coef(w[p + 2], w[p + 3]) * self.ld(w[p])
        for k in range(1, n):
            q = p + 4 * k
            acc = acc + _coef(w[q + 2], w[q + 3]) * self.ld(w[q])
        return acc, p + 4 * n

    def dterm(self, w, q):
        return ((self.lits[w[q + 3]] * self.ld(w[q])) * self.ld(w[q + 1])) * self.ld(w[q + 2])

    def sum_dterms(self, w, p, n):
        acc = self.dterm(w, p)
        for k in range(1, n):
            acc = acc + self.dterm(w, p + 4 * k)
        return acc, p + 4 * n

    def fun(self, ynew, out):
        sfx = "yn" if ynew else "y"
        self.products("rate_" + sfx)
        w, p = self.pd[sfx], 0
        for _ in range(self.ir["nvar"]):
            nP, nD, vo, oo = w[p:p + 4]
            assert nP % 4 == 0 and nD % 4 == 0
            p += 4
            P = D = 0.0
            kp = kd = 0
            while kp < nP or kd < nD:  # batches: up to 8 P terms, then up to 4 D terms
                pc, dc = min(8, nP - kp), min(4, nD - kd)
                for r in range(pc):
                    v = _coef(w[p + 2], w[p + 3]) * self.ld(w[p])
                    P = v if kp + r == 0 else P + v
                    p += 4
                for r in range(dc):
                    v = self.dterm(w, p)
                    D = v if kd + r == 0 else D + v
                    p += 4
                kp, kd = kp + pc, kd + dc
            if self.do_fun is not None:
                on = self.do_fun[:, oo // 512]
                P, D = np.where(on, P, 0.0), np.where(on, D, 0.0)
            self.st(out + oo, P - D * self.ld(vo))
        assert p == len(w)

    def jac_lu(self, ghinv):
        self.products("bjac")
        w, p, sing = self.lu_ops, 0, np.zeros(self.ws.shape[1], bool)
        lds = lambda off: self.lds[off // 512]
        while True:
            code = w[p]
            if code == 0:
                break
            if code == 1:
                slot, nt, isdiag = w[p + 1:p + 4]
                p += 4
                if nt:
                    v, p = self.sum_terms(w, p, nt)
                else:
                    v = np.zeros(self.ws.shape[1])
                g = -v
                if isdiag:
                    g = g + ghinv
                    sing |= np.abs(g) < 2.2250738585072014e-308
                self.lds[slot // 512] = g
            elif code == 5:
                T = w[p + 1]
                assert T % 4 == 0
                p += 4
                acc = None
                for q in range(T):
                    soff, fl = w[p], w[p + 1]
                    cx = _coef(w[p + 2], w[p + 3]) * self.ld(soff)
                    acc = cx if fl & 1 else acc + cx
                    if fl & 2:
                        if self.do_jvs is not None:
                            acc = np.where(self.do_jvs[:, fl >> 14], acc, 0.0)
                        g = -acc
                        if fl & 4:
                            g = g + ghinv
                            sing |= np.abs(g) < 2.2250738585072014e-308
                        self.lds[(fl >> 8) & 63] = g
                    p += 4
            elif code == 8:
                G = w[p + 1]
                assert 1 <= G <= GT.PIVB_MAX
                npk = -(-G // 4) * 4
                pk = w[p + 4:p + 4 + G]
                d = w[p + 4 + npk:p + 4 + npk + 5 * G]
                p += 4 + npk + -(-5 * G // 4) * 4
                for g in range(G):
                    word = pk[g] & 0xFFFFFFFF
                    kind, sj = word & 3, ((word >> 2) & 63) * 512
                    if kind == 1:
                        a = -lds(sj) / self.ld(d[5 * g])
                        self.lds[sj // 512] = -a
                    elif kind == 2:
                        a = -self.ld(d[5 * g])
                    else:
                        assert kind == 0  # continues the previous sub-record's multiplier
                    for k in range(4):
                        t = ((word >> (8 + 6 * k)) & 63) * 512
                        self.lds[t // 512] = lds(t) + a * self.ld(d[5 * g + 1 + k])
            else:
                assert code == 4
                n = w[p + 1]
                assert n % 2 == 0
                p += 4
                for _ in range(n):
                    self.st(w[p + 1], lds(w[p]).copy())
                    p += 2
        return sing

    def _slv_on(self, xi, xbase):
        return self.do_slv[:, xi // 512 - xbase]

    def solve_rows(self, rows, div, xbase=None):
        w, nrows = rows
        p = 0
        for _ in range(nrows):
            n, xi, dg = w[p:p + 3]
            assert n % 4 == 0
            p += 4
            x = self.ld(xi).copy()
            x0 = x
            for _ in range(n):
                x = x - self.ld(w[p]) * self.ld(w[p + 1])
                p += 2
            if div:
                x = x / self.ld(dg)
            if xbase is not None and self.do_slv is not None:
                x = np.where(self._slv_on(xi, xbase), x, x0)
            self.st(xi, x)
        assert p == len(w)

    def solve_bwd_lds(self, rows, xbase):
        """Backward rows with the last seg columns' x_j in LDS (bwd_lds)."""
        w, nrows = rows
        lds = lambda off: self.lds[off // 512]
        self.lds[:] = np.nan                        # what the LU left there is never read,
        self.lds[self.seg] = 0.0                    # but the spare slot: LDS padding operand
        p = 0
        for _ in range(nrows):
            nn, xi, dg, li = w[p:p + 4]
            nc, nh = nn & 0xFFFF, nn >> 16
            assert nc % 4 == 0 and nh % 4 == 0
            p += 4
            x = self.ld(xi).copy()
            x0 = x
            for _ in range(nc):
                x = x - self.ld(w[p]) * self.ld(w[p + 1])
                p += 2
            for _ in range(nh):
                x = x - self.ld(w[p]) * lds(w[p + 1])
                p += 2
            x = x / self.ld(dg)
            if self.do_slv is not None:
                x = np.where(self._slv_on(xi, xbase), x, x0)
            self.st(xi, x)
            if li >= 0:
                assert li // 512 < self.seg
                self.lds[li // 512] = x
        assert p == len(w)

    def solve(self, s):
        xbase = self.E["K"] + s * self.ir["nvar"]
        self.solve_rows(self.fwd[s], False, xbase)
        self.solve_bwd_lds(self.bwd[s], xbase)

    # -- adjoint routines, as adj_* in gc_rodas3.hip ---------------------------
    ROS_A = (0.0, 2.0, 0.0, 2.0, 0.0, 1.0)
    ROS_C = (4.0, 1.0, -1.0, 1.0, -1.0, -(8.0 / 3.0))
    ROS_M = (2.0, 0.0, 1.0, 1.0)

    def o(self, name, k=0):
        return (self.E[name] + k) * 512

    def solve_tr(self):
        w1, n1, w2, n2 = self.tr
        self.solve_rows((w1, n1), True)
        self.solve_rows((w2, n2), False)

    def stage_point(self, s):
        n = self.ir["nvar"]
        for k in range(n):
            y = self.ld(self.o("Y", k)).copy()
            for j in range(s):
                y = y + self.ROS_A[s * (s - 1) // 2 + j] * self.ld(self.o("K", j * n + k))
            self.st(self.o("YNEW", k), y)

    def stages(self, dh):
        """Stage vectors; dh (scalar or per lane) is Dir*H."""
        n = self.ir["nvar"]
        for s in range(4):
            Ks = self.o("K", s * n)
            src = Ks
            if s < 2:
                src = self.o("FCN0")
            else:
                self.stage_point(s)
                self.fun(True, Ks)
            hc = [self.ROS_C[s * (s - 1) // 2 + j] / dh for j in range(s)]
            for k in range(n):
                v = self.ld(src + k * 512).copy()
                for j in range(s):
                    v = v + hc[j] * self.ld(self.o("K", j * n + k))
                self.st(Ks + k * 512, v)
            self.solve(s)

    def adj_nu(self):
        w, p = self.nu_tab, 0
        for _ in range(self.ir["nreact"]):
            n, dst = w[p], w[p + 1]
            p += 4
            if n:
                v, p = self.sum_terms(w, p, n)
            else:
                v = np.zeros(self.ws.shape[1])
            self.st(dst, v)
        assert p == len(w)

    def acc(self, g, v):
        self.st(g, self.ld(g) + v)

    def adj_wk(self, point):
        w = self.wk[point]
        assert len(w) % 32 == 0
        e = lambda v, sh: ((v >> sh) & 0xFFFF) * 512
        for p in range(0, len(w), 4):
            a, b, c, d = w[p:p + 4]
            cn = self.lits[(d >> 16) & 0xFFFF] * self.ld(e(b, 0))
            g = self.ld(e(a, 16))
            self.st(e(a, 0), cn * self.ld(e(b, 16)))
            self.st(e(a, 16), g + cn * ((self.ld(e(c, 0)) * self.ld(e(c, 16))) * self.ld(e(d, 0))))
            self.st(e(b, 0), cn)

    def adj_gk_second(self):
        w, counts = self.gk2
        e = lambda v, sh: ((v >> sh) & 0xFFFF) * 512
        L = lambda v, sh: self.ld(e(v, sh))
        p = 0
        for nf, n in zip((1, 2, 3), counts):
            assert n % 8 == 0
            for _ in range(n):
                a, b, c, d = w[p:p + 4]
                if nf == 1:
                    t = L(c, 16)
                elif nf == 2:
                    t = (L(b, 16) * L(c, 16)) + (L(b, 0) * L(d, 0))
                else:
                    t = (((L(b, 16) * L(c, 0)) * L(c, 16)) + ((L(b, 0) * L(c, 0)) * L(d, 0))) \
                        + ((L(b, 0) * L(b, 16)) * L(d, 16))
                self.st(e(a, 0), L(a, 0) + L(a, 16) * t)
                p += 4
        assert p == len(w)

    def stream_term(self, word, v, acc):
        acc = v if word & 1 else acc + v
        if word & 2:
            self.st(word & ~511, acc)
        return acc

    def adj_nu_stream(self):
        w = self.nu_stream
        assert len(w) % 32 == 0
        acc = None
        for p in range(0, len(w), 4):
            acc = self.stream_term(w[p + 1], _coef(w[p + 2], w[p + 3]) * self.ld(w[p]), acc)

    def adj_lp(self, second, point):
        w = self.lp[(second, point)]
        assert len(w) % 32 == 0
        e = lambda v, sh: ((v >> sh) & 0xFFFF) * 512
        acc = None
        for p in range(0, len(w), 4):
            x, y, z, word = w[p:p + 4]
            v = self.ld(e(x, 0)) * ((self.ld(e(x, 16)) * self.ld(e(y, 0)))
                                    + (self.ld(e(y, 16)) * self.ld(e(z, 0))))
            acc = self.stream_term(word, v, acc)

    def adj_mono_gather(self, second, point):
        """The kernel's form: W, then the k records, then the gather sums."""
        if not second:
            self.adj_wk(point)
        else:
            self.adj_gk_second()
        self.adj_lp(second, point)

    def adj_mono(self, second, point):
        """Scatter form of the same derivatives (per reaction, accumulating
        into LP/GF/GK), kept to check the gather form bit for bit."""
        q = self.mono[point]
        for r in range(self.ir["nreact"]):
            rec = q[16 * r:16 * r + 16]
            nf, hk, li = rec[0], rec[1], rec[2]
            nu = self.ld(rec[3])
            c = self.lits[li] if li >= 0 else 1.0
            k = self.ld(rec[4]) if hk else 1.0
            x1 = self.ld(rec[6])
            w = nu * c * k
            if not second:
                if nf == 1:
                    self.acc(rec[12], w)
                    if hk: self.acc(rec[5], nu * c * x1)
                elif nf == 2:
                    x2 = self.ld(rec[7])
                    self.acc(rec[12], w * x2)
                    self.acc(rec[13], w * x1)
                    if hk: self.acc(rec[5], nu * c * (x1 * x2))
                else:
                    x2, x3 = self.ld(rec[7]), self.ld(rec[8])
                    self.acc(rec[12], w * (x2 * x3))
                    self.acc(rec[13], w * (x1 * x3))
                    self.acc(rec[14], w * (x1 * x2))
                    if hk: self.acc(rec[5], nu * c * (x1 * x2 * x3))
            else:
                d1 = self.ld(rec[9])
                if nf == 1:
                    if hk: self.acc(rec[5], nu * c * d1)
                elif nf == 2:
                    x2, d2 = self.ld(rec[7]), self.ld(rec[10])
                    self.acc(rec[12], w * d2)
                    self.acc(rec[13], w * d1)
                    if hk: self.acc(rec[5], nu * c * (x2 * d1 + x1 * d2))
                else:
                    x2, x3 = self.ld(rec[7]), self.ld(rec[8])
                    d2, d3 = self.ld(rec[10]), self.ld(rec[11])
                    self.acc(rec[12], w * (x3 * d2 + x2 * d3))
                    self.acc(rec[13], w * (x3 * d1 + x1 * d3))
                    self.acc(rec[14], w * (x2 * d1 + x1 * d2))
                    if hk: self.acc(rec[5], nu * c * (x2 * x3 * d1 + x1 * x3 * d2 + x1 * x2 * d3))

    scatter = False  # True: the per-reaction scatter form of the monomial derivatives

    def adj_step(self, dh):
        """Reverse sweep of one step; dh (scalar or per lane) is Dir*H."""
        n = self.ir["nvar"]
        for k in range(n):
            l = self.ld(self.o("LAM", k)).copy()
            for j in range(4):
                self.st(self.o("LK", j * n + k), self.ROS_M[j] * l)
            self.st(self.o("GY", k), l)
        for s in (3, 2, 1, 0):
            for k in range(n):
                self.st(self.o("TMP", k), self.ld(self.o("LK", s * n + k)).copy())
            self.solve_tr()
            if self.scatter:
                self.adj_nu()
            else:
                self.adj_nu_stream()
            if s >= 2:
                self.stage_point(s)
            if self.scatter:
                for k in range(n):
                    self.st(self.o("LP", k), 0.0 * self.ld(self.o("ZERO")))
                self.adj_mono(False, "YNEW" if s >= 2 else "Y")
            else:
                self.adj_mono_gather(False, "YNEW" if s >= 2 else "Y")
            for k in range(n):
                lp, mu = self.ld(self.o("LP", k)), self.ld(self.o("TMP", k))
                self.acc(self.o("GY", k), lp)
                for j in range(s):
                    o = self.o("LK", j * n + k)
                    v = self.ld(o) + (self.ROS_C[s * (s - 1) // 2 + j] / dh) * mu
                    if s >= 2:
                        v = v + self.ROS_A[s * (s - 1) // 2 + j] * lp
                    self.st(o, v)
            for k in range(n):
                self.st(self.o("DIR", k), self.ld(self.o("K", s * n + k)).copy())
            if self.scatter:
                for k in range(n):
                    self.st(self.o("LP", k), 0.0 * self.ld(self.o("ZERO")))
                self.adj_mono(True, "Y")
            else:
                self.adj_mono_gather(True, "Y")
            for k in range(n):
                self.acc(self.o("GY", k), self.ld(self.o("LP", k)))
        for k in range(n):
            self.st(self.o("LAM", k), self.ld(self.o("GY", k)).copy())

    def vec(self, name, k=0):
        n = self.ir["nvar"]
        e = self.E[name] + k * n
        return self.ws[e:e + n]


def _cells(mech, lanes, seed):
    rng = np.random.default_rng(seed)
    nvar, nfix, nreact = mech.nvar, mech.nfix, mech.nreact
    V = 10.0 ** rng.uniform(-2, 10, (lanes, nvar))
    V[:, ::17] = 0.0                      # exact zeros exercise the signed sums
    F = 10.0 ** rng.uniform(15, 19, (lanes, nfix))
    R = 10.0 ** rng.uniform(-20, -8, (lanes, nreact))
    return V, F, R


def _load(m, V, F, R):
    E, n = m.E, V.shape[1]
    m.ws[E["Y"]:E["Y"] + n] = V.T
    m.ws[E["FIX"]:E["FIX"] + F.shape[1]] = F.T
    m.ws[E["RCT"]:E["RCT"] + R.shape[1]] = R.T


def _same(a, b):
    """Bitwise equality of float arrays (NaN-free)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


LANES = 4


@pytest.mark.parametrize("seg", GT.SEGMENTS)
def test_tables_match_oracle_fullchem(mech, oracle_cr, seg):
    _check(mech, oracle_cr, seg, seed=seg)


@pytest.mark.parametrize("seg", GT.SEGMENTS)
@pytest.mark.parametrize("name", ["stress500", "auto1"])
def test_tables_match_oracle_synthetic(name, seg):
    from gcchem_amd import mechanism
    from oracle import Oracle
    m = mechanism.load(name)
    _check(m, Oracle(m.ir, pow_mode="cr"), seg, seed=7)


def _check(mech, orc, seg, seed):
    ir = mech.ir
    n = ir["nvar"]
    V, F, R = _cells(mech, LANES, seed)
    m = TableMachine(ir, LANES, seg)
    _load(m, V, F, R)
    # Fun at Y into FCN0, and at Ynew (= perturbed Y) into K_1
    m.fun(False, m.E["FCN0"] * 512)
    Vn = V * 1.25
    m.ws[m.E["YNEW"]:m.E["YNEW"] + n] = Vn.T
    m.fun(True, (m.E["K"] + n) * 512)
    ghinv = np.array([1.0e3, 3.7, 0.05, 1.0e-6])[:LANES]
    sing = m.jac_lu(ghinv)
    lu_dev = m.ws[m.E["LU"]:m.E["LU"] + ir["nnz"]]
    # right-hand sides for each stage, solved in place
    rhs = np.stack([np.sin(np.arange(n) + s) * 10.0 ** (s - 2) for s in range(4)])
    for s in range(4):
        m.ws[m.E["K"] + s * n:m.E["K"] + (s + 1) * n] = rhs[s][:, None]
        m.solve(s)
    diag = np.asarray(ir["diag"])
    for c in range(LANES):
        vdot, *_ = orc.fun_split(V[c], F[c], R[c])
        assert _same(m.vec("FCN0")[:, c], vdot)
        jvs = orc.jac_sp(V[c], F[c], R[c])
        g = -jvs
        g[diag] = g[diag] + ghinv[c]
        assert bool(sing[c]) == bool(np.any(np.abs(g[diag]) < 2.2250738585072014e-308))
        lu, ier = orc.decomp(g)
        assert ier == 0
        assert _same(lu_dev[:, c], lu)
        for s in range(4):
            assert _same(m.vec("K", s)[:, c], orc.solve(lu, rhs[s]))
    # Fun at Ynew was written into K_1 before the solves overwrote it: redo it
    m.fun(True, (m.E["K"] + n) * 512)
    for c in range(LANES):
        vdot_n, *_ = orc.fun_split(Vn[c], F[c], R[c])
        assert _same(m.vec("K", 1)[:, c], vdot_n)


@pytest.mark.parametrize("seg", (16, 32))
@pytest.mark.parametrize("name", ["fullchem", "stress500"])
def test_masked_tables_match_oracle(name, seg):
    """The masked kernels' guards, interpreted over the tables with random
    per-lane auto-reduce flags (DO_FUN, DO_JVS, DO_SLV), equal the oracle's
    guarded Fun_SPLIT (P/D starting at zero), Jac_SP (Jac0 starting at
    zero), KppDecomp and KppSolve bit for bit."""
    from gcchem_amd import mechanism
    from oracle import Oracle
    mech = mechanism.load(name)
    orc = Oracle(mech.ir, pow_mode="cr")
    ir = mech.ir
    n, nnz = ir["nvar"], ir["nnz"]
    V, F, R = _cells(mech, LANES, 11)
    rng = np.random.default_rng(12)
    m = TableMachine(ir, LANES, seg)
    m.do_fun = rng.uniform(size=(LANES, n)) < 0.7
    m.do_jvs = rng.uniform(size=(LANES, nnz)) < 0.8
    m.do_slv = rng.uniform(size=(LANES, n)) < 0.7
    m.do_fun[0], m.do_jvs[0], m.do_slv[0] = True, True, True    # lane 0: every flag on
    _load(m, V, F, R)
    m.fun(False, m.E["FCN0"] * 512)
    ghinv = np.array([1.0e3, 3.7, 0.05, 1.0e-6])[:LANES]
    m.jac_lu(ghinv)
    lu_dev = m.ws[m.E["LU"]:m.E["LU"] + nnz]
    rhs = np.stack([np.sin(np.arange(n) + s) * 10.0 ** (s - 2) for s in range(4)])
    for s in range(4):
        m.ws[m.E["K"] + s * n:m.E["K"] + (s + 1) * n] = rhs[s][:, None]
        m.solve(s)
    diag = np.asarray(ir["diag"])
    W = Oracle.mask_words
    for c in range(LANES):
        vdot, *_ = orc.fun_split(V[c], F[c], R[c], do_fun=W(m.do_fun[c]))
        assert _same(m.vec("FCN0")[:, c], vdot)
        if c == 0:
            assert _same(vdot, orc.fun_split(V[c], F[c], R[c])[0])
        assert np.all(vdot[~m.do_fun[c]] == 0.0)
        jvs = orc.jac_sp(V[c], F[c], R[c], do_jvs=W(m.do_jvs[c]))
        g = -jvs
        g[diag] = g[diag] + ghinv[c]
        lu, ier = orc.decomp(g)
        assert ier == 0 and _same(lu_dev[:, c], lu)
        for s in range(4):
            x = orc.solve(lu, rhs[s], do_slv=W(m.do_slv[c]))
            assert _same(m.vec("K", s)[:, c], x)
            assert _same(x[~m.do_slv[c]], rhs[s][~m.do_slv[c]])


def test_padding_is_exact_noop():
    """-0.0 - 0.0*0.0 keeps the sign, and an LDS padding update only touches
    the spare slot."""
    x = np.array([-0.0, 0.0, 1e-300, -3.5])
    y = x - np.float64(0.0) * np.float64(0.0)
    assert _same(y, x)


@pytest.mark.parametrize("name", ["fullchem", "stress500"])
def test_adjoint_step_tables_match_reference(name, checkforms):
    """One adjoint step through the generated tables (the kernel's loops)
    equals the NumPy reference adjoint of the same step, at the start states
    and step sizes of accepted steps of seeded cells."""
    from adjoint_ref import Monomials, adjoint_cell, stoich_matrix
    from gcchem_amd import inputs, mechanism
    from oracle import Oracle
    mech = mechanism.load(name)
    orc = Oracle(mech.ir, pow_mode="cr")
    ir = mech.ir
    n, nf, nr = ir["nvar"], ir["nfix"], ir["nreact"]
    d = inputs.random_cells(mech, n=3, seed=5) if name == "fullchem" else inputs.stress_cells(mech, 3, seed=5)
    if name != "fullchem":
        d["Tend"] = 20.0
    lanes = 3
    V, F, R, Hs = np.zeros((lanes, n)), d["C"][:, n:].copy(), d["RCONST"].copy(), []
    for c in range(lanes):
        *_, hrec = orc.solve_cell_rec(d["Tstart"], d["Tend"], d["C"][c], d["RCONST"][c],
                                      d["ATOL"], d["RTOL"], d["ICNTRL"][c], d["RCNTRL"][c])
        k = len(hrec) // 2
        V[c] = orc.replay(d["C"][c][:n], F[c], R[c], hrec[:k])
        Hs.append(hrec[k])
    Hs = np.array(Hs)
    rng = np.random.default_rng(21)
    lam = rng.standard_normal((lanes, n))
    lamF = rng.standard_normal((lanes, nf))
    m = TableMachine(ir, lanes, 16)
    _load(m, V, F, R)
    E = m.E
    m.ws[E["LAM"]:E["LAM"] + n] = lam.T
    m.ws[E["GF"]:E["GF"] + nf] = lamF.T
    m.ws[E["GK"]:E["GK"] + nr] = 0.0
    m.fun(False, m.o("FCN0"))
    assert not m.jac_lu(1.0 / (Hs * 0.5)).any()
    m.stages(Hs)
    m.adj_step(Hs)
    gV = m.ws[E["LAM"]:E["LAM"] + n].T
    gF = m.ws[E["GF"]:E["GF"] + nf].T
    gK = m.ws[E["GK"]:E["GK"] + nr].T
    # reference stoichiometry: the reference's CalcStoichNum for fullchem, the
    # generator's for the synthetic mechanism (checked separately below)
    if name == "fullchem":
        S = stoich_matrix(checkforms, n, nr)
    else:
        S = np.zeros((n, nr))
        for i, row in enumerate(GT.Tables(ir).stoichiometry()):
            for r, c in row.items():
                S[i, r] = c
    mono = Monomials(ir)
    # Same sums in different orders: every entry to 1e-9 of its vector's
    # scale and almost all to 1e-10 relative; a wrong offset or a dropped term
    # breaks both by orders of magnitude.
    for c in range(lanes):
        rV, rF, rK = adjoint_cell(orc, mono, S, V[c], F[c], R[c], [Hs[c]], lam[c], lamF[c])
        for got, ref in ((gV[c], rV), (gF[c], rF), (gK[c], rK)):
            scale = np.max(np.abs(ref)) + 1e-300
            assert np.max(np.abs(got - ref)) <= 1e-9 * scale
            rel = np.abs(got - ref) / (np.abs(ref) + 1e-300 * scale)
            assert np.mean(rel <= 1e-10) >= 0.95 or ref.size < 10


@pytest.mark.parametrize("name", ["fullchem", "stress500"])
def test_adjoint_gather_equals_scatter(name):
    """The kernel's gather form of the monomial derivatives (W records, k
    records, per-target sums in reaction order) gives the per-reaction scatter
    form's gradients bit for bit, on one adjoint step at random states."""
    from gcchem_amd import mechanism
    mech = mechanism.load(name)
    ir = mech.ir
    n, nf, nr = ir["nvar"], ir["nfix"], ir["nreact"]
    lanes = 4
    V, F, R = _cells(mech, lanes, 3)
    out = []
    for scatter in (True, False):
        m = TableMachine(ir, lanes, 16)
        m.scatter = scatter
        _load(m, V, F, R)
        E = m.E
        rs = np.random.default_rng(9)
        m.ws[E["LAM"]:E["LAM"] + n] = rs.standard_normal((n, lanes))
        m.ws[E["GF"]:E["GF"] + nf] = rs.standard_normal((nf, lanes))
        m.ws[E["GK"]:E["GK"] + nr] = 0.0
        for s in range(4):   # stage vectors and factors of a nearby matrix
            m.ws[E["K"] + s * n:E["K"] + (s + 1) * n] = rs.standard_normal((n, lanes)) * V.T * 1e-3
        m.fun(False, m.o("FCN0"))
        assert not m.jac_lu(np.full(lanes, 1e3)).any()
        m.adj_step(np.full(lanes, 0.02))
        keep = np.r_[E["LAM"]:E["NU"], E["GK"]:E["ADJ_TOTAL"]]   # NU is scratch here
        out.append(m.ws[keep].copy())
    a, b = out
    assert np.isfinite(a).all()
    assert np.array_equal(a.view(np.int64), b.view(np.int64))


def test_stoichiometry_matches_calcstoichnum(mech, checkforms):
    """The net stoichiometry the generator derives from the Fun_SPLIT sums is
    the reference's CalcStoichNum, entry for entry."""
    S = GT.Tables(mech.ir).stoichiometry()
    mine = {(i, r): c for i, row in enumerate(S) for r, c in row.items() if c != 0.0}
    ref = {(i, r): float(v) for i, r, v in checkforms["stoich"]}
    assert mine == ref
