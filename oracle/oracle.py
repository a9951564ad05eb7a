"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle is the plain-C restatement of the reference Rodas3 path in
``rodas3_oracle.c``.  Only the tests, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` load it, as the checker; the product library
never does.  The mechanism comes from the mechanism IR JSON (the same data the
code generator reads), flattened here into the sum programs of
``rodas3_oracle.h``.
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

KIND = {"lit": 0, "A": 1, "RCT": 2, "V": 3, "F": 4, "B": 5}


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile (gcc, -O2 -ffp-contract=off)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(HERE), "-B" if force else "all"],
                       check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


class _SumProg(ctypes.Structure):
    _fields_ = [("nout", ctypes.c_int),
                ("out_off", ctypes.c_void_p),
                ("term_sign", ctypes.c_void_p),
                ("term_off", ctypes.c_void_p),
                ("fac_kind", ctypes.c_void_p),
                ("fac_idx", ctypes.c_void_p)]


class _Mech(ctypes.Structure):
    _fields_ = [("nvar", ctypes.c_int), ("nfix", ctypes.c_int),
                ("nreact", ctypes.c_int), ("nnz", ctypes.c_int),
                ("nb", ctypes.c_int),
                ("lit", ctypes.c_void_p),
                ("rate", _SumProg), ("prod", _SumProg), ("dest", _SumProg),
                ("jacb", _SumProg), ("jvs", _SumProg),
                ("crow", ctypes.c_void_p), ("icol", ctypes.c_void_p),
                ("diag", ctypes.c_void_p),
                ("pow_mode", ctypes.c_int)]


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """CPU oracle for one mechanism IR."""

    def __init__(self, ir: dict | str | os.PathLike, pow_mode: str = "libm"):
        if not isinstance(ir, dict):
            ir = json.loads(Path(ir).read_text())
        self.ir = ir
        self.nvar, self.nfix = ir["nvar"], ir["nfix"]
        self.nspec, self.nreact = ir["nvar"] + ir["nfix"], ir["nreact"]
        self.nnz, self.nb = ir["nnz"], ir["nb"]
        self._keep = []
        self._lits: list[float] = []
        self._lit_index: dict[str, int] = {}
        self.lib = ctypes.CDLL(str(build()))
        m = _Mech()
        m.nvar, m.nfix, m.nreact, m.nnz, m.nb = (self.nvar, self.nfix, self.nreact,
                                                  self.nnz, self.nb)
        m.rate = self._prog([[[1, f]] for f in ir["rate"]])
        m.prod = self._prog(ir["prod"])
        m.dest = self._prog(ir["dest"])
        m.jacb = self._prog([[] if f is None else [[1, f]] for f in ir["jac_b"]])
        m.jvs = self._prog([s for _, _, s in ir["jvs"]])
        lit = _f64(self._lits if self._lits else [0.0])
        self._keep.append(lit)
        m.lit = _ptr(lit)
        for name in ("crow", "icol", "diag"):
            a = _i32(ir[name])
            self._keep.append(a)
            setattr(m, name, _ptr(a))
        m.pow_mode = {"libm": 0, "cr": 1}[pow_mode]
        self.mech = m
        self._declare()

    # -- flattening -------------------------------------------------------
    def _lit(self, text: str) -> int:
        if text not in self._lit_index:
            self._lit_index[text] = len(self._lits)
            self._lits.append(float(text.replace("d", "e").replace("D", "e")))
        return self._lit_index[text]

    def _prog(self, sums) -> _SumProg:
        out_off, term_sign, term_off, fac_kind, fac_idx = [0], [], [0], [], []
        for s in sums:
            for sign, prod in s:
                term_sign.append(sign)
                for kind, idx in prod:
                    fac_kind.append(KIND[kind])
                    fac_idx.append(self._lit(idx) if kind == "lit" else idx)
                term_off.append(len(fac_kind))
            out_off.append(len(term_sign))
        arrs = [_i32(x if x else [0]) for x in (out_off, term_sign, term_off,
                                                 fac_kind, fac_idx)]
        self._keep.extend(arrs)
        p = _SumProg()
        p.nout = len(sums)
        p.out_off, p.term_sign, p.term_off, p.fac_kind, p.fac_idx = (_ptr(a) for a in arrs)
        return p

    def _declare(self):
        L, P = self.lib, ctypes.c_void_p
        D, I = ctypes.c_double, ctypes.c_int
        L.or_work_doubles.restype = ctypes.c_long
        L.or_work_doubles.argtypes = [P]
        L.or_fun_split.argtypes = [P] * 8
        L.or_jac_sp.argtypes = [P] * 6
        L.or_decomp.argtypes = [P, P, P, P]
        L.or_solve.argtypes = [P, P, P]
        L.or_solve_tr.argtypes = [P, P, P, P]
        L.or_replay_cell.argtypes = [P, P, P, P, D, I, P, I, P, P]
        L.or_solve_cell_rec.argtypes = [P, P, P, P, D, D, P, P, P, P, P, P, P, P, P, I, P]
        L.or_integrate_batch.restype = D
        L.or_integrate_batch.argtypes = [P, I, D, D, P, P, P, P, P, P, P, P, P, P, I]
        L.or_fun_split_m.argtypes = [P] * 9
        L.or_jac_sp_m.argtypes = [P] * 7
        L.or_solve_m.argtypes = [P] * 4
        L.or_integrate_batch_m.restype = D
        L.or_integrate_batch_m.argtypes = [P, I, D, D] + [P] * 10 + [I, P, P, P]

    @property
    def _m(self):
        return ctypes.byref(self.mech)

    # -- kernels ------------------------------------------------------------
    # Masks (auto-reduce flags): uint32 words, flag k in bit k%32 of word k//32
    # (see mask_words); None = every flag on.  A guarded output whose flag is
    # off keeps the value passed in (P/D for Fun, JVS for Jac; zeros if None).
    @staticmethod
    def mask_words(flags) -> np.ndarray:
        """Pack booleans flags[..., n] into uint32 words[..., ceil(n/32)]."""
        f = np.asarray(flags, bool)
        n = f.shape[-1]
        pad = np.zeros(f.shape[:-1] + ((-n) % 32,), bool)
        bits = np.concatenate([f, pad], axis=-1).reshape(f.shape[:-1] + (-1, 32))
        return (bits.astype(np.uint64) << np.arange(32, dtype=np.uint64)).sum(-1).astype(np.uint32)

    def fun_split(self, V, F, RCT, do_fun=None, P0=None, D0=None):
        V, F, RCT = _f64(V), _f64(F), _f64(RCT)
        vdot = np.zeros(self.nvar)
        p = np.zeros(self.nvar) if P0 is None else _f64(P0).copy()
        d = np.zeros(self.nvar) if D0 is None else _f64(D0).copy()
        a = np.zeros(self.nreact)
        if do_fun is None:
            self.lib.or_fun_split(self._m, _ptr(V), _ptr(F), _ptr(RCT), _ptr(vdot),
                                  _ptr(p), _ptr(d), _ptr(a))
        else:
            mk = np.ascontiguousarray(do_fun, np.uint32)
            self.lib.or_fun_split_m(self._m, _ptr(V), _ptr(F), _ptr(RCT), _ptr(vdot),
                                    _ptr(p), _ptr(d), _ptr(a), _ptr(mk))
        return vdot, p, d, a

    def jac_sp(self, V, F, RCT, do_jvs=None, JVS0=None):
        V, F, RCT = _f64(V), _f64(F), _f64(RCT)
        jvs = np.zeros(self.nnz) if JVS0 is None else _f64(JVS0).copy()
        b = np.zeros(max(self.nb, 1))
        if do_jvs is None:
            self.lib.or_jac_sp(self._m, _ptr(V), _ptr(F), _ptr(RCT), _ptr(jvs), _ptr(b))
        else:
            mk = np.ascontiguousarray(do_jvs, np.uint32)
            self.lib.or_jac_sp_m(self._m, _ptr(V), _ptr(F), _ptr(RCT), _ptr(jvs), _ptr(b),
                                 _ptr(mk))
        return jvs

    def decomp(self, jvs):
        lu = _f64(jvs).copy()
        w = np.zeros(self.nvar)
        ier = ctypes.c_int(0)
        self.lib.or_decomp(self._m, _ptr(lu), _ptr(w), ctypes.byref(ier))
        return lu, ier.value

    def solve(self, lu, x, do_slv=None):
        lu = _f64(lu)
        x = _f64(x).copy()
        if do_slv is None:
            self.lib.or_solve(self._m, _ptr(lu), _ptr(x))
        else:
            mk = np.ascontiguousarray(do_slv, np.uint32)
            self.lib.or_solve_m(self._m, _ptr(lu), _ptr(x), _ptr(mk))
        return x

    def solve_tr(self, lu, x):
        lu, x = _f64(lu), _f64(x)
        xx = np.zeros(self.nvar)
        self.lib.or_solve_tr(self._m, _ptr(lu), _ptr(x), _ptr(xx))
        return xx

    def solve_cell_rec(self, Tstart, Tend, C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL, maxrec=100000):
        """One cell (C[NSPEC], RCONST[NREACT], ICNTRL/RCNTRL[20]) with the
        accepted step sizes recorded.  Returns (C, ISTAT, RSTAT, IERR, Hs)."""
        C = _f64(C).copy()
        RCONST, ATOL, RTOL = _f64(RCONST), _f64(ATOL), _f64(RTOL)
        ICNTRL, RCNTRL = _i32(ICNTRL), _f64(RCNTRL)
        ISTAT, RSTAT = np.zeros(8, np.int32), np.zeros(3)
        IERR, nrec = ctypes.c_int(0), ctypes.c_int(0)
        Hrec = np.zeros(maxrec)
        work = np.zeros(self.lib.or_work_doubles(self._m))
        F = C[self.nvar:].copy()
        self.lib.or_solve_cell_rec(self._m, _ptr(C), _ptr(F), _ptr(RCONST), ctypes.c_double(Tstart),
                                   ctypes.c_double(Tend), _ptr(ATOL), _ptr(RTOL), _ptr(ICNTRL),
                                   _ptr(RCNTRL), _ptr(ISTAT), _ptr(RSTAT), ctypes.byref(IERR),
                                   _ptr(work), _ptr(Hrec), int(maxrec), ctypes.byref(nrec))
        return C, ISTAT, RSTAT, IERR.value, Hrec[:min(nrec.value, maxrec)].copy()

    def replay(self, V, F, RCT, Hs, Dir=1.0, clamp=False):
        """Final VAR concentrations after the accepted steps of sizes Hs
        (frozen step-size control).  Raises if a step's matrix is singular."""
        Y = _f64(V).copy()
        F, RCT, Hs = _f64(F), _f64(RCT), _f64(Hs)
        work = np.zeros(self.lib.or_work_doubles(self._m))
        ier = ctypes.c_int(0)
        self.lib.or_replay_cell(self._m, _ptr(Y), _ptr(F), _ptr(RCT), ctypes.c_double(Dir),
                                int(Hs.size), _ptr(Hs), int(bool(clamp)), _ptr(work),
                                ctypes.byref(ier))
        if ier.value:
            raise RuntimeError(f"replay: singular matrix at step {ier.value - 1}")
        return Y

    def integrate_batch(self, Tstart, Tend, C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL,
                        Perm=None, nthreads=0, masks=None):
        """gc_gpu_integrate_batch semantics.  Arrays are cell-major:
        C[nCell, NSPEC] (updated in place on a copy, returned), RCONST[nCell,
        NREACT], ICNTRL/RCNTRL[nCell, 20], Perm 1-based (identity if None).
        masks: None or (doFun[nCell, ceil(NVAR/32)], doJvs[nCell,
        ceil(LU_NONZERO/32)], doSlv[nCell, ceil(NVAR/32)]) uint32 words, any
        of them None for every flag on.
        Returns (C, ISTAT[nCell,8], RSTAT[nCell,3], IERR[nCell], seconds)."""
        C = _f64(C).copy()
        n = C.shape[0]
        RCONST, ATOL, RTOL = _f64(RCONST), _f64(ATOL), _f64(RTOL)
        ICNTRL, RCNTRL = _i32(ICNTRL), _f64(RCNTRL)
        Perm = _i32(np.arange(1, n + 1) if Perm is None else Perm)
        assert C.shape == (n, self.nspec) and RCONST.shape == (n, self.nreact)
        assert ICNTRL.shape == (n, 20) and RCNTRL.shape == (n, 20) and Perm.shape == (n,)
        ISTAT = np.zeros((n, 8), np.int32)
        RSTAT = np.zeros((n, 3))
        IERR = np.zeros(n, np.int32)
        if masks is None:
            secs = self.lib.or_integrate_batch(
                self._m, n, float(Tstart), float(Tend), _ptr(C), _ptr(RCONST), _ptr(ATOL),
                _ptr(RTOL), _ptr(ICNTRL), _ptr(RCNTRL), _ptr(Perm), _ptr(ISTAT),
                _ptr(RSTAT), _ptr(IERR), int(nthreads))
        else:
            wv, wj = (self.nvar + 31) // 32, (self.nnz + 31) // 32
            mk = []
            for a, w in zip(masks, (wv, wj, wv)):
                if a is None:
                    mk.append(None)
                    continue
                a = np.ascontiguousarray(a, np.uint32)
                assert a.shape == (n, w), (a.shape, (n, w))
                mk.append(a)
            secs = self.lib.or_integrate_batch_m(
                self._m, n, float(Tstart), float(Tend), _ptr(C), _ptr(RCONST), _ptr(ATOL),
                _ptr(RTOL), _ptr(ICNTRL), _ptr(RCNTRL), _ptr(Perm), _ptr(ISTAT),
                _ptr(RSTAT), _ptr(IERR), int(nthreads),
                *[None if a is None else _ptr(a) for a in mk])
        return C, ISTAT, RSTAT, IERR, secs
