"""Mechanism IR -> chunked straight-line FP64 HIP for gfx950.

Emits, for one mechanism, the per-cell kernels of the Rosenbrock Rodas3 solve
as ``__device__`` functions working on one cell per lane, with all cell state
in structure-of-arrays (SoA) HBM layout: element ``k`` of the cell held by
slot ``s`` lives at ``base[k*S + s]``.  Every access is one coalesced
512-byte wave access ``global_load_dwordx2 v, voff, s[base_k]`` (uniform
element base in SGPRs, one per-lane offset register shared by all accesses).

Each routine is cut into *chunks* of bounded size.  A chunk is a
``__noinline__`` function with a chunk-local value cache: the first use of a
memory operand in a chunk loads it into a local, later uses reuse the local,
and values produced in the chunk stay in locals for the rest of the chunk.
Values that cross chunks go through the SoA workspace.  This bounds both the
compiler's work per function (compile time is linear in mechanism size) and
the live-register set, so register use and spills are a choice of the chunk
size rather than an accident of scheduling.

Functions (paths relative to the reference checkout):

* ``fun``      Fun_SPLIT  src/geos-chem/KPP/fullchem/gckpp_Function.F90:2170-5060
               pass 1 forms A(r) into a scratch SoA array, pass 2 forms
               P_VAR(i), D_VAR(i) and Vdot(i) = P_VAR(i) - D_VAR(i)*V(i).
* ``jac_prep`` Jac_SP     src/geos-chem/KPP/fullchem/gckpp_Jacobian.F90:47-20887
               fused with the matrix preparation of Integrate_Cell_S
               (Ghimj = -Jac0, diagonal += ghinv:
               src/geos-chem/KPP/fullchem/gckpp_BatchIntegrator.F90:573-577,784-795)
               and KppDecomp's singularity test
               (src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:61-64).
               Generated column by column so each column's B temporaries die
               with the column.
* ``decomp``   KppDecomp  src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:46-84
               IKJ, unrolled from the pattern, each row held in registers;
               U rows finished in the same chunk are reused from registers.
* ``solve``    KppSolve   src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:645-2311
* ``solve_tr`` KppSolveTR src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:2327-3701

Bit-faithfulness: every product and sum is a sequence of binary operations in
the reference's left-to-right order; literals are the nearest double of their
decimal text (printed with ``repr`` so the compiler reads back that double);
the output is compiled with ``-ffp-contract=off``.  Only the schedule (which
independent output is computed when) is chosen here.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path


def _lit(text: str) -> str:
    v = float(text.replace("d", "e").replace("D", "e"))
    r = repr(v)
    if "e" not in r and "." not in r and "inf" not in r and "nan" not in r:
        r += ".0"
    return r


PRELUDE = r"""
// ---- SoA access helpers (uniform element base in SGPRs + per-lane offset) ----
typedef __attribute__((address_space(1))) double gdbl_t;
typedef __attribute__((address_space(1))) char gchar_t;

#ifndef GC_SOA_DEFINED
#define GC_SOA_DEFINED
__device__ __forceinline__ gchar_t* gc_uniform_ptr(const void* p) {
  unsigned long long v = (unsigned long long)p;
  unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (gchar_t*)(((unsigned long long)hi << 32) | lo);
}
// One SoA array: element k of the calling lane's slot.  The pointer and the
// stride are wave-uniform (they are made so with readfirstlane, which is exact
// because every caller passes the same value in every lane).
struct GcSoa {
  gchar_t* base;
  unsigned long long sb;
  unsigned off;
  __device__ __forceinline__ GcSoa(const void* p, unsigned S, unsigned slot)
      : base(gc_uniform_ptr(p)),
        sb((unsigned long long)__builtin_amdgcn_readfirstlane(S) * 8ull),
        off(slot * 8u) {}
  __device__ __forceinline__ double ld(int k) const {
    return *(gdbl_t*)(base + (unsigned long long)k * sb + off);
  }
  __device__ __forceinline__ void st(int k, double v) const {
    *(gdbl_t*)(base + (unsigned long long)k * sb + off) = v;
  }
};
#endif
"""


class Chunker:
    """Emit a routine as a sequence of chunk functions.

    ``arrays`` maps an array tag to (C parameter name, SoA object name).
    Within a chunk, ``get(tag, k)`` returns a local holding element k (loading
    it on first use) and ``put(tag, k, expr, store)`` defines a new value."""

    def __init__(self, ns, name, params, arrays, ret=None, budget=250):
        self.ns, self.name, self.params, self.arrays = ns, name, params, arrays
        self.ret = ret            # None or ("int", "sing") for an OR-ed flag
        self.budget = budget
        self.chunks: list[list[str]] = []
        self.body: list[str] = []
        self.cache: dict[tuple, str] = {}
        self.cost = 0
        self.nlocal = 0

    # -- chunk management ---------------------------------------------------
    def boundary(self):
        """A point where a new chunk may start (all values are in memory)."""
        if self.cost >= self.budget:
            self._close()

    def _close(self):
        if self.body:
            self.chunks.append(self.body)
        self.body, self.cache, self.cost = [], {}, 0

    def line(self, s: str, cost: int = 0):
        self.body.append("  " + s)
        self.cost += cost

    def fresh(self, prefix="t"):
        self.nlocal += 1
        return f"{prefix}{self.nlocal}"

    def get(self, tag, k) -> str:
        key = (tag, k)
        if key not in self.cache:
            v = self.fresh("m")
            self.line(f"const double {v} = {self.arrays[tag][1]}.ld({k});", 1)
            self.cache[key] = v
        return self.cache[key]

    def put(self, tag, k, expr: str, store=True) -> str:
        v = self.fresh("v")
        self.line(f"const double {v} = {expr};")
        if store:
            self.line(f"{self.arrays[tag][1]}.st({k}, {v});", 1)
        self.cache[(tag, k)] = v
        return v

    # -- output ---------------------------------------------------------------
    def emit(self) -> str:
        self._close()
        out = []
        fname = f"{self.ns}_{self.name}"
        rtype = self.ret[0] if self.ret else "void"
        soas = ", ".join(f"{obj}({p}, S, slot)" for p, obj in self.arrays.values())
        for n, body in enumerate(self.chunks):
            out.append(f"__device__ __noinline__ {rtype} {fname}_c{n}({self.params}, "
                       f"unsigned S, unsigned slot) {{")
            out.append(f"  const GcSoa {soas};")
            if self.ret:
                out.append(f"  {self.ret[0]} {self.ret[1]} = 0;")
            out.extend(body)
            if self.ret:
                out.append(f"  return {self.ret[1]};")
            out.append("}")
        args = ", ".join(a.split()[-1].lstrip("*") for a in self.params.split(","))
        out.append(f"__device__ __forceinline__ {rtype} {fname}({self.params}, "
                   f"unsigned S, unsigned slot) {{")
        if self.ret:
            out.append(f"  {self.ret[0]} r = 0;")
            for n in range(len(self.chunks)):
                out.append(f"  r |= {fname}_c{n}({args}, S, slot);")
            out.append("  return r;")
        else:
            for n in range(len(self.chunks)):
                out.append(f"  {fname}_c{n}({args}, S, slot);")
        out.append("}")
        return "\n".join(out) + "\n"


def _product(prod, val) -> tuple[str, int]:
    """Left-to-right product; returns (expr, number of multiplies)."""
    expr = val(prod[0])
    for f in prod[1:]:
        expr = f"({expr} * {val(f)})"
    return expr, len(prod) - 1


def _sum_expr(terms, val) -> tuple[str, int]:
    """Sum in reference order: first term (negated if signed), then +/-."""
    if not terms:
        return "0.0", 0
    s0, p0 = terms[0]
    e, n = _product(p0, val)
    expr = f"(-{e})" if s0 < 0 else e
    for s, p in terms[1:]:
        pe, pn = _product(p, val)
        expr = f"({expr} {'-' if s < 0 else '+'} {pe})"
        n += pn + 1
    return expr, n


# ---------------------------------------------------------------------------


def gen_fun(ir, ns: str, budget: int) -> str:
    nvar, nreact = ir["nvar"], ir["nreact"]
    rate, prod, dest = ir["rate"], ir["prod"], ir["dest"]
    arrays = {"V": ("V", "Vv"), "F": ("F", "Fv"), "RCT": ("RCT", "Rv"),
              "A": ("Aw", "Av"), "O": ("Vdot", "Ov")}
    params = "const double* V, const double* F, const double* RCT, double* Aw, double* Vdot"
    ch = Chunker(ns, "fun", params, arrays, budget=budget)

    def val(f):
        kind, idx = f
        return _lit(idx) if kind == "lit" else ch.get(kind, idx)

    # pass 1: rate monomials A(r) (gckpp_Function.F90, "Computation of equation rates")
    for r in range(nreact):
        e, n = _product(rate[r], val)
        ch.put("A", r, e)
        ch.cost += n
        ch.boundary()
    ch._close()
    # pass 2: P_VAR, D_VAR and Vdot = P_VAR - D_VAR*V
    for i in range(nvar):
        pe, pn = _sum_expr(prod[i], val)
        de, dn = _sum_expr(dest[i], val)
        p = ch.fresh("p")
        d = ch.fresh("d")
        ch.line(f"const double {p} = {pe};", pn)
        ch.line(f"const double {d} = {de};", dn)
        ch.line(f"Ov.st({i}, {p} - {d} * {ch.get('V', i)});", 3)
        ch.boundary()
    return ("// Fun_SPLIT: src/geos-chem/KPP/fullchem/gckpp_Function.F90:2170-5060\n"
            + ch.emit())


def gen_jac_prep(ir, ns: str, budget: int) -> str:
    jvs, jb = ir["jvs"], ir["jac_b"]
    diag = set(ir["diag"])
    cols: dict[int, list[int]] = {}
    for k, (i, j, _) in enumerate(jvs):
        cols.setdefault(j, []).append(k)
    arrays = {"V": ("V", "Vv"), "F": ("F", "Fv"), "RCT": ("RCT", "Rv"), "LU": ("LU", "Lv")}
    params = "const double* V, const double* F, const double* RCT, double* LU, double ghinv"
    ch = Chunker(ns, "jac_prep", params, arrays, ret=("int", "sing"), budget=budget)
    bloc: dict[int, str] = {}

    def val(f):
        kind, idx = f
        if kind == "lit":
            return _lit(idx)
        if kind == "B":
            return bloc[idx]
        return ch.get(kind, idx)

    for j in sorted(cols):
        ks = cols[j]
        bs = sorted({f[1] for k in ks for _, p in jvs[k][2] for f in p if f[0] == "B"})
        bloc.clear()
        for b in bs:
            e, n = _product(jb[b], val)
            v = ch.fresh("b")
            ch.line(f"const double {v} = {e};", n)
            bloc[b] = v
        for k in ks:
            e, n = _sum_expr(jvs[k][2], val)
            v = ch.fresh("j")
            ch.line(f"const double {v} = {e};", n)
            if k in diag:
                g = ch.fresh("g")
                ch.line(f"const double {g} = (-{v}) + ghinv;", 1)
                ch.line(f"sing |= (__builtin_fabs({g}) < 2.2250738585072014e-308) ? 1 : 0;")
                ch.line(f"Lv.st({k}, {g});", 1)
            else:
                ch.line(f"Lv.st({k}, -{v});", 1)
        ch.boundary()
    return ("// Jac_SP: src/geos-chem/KPP/fullchem/gckpp_Jacobian.F90:47-20887, stored as\n"
            "// Ghimj = -Jac0 with ghinv on the diagonal (gckpp_BatchIntegrator.F90:573-577).\n"
            "// Nonzero return: a diagonal entry is below TINY.  KppDecomp tests each row's\n"
            "// diagonal before that row is eliminated (gckpp_LinearAlgebra.F90:61-64), which\n"
            "// is the value stored here, so the test is done at store time.\n"
            + ch.emit())


def gen_decomp(ir, ns: str, budget: int) -> str:
    nvar = ir["nvar"]
    crow, icol, diag = ir["crow"], ir["icol"], ir["diag"]
    arrays = {"LU": ("LU", "Lv")}
    ch = Chunker(ns, "decomp", "double* LU", arrays, budget=budget)
    for k in range(nvar):
        if diag[k] == crow[k]:
            continue  # no L entries: the row is unchanged
        w = {icol[kk]: ch.get("LU", kk) for kk in range(crow[k], crow[k + 1])}
        for kk in range(crow[k], diag[k]):
            j = icol[kk]
            a = ch.fresh("a")
            ch.line(f"const double {a} = -{w[j]} / {ch.get('LU', diag[j])};", 2)
            nw = ch.fresh("w")
            ch.line(f"const double {nw} = -{a};")
            w[j] = nw
            for jj in range(diag[j] + 1, crow[j + 1]):
                c = icol[jj]
                nw = ch.fresh("w")
                ch.line(f"const double {nw} = {w[c]} + {a} * {ch.get('LU', jj)};", 2)
                w[c] = nw
        for kk in range(crow[k], crow[k + 1]):
            ch.line(f"Lv.st({kk}, {w[icol[kk]]});", 1)
            ch.cache[("LU", kk)] = w[icol[kk]]
        ch.boundary()
    return ("// KppDecomp: src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:46-84 "
            "(IKJ, unpivoted)\n" + ch.emit())


def gen_solve(ir, ns: str, budget: int) -> str:
    nvar = ir["nvar"]
    crow, icol, diag = ir["crow"], ir["icol"], ir["diag"]
    arrays = {"LU": ("LU", "Lv"), "X": ("X", "Xv")}
    ch = Chunker(ns, "solve", "const double* LU, double* X", arrays, budget=budget)
    # forward substitution with the unit-diagonal L, rows in order
    for i in range(nvar):
        if diag[i] == crow[i]:
            continue
        e = ch.get("X", i)
        for j in range(crow[i], diag[i]):
            e = f"({e} - {ch.get('LU', j)} * {ch.get('X', icol[j])})"
        ch.put("X", i, e)
        ch.cost += 2 * (diag[i] - crow[i])
        ch.boundary()
    # backward substitution with U, rows NVAR..1, one divide per row
    for i in reversed(range(nvar)):
        e = ch.get("X", i)
        for j in range(diag[i] + 1, crow[i + 1]):
            e = f"({e} - {ch.get('LU', j)} * {ch.get('X', icol[j])})"
        ch.put("X", i, f"{e} / {ch.get('LU', diag[i])}")
        ch.cost += 2 * (crow[i + 1] - diag[i])
        ch.boundary()
    return "// KppSolve: src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:645-2311\n" + ch.emit()


def gen_solve_tr(ir, ns: str, budget: int) -> str:
    nvar = ir["nvar"]
    crow, icol, diag = ir["crow"], ir["icol"], ir["diag"]
    ucol = [[] for _ in range(nvar)]  # U(j,i) for column i: (k, j), ascending j
    lcol = [[] for _ in range(nvar)]  # L(j,i) for column i: (k, j), ascending j
    for j in range(nvar):
        for k in range(diag[j] + 1, crow[j + 1]):
            ucol[icol[k]].append((k, j))
        for k in range(crow[j], diag[j]):
            lcol[icol[k]].append((k, j))
    arrays = {"LU": ("LU", "Lv"), "X": ("X", "Xv"), "XX": ("XX", "Yv")}
    ch = Chunker(ns, "solve_tr", "const double* LU, const double* X, double* XX",
                 arrays, budget=budget)
    for i in range(nvar):
        e = ch.get("X", i)
        for k, j in ucol[i]:
            e = f"({e} - {ch.get('LU', k)} * {ch.get('XX', j)})"
        ch.put("XX", i, f"{e} / {ch.get('LU', diag[i])}")
        ch.cost += 2 * len(ucol[i]) + 1
        ch.boundary()
    for i in reversed(range(nvar)):
        if not lcol[i]:
            continue
        e = ch.get("XX", i)
        for k, j in lcol[i]:
            e = f"({e} - {ch.get('LU', k)} * {ch.get('XX', j)})"
        ch.put("XX", i, e)
        ch.cost += 2 * len(lcol[i])
        ch.boundary()
    return ("// KppSolveTR: src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:2327-3701\n"
            + ch.emit())


def generate(ir: dict, ns: str = "mech", budget: int = 250) -> str:
    """Whole generated header for one mechanism.  The function names are fixed
    (``mech_fun``, ``mech_jac_prep``, ...) so one kernel source serves every
    mechanism; each mechanism is compiled into its own library."""
    parts = [
        "// GENERATED by codegen/gen_hip.py from the mechanism IR -- do not edit.",
        f"// mechanism: {ir['name']}  NVAR={ir['nvar']} NFIX={ir['nfix']} "
        f"NREACT={ir['nreact']} LU_NONZERO={ir['nnz']}",
        "#pragma once",
        "#include <hip/hip_runtime.h>",
        f"#define MECH_NAME \"{ir['name']}\"",
        f"#define MECH_NVAR {ir['nvar']}",
        f"#define MECH_NFIX {ir['nfix']}",
        f"#define MECH_NREACT {ir['nreact']}",
        f"#define MECH_NNZ {ir['nnz']}",
        PRELUDE,
        gen_fun(ir, ns, budget),
        gen_jac_prep(ir, ns, budget),
        gen_decomp(ir, ns, budget),
        gen_solve(ir, ns, budget),
        gen_solve_tr(ir, ns, budget),
    ]
    return "\n".join(parts)


def main(argv):
    ir = json.loads(Path(argv[1]).read_text())
    budget = int(argv[3]) if len(argv) > 3 else 250
    Path(argv[2]).write_text(generate(ir, budget=budget))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
