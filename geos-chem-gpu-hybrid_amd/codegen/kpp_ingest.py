"""Mechanism ingest: KPP-generated Fortran (read as text) -> mechanism IR.

The KPP preprocessor writes a chemical mechanism out as straight-line
Fortran.  This module reads those files as plain text (nothing in them is
compiled or executed) and turns them into a small JSON "mechanism IR" that
the HIP code generator and the CPU oracle both consume.  The IR keeps every
floating-point expression in the exact left-to-right form of the generated
code, because bit-level agreement with the reference integrator depends on
the order of every product and sum.

Reference forms read (paths relative to the reference checkout):

* Fun_SPLIT rate monomials and P_VAR/D_VAR sums:
  src/geos-chem/KPP/fullchem/gckpp_Function.F90:2170-5060
* Jac_SP B temporaries and JVS sums, with the sparsity pattern taken from the
  ``! JVS(k) = Jac_FULL(i,j)`` comments:
  src/geos-chem/KPP/fullchem/gckpp_Jacobian.F90:47-20887
* KppSolve statement order (checked against the pattern-derived order):
  src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:645-2311

Independent forms used only for checking the IR (written to test fixtures):
Fun (aggregate) gckpp_Function.F90:47-2150, CalcStoichNum
gckpp_Function.F90:7271-12790, Jac_SP_Vec gckpp_Jacobian.F90:20903-21777,
JacTR_SP_Vec gckpp_Jacobian.F90:21793-22738, KppSolveTR
gckpp_LinearAlgebra.F90:2327-3701.

Expression IR
-------------
A *sum* is a list of ``[sign, product]`` with sign +1/-1; a *product* is a
list of factors evaluated left to right; a factor is ``[kind, index]`` with
kind in {"A", "RCT", "V", "F", "B"} (0-based index) or ``["lit", text]``
for a numeric literal.  Literals keep their decimal text; they are read as
the nearest double, matching the model build's ``-fdefault-real-8``
(src/geos-chem/KPP/fullchem/CMakeLists.txt:88-92).  The value of a sum is
``s1*p1`` (negation, exact) followed by ``acc = acc +/- p_k`` in order.

Usage: python kpp_ingest.py <kpp_dir> <out.json>
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

_GUARD = re.compile(r"^\s*IF\s*\(DO_\w+\(\d+\)\)\s*&\s*$")
_TOKEN = re.compile(
    r"\s*(?:(?P<name>[A-Za-z_]\w*)\((?P<args>\d+(?:,\d+)?)\)"
    r"|(?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eEdD][+-]?\d+)?)"
    r"|(?P<op>[-+*/()=]))"
)


class ParseError(ValueError):
    pass


def _subroutine_lines(text: str, name: str) -> list[str]:
    lines = text.splitlines()
    start = end = None
    for n, line in enumerate(lines):
        s = line.strip()
        if start is None and re.match(rf"SUBROUTINE\s+{name}\s*\(", s):
            start = n
        elif start is not None and re.match(rf"END\s+SUBROUTINE\s+{name}\b", s):
            end = n
            break
    if start is None or end is None:
        raise ParseError(f"subroutine {name} not found")
    return lines[start + 1:end]


def statements(text: str, name: str):
    """Yield (lhs_name, lhs_args, rhs_text, last_comment) for each assignment.

    Continuation lines (trailing '&', leading '&') are joined; the all-TRUE
    DO_* guard lines of the auto-reduce masks are dropped (the model builds
    with auto-reduce off, rundir/geoschem_config.yml:63)."""
    body = _subroutine_lines(text, name)
    buf = ""
    comment = None
    out = []
    for raw in body:
        line = raw.rstrip()
        s = line.strip()
        if not s:
            continue
        if s.startswith("!"):
            comment = s
            continue
        if _GUARD.match(line):
            continue
        if buf:
            if not s.startswith("&"):
                raise ParseError(f"bad continuation in {name}: {s[:60]}")
            s = s[1:]
        cont = s.endswith("&")
        if cont:
            s = s[:-1]
        buf += s
        if cont:
            continue
        stmt, buf = buf, ""
        m = re.match(r"^\s*([A-Za-z_]\w*)\(([\d,]+)\)\s*=(.*)$", stmt)
        if not m:
            # declarations, "Vdot = P_VAR - D_VAR*V", "StoichNum(:,:) = 0." ...
            out.append((None, None, stmt.strip(), comment))
        else:
            args = tuple(int(a) for a in m.group(2).split(","))
            out.append((m.group(1), args, m.group(3).strip(), comment))
        comment = None
    return out


def tokenize(expr: str):
    pos = 0
    toks = []
    expr = expr.strip()
    while pos < len(expr):
        m = _TOKEN.match(expr, pos)
        if not m or m.end() == pos:
            raise ParseError(f"cannot tokenize at {expr[pos:pos+30]!r}")
        pos = m.end()
        if m.group("name"):
            args = tuple(int(a) for a in m.group("args").split(","))
            toks.append(("ref", m.group("name"), args))
        elif m.group("num"):
            toks.append(("lit", m.group("num")))
        else:
            toks.append(("op", m.group("op")))
    return toks


_KINDS = {"A": "A", "RCT": "RCT", "V": "V", "F": "F", "B": "B",
          "JVS": "JVS", "X": "X", "XX": "XX", "UV": "UV"}


def parse_sum(expr: str, allowed: set[str]):
    """Parse 'a*b - c*d + ...' (no parentheses, no division) into the IR."""
    toks = tokenize(expr)
    terms = []
    i = 0
    sign = 1
    if toks and toks[0] == ("op", "-"):
        sign, i = -1, 1
    elif toks and toks[0] == ("op", "+"):
        i = 1
    while True:
        prod = []
        while True:
            if i >= len(toks):
                raise ParseError(f"unexpected end in {expr!r}")
            t = toks[i]
            if t[0] == "ref":
                kind = t[1]
                if kind not in allowed or len(t[2]) != 1:
                    raise ParseError(f"unexpected factor {t} in {expr!r}")
                prod.append([kind, t[2][0] - 1])
            elif t[0] == "lit":
                prod.append(["lit", t[1]])
            else:
                raise ParseError(f"unexpected token {t} in {expr!r}")
            i += 1
            if i < len(toks) and toks[i] == ("op", "*"):
                i += 1
                continue
            break
        terms.append([sign, prod])
        if i >= len(toks):
            break
        t = toks[i]
        if t == ("op", "+"):
            sign = 1
        elif t == ("op", "-"):
            sign = -1
        else:
            raise ParseError(f"unexpected operator {t} in {expr!r}")
        i += 1
    # "0" alone means an empty sum
    if len(terms) == 1 and terms[0][0] == 1 and terms[0][1] == [["lit", "0"]]:
        return []
    return terms


def _lit_value(text: str) -> float:
    return float(text.replace("d", "e").replace("D", "e"))


# ---------------------------------------------------------------------------
# Individual reference forms
# ---------------------------------------------------------------------------

def parse_fun_split(text: str, nvar: int | None = None):
    """Fun_SPLIT: A(r) monomials, P_VAR(i), D_VAR(i) sums (gckpp_Function.F90:2170-5060)."""
    rates, prod, dest = {}, {}, {}
    tail = None
    for lhs, args, rhs, _ in statements(text, "Fun_SPLIT"):
        if lhs is None:
            if rhs.startswith("REAL") or rhs.startswith("IF ( PRESENT"):
                continue
            if rhs.replace(" ", "") == "Vdot=P_VAR-D_VAR*V":
                tail = rhs
                continue
            raise ParseError(f"unexpected statement in Fun_SPLIT: {rhs[:60]}")
        if lhs == "A":
            s = parse_sum(rhs, {"RCT", "V", "F"})
            if len(s) != 1 or s[0][0] != 1:
                raise ParseError(f"rate {args} is not a monomial: {rhs}")
            rates[args[0] - 1] = s[0][1]
        elif lhs == "P_VAR":
            prod[args[0] - 1] = parse_sum(rhs, {"A"})
        elif lhs == "D_VAR":
            dest[args[0] - 1] = parse_sum(rhs, {"RCT", "V", "F"})
        else:
            raise ParseError(f"unexpected lhs {lhs} in Fun_SPLIT")
    if tail is None:
        raise ParseError("Fun_SPLIT: missing 'Vdot = P_VAR - D_VAR*V'")
    nreact = max(rates) + 1
    nv = max(prod) + 1
    if sorted(rates) != list(range(nreact)) or sorted(prod) != list(range(nv)) \
            or sorted(dest) != list(range(nv)):
        raise ParseError("Fun_SPLIT: rates / P_VAR / D_VAR are not contiguous")
    for r, f in rates.items():
        head = f[0]
        if not (head == ["RCT", r] or head[0] == "lit"):
            raise ParseError(f"rate {r+1}: first factor must be RCT({r+1}) or a literal")
        for fac in f[1:]:
            if fac[0] not in ("V", "F"):
                raise ParseError(f"rate {r+1}: unexpected factor {fac}")
    return ([rates[r] for r in range(nreact)], [prod[i] for i in range(nv)],
            [dest[i] for i in range(nv)])


def parse_fun_aggregate(text: str):
    """Fun (aggregate form): Vdot(i) as sums of A(r) (gckpp_Function.F90:47-2150)."""
    rates, vdot = {}, {}
    for lhs, args, rhs, _ in statements(text, "Fun"):
        if lhs is None:
            if rhs.startswith("REAL") or rhs.startswith("IF ( PRESENT"):
                continue
            raise ParseError(f"unexpected statement in Fun: {rhs[:60]}")
        if lhs == "A":
            rates[args[0] - 1] = parse_sum(rhs, {"RCT", "V", "F"})[0][1]
        elif lhs == "Vdot":
            vdot[args[0] - 1] = parse_sum(rhs, {"A"})
        else:
            raise ParseError(f"unexpected lhs {lhs} in Fun")
    return [rates[r] for r in range(len(rates))], [vdot[i] for i in range(len(vdot))]


def parse_jac_sp(text: str):
    """Jac_SP: B temporaries, JVS sums and the (i,j) pattern (gckpp_Jacobian.F90:47-20887)."""
    bvals, bmeta, jvs = {}, {}, {}
    for lhs, args, rhs, comment in statements(text, "Jac_SP"):
        if lhs is None:
            if rhs.startswith("REAL"):
                continue
            raise ParseError(f"unexpected statement in Jac_SP: {rhs[:60]}")
        if lhs == "B":
            s = parse_sum(rhs, {"RCT", "V", "F"})
            if len(s) != 1 or s[0][0] != 1:
                raise ParseError(f"B({args[0]}) is not a monomial: {rhs}")
            bvals[args[0] - 1] = s[0][1]
            m = re.match(r"!\s*B\((\d+)\)\s*=\s*dA\((\d+)\)/dV\((\d+)\)", comment or "")
            if not m or int(m.group(1)) != args[0]:
                raise ParseError(f"B({args[0]}): missing dA/dV comment")
            bmeta[args[0] - 1] = (int(m.group(2)) - 1, int(m.group(3)) - 1)
        elif lhs == "JVS":
            m = re.match(r"!\s*JVS\((\d+)\)\s*=\s*Jac_FULL\((\d+),(\d+)\)", comment or "")
            if not m or int(m.group(1)) != args[0]:
                raise ParseError(f"JVS({args[0]}): missing Jac_FULL comment")
            jvs[args[0] - 1] = (int(m.group(2)) - 1, int(m.group(3)) - 1,
                                parse_sum(rhs, {"B"}))
        else:
            raise ParseError(f"unexpected lhs {lhs} in Jac_SP")
    # KPP leaves some B indices undefined (never written, never read); they
    # are kept as null entries so B numbering matches the generated text.
    nb, nnz = max(bvals) + 1, len(jvs)
    if sorted(jvs) != list(range(nnz)):
        raise ParseError("Jac_SP: JVS not contiguous")
    for k, (_, _, s) in jvs.items():
        for _, prod in s:
            for fac in prod:
                if fac[0] == "B" and fac[1] not in bvals:
                    raise ParseError(f"JVS({k+1}) reads undefined B({fac[1]+1})")
    return ([bvals.get(k) for k in range(nb)],
            [list(bmeta[k]) if k in bmeta else None for k in range(nb)],
            [list(jvs[k]) for k in range(nnz)])


def pattern_tables(pattern, nvar):
    """CSR tables (0-based) of the row-major LU pattern: crow, icol, diag."""
    rows = [p[0] for p in pattern]
    cols = [p[1] for p in pattern]
    for k in range(1, len(pattern)):
        if (rows[k], cols[k]) <= (rows[k - 1], cols[k - 1]):
            raise ParseError("pattern is not sorted row-major")
    crow = [0] * (nvar + 1)
    for r in rows:
        crow[r + 1] += 1
    for i in range(nvar):
        crow[i + 1] += crow[i]
    diag = []
    for i in range(nvar):
        ks = [k for k in range(crow[i], crow[i + 1]) if cols[k] == i]
        if len(ks) != 1:
            raise ParseError(f"row {i+1} has no diagonal entry")
        diag.append(ks[0])
    return crow, cols, diag


def check_lu_closed(crow, icol, diag, nvar):
    """The IKJ elimination of gckpp_LinearAlgebra.F90:61-82 writes only into
    positions of the pattern (otherwise fill would be lost)."""
    for k in range(nvar):
        row = set(icol[crow[k]:crow[k + 1]])
        for kk in range(crow[k], diag[k]):
            j = icol[kk]
            for jj in range(diag[j] + 1, crow[j + 1]):
                if icol[jj] not in row:
                    raise ParseError(f"LU fill ({k+1},{icol[jj]+1}) missing from pattern")


def parse_kpp_solve(text: str):
    """KppSolve statements as (row, [(k, col)...], diag_k or None) in source order
    (gckpp_LinearAlgebra.F90:645-2311)."""
    fwd, bwd = [], []
    for lhs, args, rhs, _ in statements(text, "KppSolve"):
        if lhs is None:
            if rhs.startswith("REAL"):
                continue
            raise ParseError(f"unexpected statement in KppSolve: {rhs[:60]}")
        i = args[0] - 1
        r = rhs.replace(" ", "")
        m = re.match(r"^\(X\((\d+)\)(.*)\)/\(JVS\((\d+)\)\)$", r) or \
            re.match(r"^X\((\d+)\)()/JVS\((\d+)\)$", r)
        if m:
            if int(m.group(1)) - 1 != i:
                raise ParseError(f"KppSolve row {i+1}: rhs names another X")
            terms = re.findall(r"-JVS\((\d+)\)\*X\((\d+)\)", m.group(2))
            if "".join(f"-JVS({a})*X({b})" for a, b in terms) != m.group(2):
                raise ParseError(f"KppSolve row {i+1}: unexpected backward form")
            bwd.append([i, [[int(a) - 1, int(b) - 1] for a, b in terms], int(m.group(3)) - 1])
            continue
        m = re.match(r"^X\((\d+)\)(.*)$", r)
        if not m or int(m.group(1)) - 1 != i:
            raise ParseError(f"KppSolve row {i+1}: unexpected form {r[:40]}")
        terms = re.findall(r"-JVS\((\d+)\)\*X\((\d+)\)", m.group(2))
        if "".join(f"-JVS({a})*X({b})" for a, b in terms) != m.group(2):
            raise ParseError(f"KppSolve row {i+1}: unexpected forward form")
        fwd.append([i, [[int(a) - 1, int(b) - 1] for a, b in terms], None])
    return fwd, bwd


def parse_kpp_solve_tr(text: str):
    """KppSolveTR statements (gckpp_LinearAlgebra.F90:2327-3701)."""
    out = []
    for lhs, args, rhs, _ in statements(text, "KppSolveTR"):
        if lhs is None:
            if rhs.startswith("REAL"):
                continue
            raise ParseError(f"unexpected statement in KppSolveTR: {rhs[:60]}")
        r = rhs.replace(" ", "")
        i = args[0] - 1
        m = re.match(r"^\((XX?)\((\d+)\)(.*)\)/\(JVS\((\d+)\)\)$", r) or \
            re.match(r"^(XX?)\((\d+)\)()/JVS\((\d+)\)$", r)
        if m:
            terms = re.findall(r"-JVS\((\d+)\)\*XX\((\d+)\)", m.group(3))
            if "".join(f"-JVS({a})*XX({b})" for a, b in terms) != m.group(3):
                raise ParseError(f"KppSolveTR {lhs}({i+1}): unexpected form")
            out.append([lhs, i, m.group(1), int(m.group(2)) - 1,
                        [[int(a) - 1, int(b) - 1] for a, b in terms], int(m.group(4)) - 1])
            continue
        m = re.match(r"^(XX?)\((\d+)\)(.*)$", r)
        if not m:
            raise ParseError(f"KppSolveTR {lhs}({i+1}): unexpected form {r[:40]}")
        terms = re.findall(r"-JVS\((\d+)\)\*XX\((\d+)\)", m.group(3))
        if "".join(f"-JVS({a})*XX({b})" for a, b in terms) != m.group(3):
            raise ParseError(f"KppSolveTR {lhs}({i+1}): unexpected form")
        out.append([lhs, i, m.group(1), int(m.group(2)) - 1,
                    [[int(a) - 1, int(b) - 1] for a, b in terms], None])
    return out


def parse_jac_vec(text: str, name: str, lhs_name: str):
    """Jac_SP_Vec / JacTR_SP_Vec: out(i) = sum JVS(k)*UV(j) (gckpp_Jacobian.F90:20903-22738)."""
    out = {}
    for lhs, args, rhs, _ in statements(text, name):
        if lhs is None:
            if rhs.startswith("REAL"):
                continue
            raise ParseError(f"unexpected statement in {name}: {rhs[:60]}")
        if lhs != lhs_name:
            raise ParseError(f"unexpected lhs {lhs} in {name}")
        s = parse_sum(rhs, {"JVS", "UV"})
        terms = []
        for sign, prod in s:
            if sign != 1 or len(prod) != 2 or prod[0][0] != "JVS" or prod[1][0] != "UV":
                raise ParseError(f"{name}: unexpected term in row {args[0]}")
            terms.append([prod[0][1], prod[1][1]])
        out[args[0] - 1] = terms
    return [out[i] for i in range(len(out))]


def parse_stoich(text: str):
    """CalcStoichNum entries (i, r, value text) (gckpp_Function.F90:7271-12790)."""
    out = []
    for lhs, args, rhs, _ in statements(text, "CalcStoichNum"):
        if lhs is None:
            if rhs.startswith("REAL") or rhs.replace(" ", "") == "StoichNum(:,:)=0.":
                continue
            raise ParseError(f"unexpected statement in CalcStoichNum: {rhs[:60]}")
        toks = tokenize(rhs)
        sign = 1
        if toks[0] == ("op", "-"):
            sign, toks = -1, toks[1:]
        if len(toks) != 1 or toks[0][0] != "lit":
            raise ParseError(f"StoichNum{args}: unexpected value {rhs}")
        out.append([args[0] - 1, args[1] - 1, ("-" if sign < 0 else "") + toks[0][1]])
    return out


# ---------------------------------------------------------------------------
# Whole-mechanism ingest
# ---------------------------------------------------------------------------

def ingest(kpp_dir: str | Path, name: str = "fullchem", nfix: int | None = None):
    kpp_dir = Path(kpp_dir)
    fun_txt = (kpp_dir / "gckpp_Function.F90").read_text()
    jac_txt = (kpp_dir / "gckpp_Jacobian.F90").read_text()
    la_txt = (kpp_dir / "gckpp_LinearAlgebra.F90").read_text()

    rates, prod, dest = parse_fun_split(fun_txt)
    bvals, bmeta, jvs = parse_jac_sp(jac_txt)
    nvar = len(prod)
    nreact = len(rates)
    pattern = [(i, j) for i, j, _ in jvs]
    crow, icol, diag = pattern_tables(pattern, nvar)
    check_lu_closed(crow, icol, diag, nvar)

    # KppSolve must follow the order that the CSR tables imply.
    fwd, bwd = parse_kpp_solve(la_txt)
    exp_fwd = [[i, [[kk, icol[kk]] for kk in range(crow[i], diag[i])], None]
               for i in range(nvar) if diag[i] > crow[i]]
    exp_bwd = [[i, [[kk, icol[kk]] for kk in range(diag[i] + 1, crow[i + 1])], diag[i]]
               for i in reversed(range(nvar))]
    if fwd != exp_fwd or bwd != exp_bwd:
        raise ParseError("KppSolve order differs from the CSR-derived order")

    # Fixed-species count: highest F index used (F(1..NFIX)).
    fmax = -1
    for f in rates:
        for k, v in f:
            if k == "F":
                fmax = max(fmax, v)
    nfix_used = fmax + 1
    if nfix is None:
        nfix = nfix_used

    return {
        "name": name,
        "generated_from": {
            "Fun_SPLIT": "src/geos-chem/KPP/fullchem/gckpp_Function.F90:2170-5060",
            "Jac_SP": "src/geos-chem/KPP/fullchem/gckpp_Jacobian.F90:47-20887",
            "KppDecomp": "src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:46-84",
            "KppSolve": "src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:645-2311",
        },
        "nvar": nvar, "nfix": nfix, "nspec": nvar + nfix, "nreact": nreact,
        "nnz": len(jvs), "nb": len(bvals),
        "rate": rates, "prod": prod, "dest": dest,
        "jac_b": bvals, "jac_b_meta": bmeta,
        "jvs": [[i, j, s] for i, j, s in jvs],
        "crow": crow, "icol": icol, "diag": diag,
    }


def ingest_checkforms(kpp_dir: str | Path):
    """Independent reference forms used by the tests to pin the IR."""
    kpp_dir = Path(kpp_dir)
    fun_txt = (kpp_dir / "gckpp_Function.F90").read_text()
    jac_txt = (kpp_dir / "gckpp_Jacobian.F90").read_text()
    la_txt = (kpp_dir / "gckpp_LinearAlgebra.F90").read_text()
    agg_rates, agg_vdot = parse_fun_aggregate(fun_txt)
    return {
        "fun_rate": agg_rates,
        "fun_vdot": agg_vdot,
        "stoich": parse_stoich(fun_txt),
        "jac_vec": parse_jac_vec(jac_txt, "Jac_SP_Vec", "JUV"),
        "jactr_vec": parse_jac_vec(jac_txt, "JacTR_SP_Vec", "JTUV"),
        "solve_tr": parse_kpp_solve_tr(la_txt),
    }


def main(argv):
    if len(argv) != 3:
        print(__doc__)
        return 2
    ir = ingest(argv[1])
    Path(argv[2]).write_text(json.dumps(ir, separators=(",", ":")))
    print(f"{ir['name']}: nvar={ir['nvar']} nfix={ir['nfix']} nreact={ir['nreact']} "
          f"nnz={ir['nnz']} nb={ir['nb']}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
