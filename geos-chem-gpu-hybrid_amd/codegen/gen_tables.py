"""Mechanism IR -> operand tables for the table-driven gfx950 Rodas3 kernels.

The device routines (csrc/gc_rodas3.hip) are small loops that walk these
tables.  Every table word is wave-uniform, so it arrives through scalar loads
(one s_load_dwordx16 fetches four records), and every data operand is one
coalesced ``buffer_load_dwordx2 v, lane_off, rsrc, soffset`` whose soffset is
the table word itself.  The hot loops run fixed-shape records four at a time,
so a wave keeps several table fetches and data loads in flight.  The code is
a few kilobytes whatever the mechanism and stays in the instruction cache.

Per-wave block layout.  The workspace is cut into blocks, one per 64-lane
wave.  Element ``e`` of lane ``l`` sits at byte ``e*512 + l*8`` of its wave's
block; the element offsets of the arrays are the ``E_*`` constants.  Tables
store byte offsets ``e*512``.

Exactness rules (the values must equal the reference's left-to-right
Fortran evaluation bit for bit):

* products keep their data-factor order; a literal factor is applied first:
  exact when the literal is the first or second factor (a single multiply
  commutes) or a power of two anywhere (scaling is exact); anything else is
  rejected here;
* a sum term's sign is folded into its coefficient: ``acc - c*x`` equals
  ``acc + (-c)*x`` and ``-(c*x)`` equals ``(-c)*x`` exactly; terms without a
  coefficient use 1.0 (``1.0*x == x``);
* padding entries are exact no-ops: solve pairs ``0.0*0.0`` subtracted from a
  row value, LU updates into a spare LDS slot.  Sums are never padded (adding
  +0.0 could flip the sign of a -0.0 accumulator).

Tables (int32 words):

* product groups ``prod_<name>``: 8-word records ``{dst, lit, f0, f1, f2, f3,
  0, 0}`` sorted by shape (number of data factors 1-4, literal or not);
  ``prodgrp_<name>`` gives the record count of each of the 8 shapes
  (shape = (nf-1)*2 + has_lit).  Value = lit*f0*f1*... left to right.
  Products: ``rate_y``/``rate_yn`` (A(r), V = Y or Ynew; gckpp_Function.F90:
  2215-3248), ``bjac`` (Jac_SP's B; gckpp_Jacobian.F90:63-3673; a B that is a rate
  constant alone is not formed: its JVS terms read RCONST directly, see
  Tables.b_alias).
* ``pd_y``/``pd_yn``: per species 4-word header ``{nP, nD, V(i), i*512}``
  then its 4-word P terms ``{soff, 0, coef_lo, coef_hi}`` (P_VAR(i) over A)
  and D product terms ``{f1, f2, f3, ci}`` (D_VAR(i), each term
  ((lits[ci]*f1)*f2)*f3 evaluated in place, missing factors reading ONE),
  nP and nD padded to multiples of 4 with terms of value -0.0 (an exact
  no-op: x + (-0.0) == x), interleaved in the kernel's batch order: batch j
  holds P terms 8j..8j+7 then D terms 4j..4j+3 (fewer where the sums end),
  so one batch's loads of both sums go out together;
  Vdot(i) = P_VAR(i) - D_VAR(i)*V(i).  (gckpp_Function.F90:3252-5058)
* ``lu<S>``: fused Jac_SP JVS sums + matrix preparation + KppDecomp with the
  active row segment (at most S entries, plus one spare slot) in LDS; 4-word
  ops: ENTRYG {5, T, 0, 0} + T terms {soff, flags, coef} holding every JVS
  entry of the segment (flags: first/last term of an entry, diagonal, slot
  index << 8; T a multiple of 4); the pivot records of the segment (IKJ
  order: PIVL, a pivot j inside the segment, multiplier -W(j)/U(j,j) stored
  back as L(k,j); PIVS, a pivot of an earlier segment of the same row,
  multiplier -L(k,j) read back from LU) cut into sub-records of four
  updates (a record of more updates continues in sub-records of kind 0 with
  the same multiplier; padding updates hit the spare slot with U = 0.0),
  batched as PIVB {8, G, 0, 0} + G packed words {kind | slot_j << 2 |
  t0 << 8 | t1 << 14 | t2 << 20 | t3 << 26} (kind 1 PIVL, 2 PIVS, 0
  continuation; LDS slot indices; padded to a multiple of 4 words) + per
  sub-record {operand, U(j,c0), U(j,c1), U(j,c2), U(j,c3)} (operand U(j,j),
  L(k,j) or ZERO; padded to a multiple of 4 words), at most PIVB_MAX
  sub-records per batch: the kernel issues a batch's data loads together,
  then applies the sub-records in order; STORE {4, n, 0, 0} + n/2 {slot, LU,
  slot, LU} (n a multiple of 2); END {0,...}.
  (gckpp_Jacobian.F90:3674-20887, gckpp_BatchIntegrator.F90:573-577,
  gckpp_LinearAlgebra.F90:46-84)
* ``fwd<s>`` (stage vector K_s) and ``tr1``/``tr2`` (the TMP vector): solve
  rows, header {n, x_i, U(i,i) or 0, 0} then n/2 {LU, x_j, LU, x_j} with n a
  multiple of 4, absolute offsets.  ``bwd<s>_<S>``: the backward rows of
  K_s with the x_j of the last S columns read from LDS slots (the LU
  segment's, idle during solves; see Tables.bwd_lds).
  (gckpp_LinearAlgebra.F90:645-2311, 2327-3701)
* discrete adjoint: ``adj_nu`` (nu = S^T mu per reaction, 4-word terms
  {TMP(i), NU(r) | flags, coef}, S the net
  stoichiometry of the Fun_SPLIT sums, cf. CalcStoichNum,
  gckpp_Function.F90:7271-12790), then the derivatives of the rate monomials
  A(r) = c*k_r*x1[*x2[*x3]] weighted by nu:
  ``adj_wk_y``/``adj_wk_yn`` (per reaction W(r) = (c*nu(r))*k_r into A's
  elements, free during an adjoint step, and the first-order dL/dk
  GK += (c*nu)*((x1*x2)*x3); 4-word records of element indices, see
  Tables.adj_wk), ``adj_gk2`` + ``adjgrp_gk2`` (second order along the
  direction DIR, 12-word records {GK, NU, c_lo, c_hi, x1, x2, x3, d1, d2, d3,
  0, 0} grouped by factor count) and the gather sums ``adj_lp1_y``/
  ``adj_lp1_yn``/``adj_lp2``: per gradient target (LP(i) of every variable,
  then GF(j) of the fixed species that have terms) 4-word terms {W | a<<16,
  b | c<<16, d, dst | flags} (element indices), each adding W*((a*b)+(c*d)),
  in reaction order, from init (ZERO for LP, GF itself).  ``adj_nu`` and the
  gather sums are streams of terms across all their targets (flags mark the
  first and last term of each target), so a batch of loads spans targets.
"""
from __future__ import annotations

import json
import struct
from fractions import Fraction
import sys
from pathlib import Path

SEGMENTS = (16, 24, 32)  # LU row-segment sizes compiled (LDS entries per lane)


PIVB_MAX = 6   # pivot sub-records per PIVB batch (the kernel has forms for 1..6)


def mask_words(n: int) -> int:
    """32-bit words of a mask of n flags (flag k: bit k % 32 of word k // 32)."""
    return (n + 31) // 32


def layout(ir) -> dict:
    nvar, nfix, nreact, nnz, nb = ir["nvar"], ir["nfix"], ir["nreact"], ir["nnz"], ir["nb"]
    E = {}
    E["Y"] = 0
    E["YNEW"] = nvar
    E["FCN0"] = 2 * nvar
    E["K"] = 3 * nvar          # 4 stage vectors
    E["DFDT"] = 7 * nvar
    E["TMP"] = 8 * nvar        # scratch vector (transposed solves)
    E["RCT"] = 9 * nvar
    E["FIX"] = E["RCT"] + nreact
    E["A"] = E["FIX"] + nfix
    E["B"] = E["A"] + nreact
    E["LU"] = E["B"] + nb
    E["ZERO"] = E["LU"] + nnz  # holds 0.0 (padding operand)
    E["SINK"] = E["ZERO"] + 1  # written by padding stores, never read
    E["ONE"] = E["SINK"] + 1   # holds 1.0 (missing factor of a D_VAR term)
    # auto-reduce masks of the cell (masked kernels only): 32 flags per
    # element, in its low 32 bits -- DO_FUN(NVAR), DO_JVS(LU_NONZERO), DO_SLV(NVAR)
    E["MFUN"] = E["ONE"] + 1
    E["MJVS"] = E["MFUN"] + mask_words(nvar)
    E["MSLV"] = E["MJVS"] + mask_words(nnz)
    E["TOTAL"] = E["MSLV"] + mask_words(nvar)
    # discrete adjoint: the forward block followed by the adjoint vectors
    E["LAM"] = E["TOTAL"]          # lambda = dL/dY at the end of the step
    E["LK"] = E["LAM"] + nvar      # 4 stage adjoints lambda_K
    E["GY"] = E["LK"] + 4 * nvar   # dL/dY accumulated over the step
    E["LP"] = E["GY"] + nvar       # J(P)^T mu of one stage / second-order part
    E["DIR"] = E["LP"] + nvar      # stage vector K_s as a direction
    E["NU"] = E["DIR"] + nvar      # S^T mu per reaction
    E["GK"] = E["NU"] + nreact     # dL/dRCONST
    E["GF"] = E["GK"] + nreact     # dL/dFIX
    E["NZERO"] = E["GF"] + nfix    # holds -0.0 (exact no-op term of the gather sums)
    E["ADJ_TOTAL"] = E["NZERO"] + 1
    return E


def _is_pow2(v: float) -> bool:
    if v == 0.0 or v != v:
        return False
    m, _ = __import__("math").frexp(abs(v))
    return m == 0.5


def _coef_words(c: float):
    lo, hi = struct.unpack("<II", struct.pack("<d", c))
    # as signed int32 for the C initialiser
    return [lo - (1 << 32) if lo >= 1 << 31 else lo, hi - (1 << 32) if hi >= 1 << 31 else hi]


class Tables:
    def __init__(self, ir):
        self.ir = ir
        self.E = layout(ir)
        self.lits: list[float] = []
        self.lit_index: dict[bytes, int] = {}

    def off(self, arr: str, k: int) -> int:
        return (self.E[arr] + k) * 512

    def lit_idx(self, v: float) -> int:
        key = struct.pack("<d", v)  # by bit pattern: -0.0 is its own literal
        if key not in self.lit_index:
            self.lit_index[key] = len(self.lits)
            self.lits.append(v)
        return self.lit_index[key]

    @staticmethod
    def lit_value(text: str) -> float:
        return float(text.replace("d", "e").replace("D", "e"))

    def data_off(self, f, vbase: str) -> int:
        kind, idx = f
        return self.off({"V": vbase, "F": "FIX", "RCT": "RCT", "A": "A", "B": "B"}[kind], idx)

    # -- products ------------------------------------------------------------
    def product_shape(self, prod, vbase):
        """(lit or None, [data offsets]) with the exactness rule of the header."""
        lits = [n for n, f in enumerate(prod) if f[0] == "lit"]
        if len(lits) > 1:
            raise ValueError(f"product with two literals: {prod}")
        lit = None
        if lits:
            n = lits[0]
            lit = self.lit_value(prod[n][1])
            if n >= 2 and not _is_pow2(lit):
                raise ValueError(f"literal in position {n + 1} of {prod} is not exact to move")
        data = [self.data_off(f, vbase) for f in prod if f[0] != "lit"]
        if not 1 <= len(data) <= 4:
            raise ValueError(f"product with {len(data)} data factors: {prod}")
        return lit, data

    def product_group(self, items):
        """items: list of (prod, vbase, dst).  Returns (words, counts[8])."""
        recs = [[] for _ in range(8)]
        for prod, vbase, dst in items:
            lit, data = self.product_shape(prod, vbase)
            shape = (len(data) - 1) * 2 + (1 if lit is not None else 0)
            li = self.lit_idx(lit) if lit is not None else 0
            recs[shape].append([dst, li, *data, *([0] * (4 - len(data))), 0, 0])
        words = [w for group in recs for r in group for w in r]
        return words, [len(g) for g in recs]

    def rates(self, vbase):
        return self.product_group([(f, vbase, self.off("A", r))
                                   for r, f in enumerate(self.ir["rate"])])

    def b_alias(self, b):
        """The RCT element B(b) equals when Jac_SP's B(b) is a rate constant
        alone (B(b) = RCONST(r), gckpp_Jacobian.F90:63-3673), else None: the
        JVS terms then read RCONST(r) itself, bit for bit the same value, and
        B(b) is neither computed nor stored."""
        f = self.ir["jac_b"][b]
        if f is not None and len(f) == 1 and f[0][0] == "RCT":
            return self.off("RCT", f[0][1])
        return None

    def bjac(self):
        return self.product_group([(f, "Y", self.off("B", b))
                                   for b, f in enumerate(self.ir["jac_b"])
                                   if f is not None and self.b_alias(b) is None])

    # -- sums ------------------------------------------------------------------
    def sum_term(self, sign, prod, resolve) -> list[int]:
        """[X] or [lit, X] -> {soff, 0, coef}; coef carries the sign."""
        if len(prod) == 1:
            coef, x = 1.0, prod[0]
        elif len(prod) == 2 and prod[0][0] == "lit":
            coef, x = self.lit_value(prod[0][1]), prod[1]
        else:
            raise ValueError(f"sum term not of the form [lit*]X: {prod}")
        if sign < 0:
            coef = -coef
        return [resolve(x), 0, *_coef_words(coef)]

    def d_term(self, sign, prod, vbase) -> list[int]:
        """D_VAR term {f1, f2, f3, ci}: value ((c*f1)*f2)*f3 with c = lits[ci]
        the signed literal (or +-1) and missing factors reading ONE.  Exact:
        moving the literal in front is the product_shape rule, the sign of a
        product is exact, x*1.0 == x."""
        lit, data = self.product_shape(prod, vbase)
        if len(data) > 3:
            raise ValueError(f"D_VAR term with {len(data)} data factors: {prod}")
        c = (lit if lit is not None else 1.0) * (-1.0 if sign < 0 else 1.0)
        one = self.off("ONE", 0)
        return [*data, *([one] * (3 - len(data))), self.lit_idx(c)]

    def pd(self, vbase) -> list[int]:
        one = self.off("ONE", 0)
        ppad = [one, 0, *_coef_words(-0.0)]              # -0.0 * 1.0
        dpad = [one, one, one, self.lit_idx(-0.0)]       # ((-0.0 * 1) * 1) * 1
        out = []
        for i in range(self.ir["nvar"]):
            P = [self.sum_term(s, p, lambda x: self.data_off(x, vbase))
                 for s, p in self.ir["prod"][i]]
            D = [self.d_term(s, p, vbase) for s, p in self.ir["dest"][i]]
            P += [ppad] * (-len(P) % 4)
            D += [dpad] * (-len(D) % 4)
            out += [len(P), len(D), self.off(vbase, i), i * 512]
            for j in range(max(-(-len(P) // 8), len(D) // 4)):
                for w in P[8 * j:8 * j + 8] + D[4 * j:4 * j + 4]:
                    out += w
        return out

    # -- Jacobian + LU ---------------------------------------------------------
    def entry_group(self, entries, spare) -> list[int]:
        """All JVS entries of one row segment as one ENTRYG record {5, T, 0, 0}
        + T terms {soff, flags, coef_lo, coef_hi} (T a multiple of 4), the
        kernel's loads going out eight or four terms at a time.  flags: 1 first
        term of an entry, 2 last term, 4 diagonal entry, slot index << 8 (6
        bits), the entry's JVS index k << 14 (for its DO_JVS flag).  An entry
        with no term (a structural zero of the LU pattern) is one term
        1.0*ZERO, whose +0.0 is the empty sum; padding terms write the spare
        slot."""
        if self.ir["nnz"] >= 1 << 18:
            raise ValueError("JVS index beyond the 18 bits of an entry term's flags")
        def resolve(x):
            if x[0] == "B" and self.b_alias(x[1]) is not None:
                return self.b_alias(x[1])
            return self.data_off(x, "Y")
        zero = self.off("ZERO", 0)
        words = []
        for sl, dg, terms, k in entries:
            if sl // 512 >= 64:
                raise ValueError("LDS slot beyond the 6 bits of an entry term's flags")
            flags = (4 if dg else 0) | ((sl // 512) << 8) | (k << 14)
            if not terms:
                words += [zero, flags | 3, *_coef_words(1.0)]
                continue
            for q, (sg, prod) in enumerate(terms):
                w = self.sum_term(sg, prod, resolve)
                w[1] = flags | (1 if q == 0 else 0) | (2 if q == len(terms) - 1 else 0)
                words += w
        while len(words) % 16:
            words += [zero, 3 | ((spare // 512) << 8), *_coef_words(1.0)]
        return [5, len(words) // 4, 0, 0] + words

    def lu(self, seg: int) -> list[int]:
        ir = self.ir
        nvar = ir["nvar"]
        crow, icol, diag, jvs = ir["crow"], ir["icol"], ir["diag"], ir["jvs"]
        spare = seg * 512
        ops: list[int] = []

        def pairs(ps, mult, pad):
            ps = list(ps)
            while len(ps) % mult:
                ps.append(pad)
            w = []
            for a, b in ps:
                w += [a, b]
            return len(ps), w

        pad = (spare, self.off("ZERO", 0))

        def emit(recs):
            subs = []  # (kind, slot_j, operand, four (t, U) updates)
            for kind, f1, opnd, ups in recs:
                ups = list(ups)
                first = True
                while first or ups:
                    chunk, ups = ups[:4], ups[4:]
                    chunk += [pad] * (4 - len(chunk))
                    if first:
                        subs.append((1, f1, opnd, chunk) if kind == 2 else (2, 0, opnd, chunk))
                    else:
                        subs.append((0, 0, self.off("ZERO", 0), chunk))
                    first = False
            out = []
            for q in range(0, len(subs), PIVB_MAX):
                grp = subs[q:q + PIVB_MAX]
                packed, data = [], []
                for kind, sj, opnd, chunk in grp:
                    w = kind | (sj // 512) << 2
                    for k, (t, u) in enumerate(chunk):
                        assert t // 512 < 64
                        w |= (t // 512) << (8 + 6 * k)
                    packed.append(w - (1 << 32) if w >= 1 << 31 else w)
                    data += [opnd] + [u for _, u in chunk]
                packed += [0] * (-len(packed) % 4)
                data += [0] * (-len(data) % 4)
                out += [8, len(grp), 0, 0] + packed + data
            return out

        for k in range(nvar):
            cols = icol[crow[k]:crow[k + 1]]
            pos = {c: crow[k] + n for n, c in enumerate(cols)}
            for s0 in range(0, len(cols), seg):
                segcols = cols[s0:s0 + seg]
                slot = {c: n * 512 for n, c in enumerate(segcols)}
                lo = segcols[0]
                ops += self.entry_group([(slot[c], c == k, jvs[pos[c]][2], pos[c])
                                         for c in segcols], spare)
                recs = []
                for kk in range(crow[k], diag[k]):
                    j = icol[kk]
                    if j > segcols[-1]:
                        break
                    ups = [(slot[icol[jj]], self.off("LU", jj))
                           for jj in range(diag[j] + 1, crow[j + 1]) if icol[jj] in slot]
                    if j in slot:
                        recs.append((2, slot[j], self.off("LU", diag[j]), ups))
                    elif ups:
                        assert j < lo
                        recs.append((3, 0, self.off("LU", kk), ups))
                ops += emit(recs)
                st = [(slot[c], self.off("LU", pos[c])) for c in segcols]
                n, w = pairs(st, 2, (spare, self.off("SINK", 0)))
                ops += [4, n, 0, 0] + w
        ops += [0, 0, 0, 0]
        return ops

    # -- discrete adjoint ------------------------------------------------------
    def rate_monomial(self, r):
        """(literal or None, has RCT, [(kind, idx)] data factors) of A(r)."""
        lit, has_k, data = None, False, []
        for kind, idx in self.ir["rate"][r]:
            if kind == "lit":
                lit = self.lit_value(idx)
            elif kind == "RCT":
                if idx != r:
                    raise ValueError(f"A({r + 1}) reads RCT({idx + 1})")
                has_k = True
            else:
                data.append((kind, idx))
        if not 1 <= len(data) <= 3:
            raise ValueError(f"A({r + 1}) has {len(data)} data factors")
        return lit, has_k, data

    def stoichiometry(self):
        """Net stoichiometric coefficients S[i][r] (dict per species) from the
        Fun_SPLIT sums: a P_VAR term c*A(r) adds c; a D_VAR term of species i
        is m*A(r)/V(i), found by its monomial, and subtracts m.  The decimal
        literals are summed exactly and rounded once, as a literal net
        coefficient of CalcStoichNum would be."""
        ir = self.ir
        exact = lambda text: Fraction(str(text).replace("d", "e").replace("D", "e"))
        mono = {}
        for r in range(ir["nreact"]):
            lit, has_k, data = self.rate_monomial(r)
            key = (has_k and r, tuple(sorted(data)))
            mono.setdefault(key, []).append((r, lit))
        S = [dict() for _ in range(ir["nvar"])]
        for i, terms in enumerate(ir["prod"]):
            for sign, prod in terms:
                c = Fraction(1)
                (kind, r), = [f for f in prod if f[0] == "A"]
                for f in prod:
                    if f[0] == "lit":
                        c = exact(f[1])
                S[i][r] = S[i].get(r, 0) + sign * c
        for i, terms in enumerate(ir["dest"]):
            for sign, prod in terms:
                lit, k, data = None, None, [("V", i)]
                for kind, idx in prod:
                    if kind == "lit":
                        lit = exact(idx)
                    elif kind == "RCT":
                        k = idx
                    else:
                        data.append((kind, idx))
                key = (k if k is not None else False, tuple(sorted(data)))
                cands = mono.get(key, [])
                if k is None:
                    cands = [(r, rl) for r, rl in cands if rl is not None]
                if len(cands) != 1:
                    raise ValueError(f"D_VAR({i + 1}) term {prod}: {len(cands)} matching rates")
                r, rl = cands[0]
                m = lit if lit is not None else Fraction(1)
                if k is None:  # literal rate: its literal is part of the D term
                    m = m / exact(self._rate_lit_text(r))
                S[i][r] = S[i].get(r, 0) - sign * m
        return [{r: float(c) for r, c in row.items()} for row in S]

    def _rate_lit_text(self, r):
        (text,) = [v for kind, v in self.ir["rate"][r] if kind == "lit"]
        return text

    def adj_nu(self):
        """Per reaction {n, NU(r), 0, 0} + n terms {TMP(i), 0, S(i,r)}: the
        reaction weights nu = S^T mu of a vector mu held in TMP."""
        cols = [[] for _ in range(self.ir["nreact"])]
        for i, row in enumerate(self.stoichiometry()):
            for r, c in sorted(row.items()):
                if c != 0.0:
                    cols[r].append((i, c))
        out = []
        for r, col in enumerate(cols):
            out += [len(col), self.off("NU", r), 0, 0]
            for i, c in col:
                out += [self.off("TMP", i), 0, *_coef_words(c)]
        return out

    def adj_mono(self, vbase):
        """Per reaction a 16-word record {nf, has_k, lit, NU(r), RCT(r), GK(r),
        x1..x3, d1..d3, g1..g3, 0}: the monomial's data factors at the point
        vbase (Y or YNEW; FIX for fixed species), their direction components
        (DIR for variables, ZERO for fixed species) and their gradient targets
        (LP for variables, GF for fixed species).  Unused slots read ZERO and
        write SINK."""
        z, sink = self.off("ZERO", 0), self.off("SINK", 0)
        out = []
        for r in range(self.ir["nreact"]):
            lit, has_k, data = self.rate_monomial(r)
            x, d, g = [], [], []
            for kind, idx in data:
                if kind == "V":
                    x.append(self.off(vbase, idx))
                    d.append(self.off("DIR", idx))
                    g.append(self.off("LP", idx))
                else:
                    x.append(self.off("FIX", idx))
                    d.append(z)
                    g.append(self.off("GF", idx))
            while len(x) < 3:
                x.append(z)
                d.append(z)
                g.append(sink)
            out += [len(data), 1 if has_k else 0,
                    self.lit_idx(lit) if lit is not None else -1, self.off("NU", r),
                    self.off("RCT", r) if has_k else z, self.off("GK", r) if has_k else sink,
                    *x, *d, *g, 0]
        return out

    def adj_nu_stream(self):
        """nu = S^T mu as one stream of 4-word terms {TMP(i), NU(r) | flags,
        coef_lo, coef_hi} over all reactions in order (flags: 1 first term of
        NU(r): the sum restarts at coef*mu; 2 last term: NU(r) is stored), a
        multiple of 8 terms.  A reaction without terms stores 1.0*ZERO; padding
        terms store into SINK."""
        cols = [[] for _ in range(self.ir["nreact"])]
        for i, row in enumerate(self.stoichiometry()):
            for r, c in sorted(row.items()):
                if c != 0.0:
                    cols[r].append((self.off("TMP", i), c))
        out = []
        for r, col in enumerate(cols):
            col = col or [(self.off("ZERO", 0), 1.0)]
            for k, (x, c) in enumerate(col):
                fl = (1 if k == 0 else 0) | (2 if k == len(col) - 1 else 0)
                out += [x, self.off("NU", r) | fl, *_coef_words(c)]
        while len(out) % 32:
            out += [self.off("ZERO", 0), self.off("SINK", 0) | 3, *_coef_words(0.0)]
        return out

    def _monomial_parts(self, r, vbase):
        """(c, has_k, factors) of A(r): the literal coefficient (1.0 if none)
        and per data factor (value offset at vbase or FIX, direction component
        DIR or ZERO, gradient target LP or GF)."""
        lit, has_k, data = self.rate_monomial(r)
        fac = []
        for kind, idx in data:
            if kind == "V":
                fac.append((self.off(vbase, idx), self.off("DIR", idx), self.off("LP", idx)))
            else:
                fac.append((self.off("FIX", idx), self.off("ZERO", 0), self.off("GF", idx)))
        return (1.0 if lit is None else lit), has_k, fac

    def adj_wk(self, vbase):
        """W(r) = (c*NU(r))*k_r into A's elements together with the first-order
        dL/dk record GK(r) += (c*NU(r))*((x1*x2)*x3); c*NU(r) is left in NU(r)
        for the second-order records.  4-word records of
        element indices {A(r) | GK(r)<<16, NU(r) | K<<16, x1 | x2<<16,
        x3 | lit<<16} (K = RCT(r), or ONE without a rate constant, whose GK
        is SINK; missing factors ONE; lit indexes gc_lits, 1.0 without a
        literal), a multiple of 8 records (padding ones write SINK)."""
        if self.E["ADJ_TOTAL"] > 0xFFFF:
            raise ValueError("element index beyond 16 bits in the W records")
        e = lambda o: o // 512
        one, sink = self.off("ONE", 0), self.off("SINK", 0)
        out = []
        for r in range(self.ir["nreact"]):
            c, has_k, fac = self._monomial_parts(r, vbase)
            x = [f[0] for f in fac] + [one] * (3 - len(fac))
            gk = self.off("GK", r) if has_k else sink
            kx = self.off("RCT", r) if has_k else one
            out += [e(self.off("A", r)) | e(gk) << 16, e(self.off("NU", r)) | e(kx) << 16,
                    e(x[0]) | e(x[1]) << 16, e(x[2]) | self.lit_idx(c) << 16]
        while len(out) % 32:
            out += [e(sink) | e(sink) << 16, e(sink) | e(one) << 16,
                    e(one) | e(one) << 16, e(one) | self.lit_idx(1.0) << 16]
        return [v - (1 << 32) if v >= 1 << 31 else v for v in out]

    def adj_gk2(self):
        """Second-order dL/dk records along DIR, grouped by factor count
        (1, 2, 3): GK(r) += CN*d1, CN*((x2*d1)+(x1*d2)),
        CN*((((x2*x3)*d1)+((x1*x3)*d2))+((x1*x2)*d3)), CN = c*nu(r) (left in
        NU(r) by the W records); 4-word records of element indices
        {GK | NU<<16, x1 | x2<<16, x3 | d1<<16, d2 | d3<<16}, each group a
        multiple of 8 records (padding ones write SINK)."""
        e = lambda o: o // 512
        z, one, sink = self.off("ZERO", 0), self.off("ONE", 0), self.off("SINK", 0)
        groups = [[], [], []]
        for r in range(self.ir["nreact"]):
            c, has_k, fac = self._monomial_parts(r, "Y")
            if not has_k:
                continue
            x = [f[0] for f in fac] + [one] * (3 - len(fac))
            d = [f[1] for f in fac] + [z] * (3 - len(fac))
            groups[len(fac) - 1].append([e(self.off("GK", r)) | e(self.off("NU", r)) << 16,
                                         e(x[0]) | e(x[1]) << 16, e(x[2]) | e(d[0]) << 16,
                                         e(d[1]) | e(d[2]) << 16])
        for g in groups:
            while len(g) % 8:
                g.append([e(sink) | e(z) << 16, e(one) | e(one) << 16, e(one) | e(z) << 16,
                          e(z) | e(z) << 16])
        words = [w for g in groups for rec in g for w in rec]
        return [v - (1 << 32) if v >= 1 << 31 else v for v in words], [len(g) for g in groups]

    def adj_lp(self, second, vbase):
        """Gather sums of the monomial derivatives into their gradient
        targets: first order (d/dx_f of W*prod x) at vbase, or second order
        (d/dx_f of W*sum_g prod_{h!=g} x_h d_g) at Y along DIR.  One stream of
        4-word terms {W | a<<16, b | c<<16, d, dst | flags} (W..d element
        indices), each W*((a*b)+(c*d)); per target its initial value (ZERO
        for LP, GF itself) as the first term and then its terms in reaction
        order (flags as adj_nu_stream); a multiple of 8 terms, padding terms
        storing into SINK."""
        one, nz = self.off("ONE", 0), self.off("NZERO", 0)
        W = lambda r: self.off("A", r)
        terms: dict[int, list] = {}
        for r in range(self.ir["nreact"]):
            _, _, fac = self._monomial_parts(r, "Y" if second else vbase)
            nf = len(fac)
            x = [f[0] for f in fac]
            dd = [f[1] for f in fac]
            for f in range(nf):
                o = [h for h in range(nf) if h != f]   # the other factors, in order
                if not second:
                    if nf == 1:
                        t = (one, one, nz, one)
                    elif nf == 2:
                        t = (x[o[0]], one, nz, one)
                    else:
                        t = (x[o[0]], x[o[1]], nz, one)
                else:
                    if nf == 1:
                        continue
                    if nf == 2:
                        t = (dd[o[0]], one, nz, one)
                    else:   # w*(x_b*d_a + x_a*d_b) for the others a < b
                        a, b = o
                        t = (x[b], dd[a], x[a], dd[b])
                terms.setdefault(fac[f][2], []).append((W(r), *t))
        if self.E["ADJ_TOTAL"] > 0xFFFF:
            raise ValueError("element index beyond 16 bits in the gather terms")
        e = lambda o: o // 512
        out = []
        targets = [self.off("LP", i) for i in range(self.ir["nvar"])]
        targets += [g for g in (self.off("GF", j) for j in range(self.ir["nfix"])) if g in terms]
        for g in targets:
            init = self.off("ZERO", 0) if g < self.off("GF", 0) else g
            # the initial value as a term: init*((1*1)+(-0*1)) == init
            ts = [(init, one, one, nz, one)] + list(terms.get(g, []))
            for k, (w, a, b, c, d) in enumerate(ts):
                fl = (1 if k == 0 else 0) | (2 if k == len(ts) - 1 else 0)
                out += [e(w) | e(a) << 16, e(b) | e(c) << 16, e(d), g | fl]
        while len(out) % 32:
            out += [e(nz) | e(one) << 16, e(one) | e(nz) << 16, e(one), self.off("SINK", 0) | 3]
        return [v - (1 << 32) if v >= 1 << 31 else v for v in out]   # as int32

    # -- solves ----------------------------------------------------------------
    # One table per target vector, absolute offsets: padding pairs read ZERO
    # for both operands, so x - 0.0*0.0 = x exactly.
    def _rows(self, rows) -> list[int]:
        z = self.off("ZERO", 0)
        out = []
        for xi, d, prs in rows:
            prs = list(prs)
            while len(prs) % 4:
                prs.append((z, z))
            out += [len(prs), xi, d, 0]
            for a, b in prs:
                out += [a, b]
        return out

    def vec(self, vbase: str, stage: int, i: int) -> int:
        return self.off(vbase, stage * self.ir["nvar"] + i)

    def fwd(self, vbase, stage=0):
        ir = self.ir
        crow, icol, diag = ir["crow"], ir["icol"], ir["diag"]
        x = lambda i: self.vec(vbase, stage, i)
        rows = [(x(i), 0, [(self.off("LU", j), x(icol[j])) for j in range(crow[i], diag[i])])
                for i in range(ir["nvar"]) if diag[i] > crow[i]]
        return self._rows(rows), len(rows)

    def bwd(self, vbase, stage=0):
        ir = self.ir
        crow, icol, diag = ir["crow"], ir["icol"], ir["diag"]
        x = lambda i: self.vec(vbase, stage, i)
        rows = [(x(i), self.off("LU", diag[i]),
                 [(self.off("LU", j), x(icol[j])) for j in range(diag[i] + 1, crow[i + 1])])
                for i in reversed(range(ir["nvar"]))]
        return self._rows(rows), len(rows)

    def bwd_lds(self, vbase, stage, seg):
        """Backward rows whose x_j of the last `seg` columns come from LDS.

        Those columns carry most of the backward solve's x loads (the dense
        trailing block, referenced by many earlier rows), and their rows are
        the first `seg` rows solved, so each such x_i is put into LDS slot
        i - (nvar - seg) as it is solved and read from there afterwards.  The
        terms keep KppSolve's ascending-column order: cold columns first, then
        the LDS suffix.  Row: {nc | nh << 16, x_i, U(i,i), lds_i or -1} +
        nc/2 {LU, x_j, LU, x_j} + nh/2 {LU, slot_j*512, LU, slot_j*512}; nc
        and nh are padded to multiples of 4, cold padding with (ZERO, ZERO),
        LDS padding with (ZERO, spare slot seg), which the solve sets to 0.0.
        """
        ir = self.ir
        n = ir["nvar"]
        crow, icol, diag = ir["crow"], ir["icol"], ir["diag"]
        h0 = n - seg
        z = self.off("ZERO", 0)
        x = lambda i: self.vec(vbase, stage, i)
        out = []
        for i in reversed(range(n)):
            cold, hot = [], []
            for k in range(diag[i] + 1, crow[i + 1]):
                j = icol[k]
                if j >= h0:
                    hot.append((self.off("LU", k), (j - h0) * 512))
                else:
                    assert not hot, "U columns of a row must ascend"
                    cold.append((self.off("LU", k), x(j)))
            while len(cold) % 4:
                cold.append((z, z))
            while len(hot) % 4:
                hot.append((z, seg * 512))
            out += [len(cold) | len(hot) << 16, x(i), self.off("LU", diag[i]),
                    (i - h0) * 512 if i >= h0 else -1]
            for a, b in cold + hot:
                out += [a, b]
        return out, n

    def tr(self, vbase="TMP"):
        """KppSolveTR in place on one vector: phase 1 U^T (rows ascending,
        U(j,i) for ascending j, divide), phase 2 L^T (rows descending, L(j,i)
        ascending)."""
        ir = self.ir
        n = ir["nvar"]
        crow, icol, diag = ir["crow"], ir["icol"], ir["diag"]
        x = lambda i: self.off(vbase, i)
        ucol = [[] for _ in range(n)]
        lcol = [[] for _ in range(n)]
        for j in range(n):
            for k in range(diag[j] + 1, crow[j + 1]):
                ucol[icol[k]].append((self.off("LU", k), x(j)))
            for k in range(crow[j], diag[j]):
                lcol[icol[k]].append((self.off("LU", k), x(j)))
        r1 = [(x(i), self.off("LU", diag[i]), ucol[i]) for i in range(n)]
        r2 = [(x(i), 0, lcol[i]) for i in reversed(range(n)) if lcol[i]]
        return self._rows(r1), len(r1), self._rows(r2), len(r2)


def _arr(name, vals, ctype="int") -> str:
    """A table as a 64-byte aligned device array.  Int tables are padded with
    zeros to a multiple of 16 words plus 32, so the kernels may fetch 16-byte
    groups of words (every record starts on a 4-word boundary) and read up to
    32 words past the last record (the next record's words fetched ahead of
    time) without leaving the array."""
    vals = list(vals)
    if ctype == "int":
        vals += [0] * ((-len(vals)) % 16 + 32)
    body = ",".join(str(v) for v in vals) if vals else "0"
    return f"__device__ const {ctype} {name}[] __attribute__((aligned(64))) = {{{body}}};\n"


def generate(ir: dict) -> str:
    t = Tables(ir)
    E = t.E
    parts = [
        "// GENERATED by codegen/gen_tables.py from the mechanism IR -- do not edit.",
        f"// mechanism: {ir['name']}  NVAR={ir['nvar']} NFIX={ir['nfix']} "
        f"NREACT={ir['nreact']} LU_NONZERO={ir['nnz']}",
        "#pragma once",
        f'#define MECH_NAME "{ir["name"]}"',
        f"#define MECH_NVAR {ir['nvar']}",
        f"#define MECH_NFIX {ir['nfix']}",
        f"#define MECH_NREACT {ir['nreact']}",
        f"#define MECH_NNZ {ir['nnz']}",
        f"#define MECH_MW_FUN {mask_words(ir['nvar'])}",
        f"#define MECH_MW_JVS {mask_words(ir['nnz'])}",
        f"#define MECH_MW_SLV {mask_words(ir['nvar'])}",
    ]
    for k, v in E.items():
        parts.append(f"#define MECH_E_{k} {v}")
    body = []
    sizes = {}
    for name, (words, counts) in {"rate_y": t.rates("Y"), "rate_yn": t.rates("YNEW"),
                                  "bjac": t.bjac()}.items():
        body.append(_arr(f"gc_prod_{name}", words))
        body.append(_arr(f"gc_prodgrp_{name}", counts))
        sizes[name] = len(words)
    pd_y, pd_yn = t.pd("Y"), t.pd("YNEW")
    body += [_arr("gc_pd_y", pd_y), _arr("gc_pd_yn", pd_yn)]
    nf = nbw = 0
    for stage in range(4):
        fw, nf = t.fwd("K", stage)
        body.append(_arr(f"gc_fwd{stage}", fw))
        sizes[f"fwd{stage}"] = len(fw)
        for s in SEGMENTS:
            bw, nbw = t.bwd_lds("K", stage, s)
            body.append(_arr(f"gc_bwd{stage}_{s}", bw))
        sizes[f"bwd{stage}"] = len(bw)
    tr1, n1, tr2, n2 = t.tr("TMP")
    body += [_arr("gc_tr1", tr1), _arr("gc_tr2", tr2)]
    body.append(_arr("gc_adj_nu", t.adj_nu_stream()))
    wk_y, wk_yn = t.adj_wk("Y"), t.adj_wk("YNEW")
    body += [_arr("gc_adj_wk_y", wk_y), _arr("gc_adj_wk_yn", wk_yn)]
    words, counts = t.adj_gk2()
    body += [_arr("gc_adj_gk2", words), _arr("gc_adjgrp_gk2", counts)]
    nterm = {"NU": len(t.adj_nu_stream()) // 4}
    for name, (second, vb) in {"lp1_y": (False, "Y"), "lp1_yn": (False, "YNEW"),
                               "lp2": (True, "Y")}.items():
        words = t.adj_lp(second, vb)
        nterm[name.upper()] = len(words) // 4
        body.append(_arr(f"gc_adj_{name}", words))
    parts += [f"#define MECH_NWK {len(wk_y) // 4}"]
    parts += [f"#define MECH_NT_{k} {v}" for k, v in nterm.items()]
    for s in SEGMENTS:
        ops = t.lu(s)
        body.append(_arr(f"gc_lu{s}", ops))
        sizes[f"lu{s}"] = len(ops)
    body.insert(0, _arr("gc_lits", [repr(x) for x in t.lits] or ["0.0"], "double"))
    parts += [f"#define MECH_NFWD {nf}", f"#define MECH_NBWD {nbw}",
              f"#define MECH_NTR1 {n1}", f"#define MECH_NTR2 {n2}"]
    sizes.update(pd=len(pd_y))
    parts.append("// table words: " + ", ".join(f"{k}={v}" for k, v in sizes.items()))
    return "\n".join(parts) + "\n" + "".join(body)


def main(argv):
    ir = json.loads(Path(argv[1]).read_text())
    Path(argv[2]).write_text(generate(ir))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
