"""Build mechanism IRs from reaction lists, in the conventions KPP generates.

Given species and reactions (reactants and products with stoichiometric
coefficients), this produces the same IR ``codegen/kpp_ingest.py`` extracts
from KPP's Fortran, so synthetic mechanisms run through the same code
generator, oracle and kernels as fullchem:

* rate     A(r) = RCT(r) * V(a) * V(b) ... * F(c)          (reactant factors)
* P_VAR(i) = sum over reactions with net production of i of  coef*A(r)
* D_VAR(i) = sum over reactions consuming i of RCT(r)*m*(other factors), where
             the reactant factor V(i) is removed and m is its multiplicity
             (KPP writes 2 X -> ... as D = RCT*2*V(X))
* B(r,s)   = dA(r)/dV(s) = RCT(r)*m*(other factors)
* JVS(i,j) = sum over r of S(i,r)*B(r,j),  S = products - reactants
* LU pattern: the Jacobian pattern plus diagonal plus the fill of the
  unpivoted IKJ factorisation, species first reordered by a greedy
  minimum-degree rule (KPP reorders species to reduce fill as well).

Terms of every sum are in ascending reaction order, as in KPP's output.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Reaction:
    reactants: dict          # species index -> multiplicity (variable species)
    products: dict           # species index -> coefficient (float)
    fixed: dict = field(default_factory=dict)  # fixed species index -> multiplicity
    literal: str | None = None  # literal rate constant text instead of RCT(r)


def _coef_text(c: float) -> str:
    if float(c).is_integer():
        return str(int(c))
    return repr(float(c))


def build_ir(name: str, nvar: int, nfix: int, reactions: list[Reaction],
             reorder: bool = True) -> dict:
    R = len(reactions)
    # ---- species order (fill-reducing) ---------------------------------------
    adj = [set() for _ in range(nvar)]
    for rx in reactions:
        touched = set(rx.reactants) | set(rx.products)
        for s in rx.reactants:
            for i in touched:
                adj[i].add(s)
                adj[s].add(i)
    if reorder:
        order = _min_degree_order(adj, nvar)
    else:
        order = list(range(nvar))
    new = {old: k for k, old in enumerate(order)}
    rxs = [Reaction({new[s]: m for s, m in rx.reactants.items()},
                    {new[s]: c for s, c in rx.products.items()},
                    dict(rx.fixed), rx.literal) for rx in reactions]

    # ---- rates -----------------------------------------------------------------
    rate = []
    for r, rx in enumerate(rxs):
        head = ["lit", rx.literal] if rx.literal else ["RCT", r]
        f = [head]
        for s in sorted(rx.reactants):
            f += [["V", s]] * rx.reactants[s]
        for s in sorted(rx.fixed):
            f += [["F", s]] * rx.fixed[s]
        rate.append(f)

    def without(r, s, m):
        """Factors of A(r) with one V(s) removed, times the factor m."""
        rx = rxs[r]
        head = ["lit", rx.literal] if rx.literal else ["RCT", r]
        f = [head]
        if m != 1:
            f.append(["lit", _coef_text(m)])
        for t in sorted(rx.reactants):
            k = rx.reactants[t] - (1 if t == s else 0)
            f += [["V", t]] * k
        for t in sorted(rx.fixed):
            f += [["F", t]] * rx.fixed[t]
        return f

    S = np.zeros((nvar, R))
    for r, rx in enumerate(rxs):
        for s, m in rx.reactants.items():
            S[s, r] -= m
        for s, c in rx.products.items():
            S[s, r] += c

    prod = [[] for _ in range(nvar)]
    dest = [[] for _ in range(nvar)]
    for r, rx in enumerate(rxs):
        for i, c in sorted(rx.products.items()):
            net = c - rx.reactants.get(i, 0)
            if net > 0:
                prod[i].append([1, [["A", r]] if net == 1 else [["lit", _coef_text(net)], ["A", r]]])
        for i in sorted(rx.reactants):
            net = rx.products.get(i, 0.0) - rx.reactants[i]
            if net < 0:
                # loss rate = -net * A(r) = D * V(i)
                dest[i].append([1, without(r, i, -net)])
    # terms in ascending reaction order already (reactions visited in order)

    # ---- Jacobian ---------------------------------------------------------------
    bmeta, jb = [], []
    bidx = {}
    for r, rx in enumerate(rxs):
        for s in sorted(rx.reactants):
            bidx[(r, s)] = len(jb)
            jb.append(without(r, s, rx.reactants[s]))
            bmeta.append([r, s])
    pattern = set((i, i) for i in range(nvar))
    terms = {}
    for (r, s), b in bidx.items():
        for i in range(nvar):
            if S[i, r] != 0:
                pattern.add((i, s))
                terms.setdefault((i, s), []).append((r, b, S[i, r]))
    # symbolic IKJ fill
    rows = [set() for _ in range(nvar)]
    for (i, j) in pattern:
        rows[i].add(j)
    for k in range(nvar):
        # eliminate columns j < k in increasing order, including fill that
        # earlier eliminations of this row create
        heap = sorted(c for c in rows[k] if c < k)
        import heapq
        heapq.heapify(heap)
        seen = set(heap)
        while heap:
            j = heapq.heappop(heap)
            for c in rows[j]:
                if c > j and c not in rows[k]:
                    rows[k].add(c)
                    if c < k and c not in seen:
                        seen.add(c)
                        heapq.heappush(heap, c)
    jvs = []
    for i in range(nvar):
        for j in sorted(rows[i]):
            sm = []
            for r, b, c in sorted(terms.get((i, j), [])):
                sign = 1 if c > 0 else -1
                mag = abs(c)
                p = [["B", b]] if mag == 1 else [["lit", _coef_text(mag)], ["B", b]]
                sm.append([sign, p])
            jvs.append([i, j, sm])
    crow = [0]
    icol, diag = [], []
    for i in range(nvar):
        cols = sorted(rows[i])
        for c in cols:
            if c == i:
                diag.append(len(icol))
            icol.append(c)
        crow.append(len(icol))
    return {
        "name": name, "nvar": nvar, "nfix": nfix, "nspec": nvar + nfix, "nreact": R,
        "nnz": len(jvs), "nb": len(jb), "rate": rate, "prod": prod, "dest": dest,
        "jac_b": jb, "jac_b_meta": bmeta, "jvs": jvs, "crow": crow, "icol": icol,
        "diag": diag, "species_order": order,
        "generated_from": {"synthetic": name},
    }


def _min_degree_order(adj, n):
    """Greedy minimum degree on the symmetric pattern, with elimination
    cliques (a standard fill-reducing heuristic)."""
    g = [set(a) - {i} for i, a in enumerate(adj)]
    alive = set(range(n))
    order = []
    while alive:
        v = min(alive, key=lambda x: (len(g[x]), x))
        order.append(v)
        nb = g[v]
        for a in nb:
            g[a] |= nb
            g[a].discard(a)
            g[a].discard(v)
        alive.remove(v)
        for a in nb:
            g[a] &= alive
    return order


def stress_mechanism(nvar: int = 500, nreact: int = 1500, seed: int = 4,
                     nhub: int = 8) -> dict:
    """The stress mechanism's IR (reactions of stress_reactions)."""
    reactions, kinds_out = stress_reactions(nvar, nreact, seed, nhub)
    ir = build_ir(f"stress{nvar}", nvar, 1, reactions)
    ir["reaction_kind"] = kinds_out
    return ir


def stoichiometry(reactions: list[Reaction], nvar: int, order=None) -> np.ndarray:
    """Net stoichiometric matrix S[i, r] = products - reactants of a reaction
    list, with species renumbered by ``order`` (an IR's species_order: new
    index k holds old species order[k])."""
    new = {old: k for k, old in enumerate(order)} if order is not None else None
    S = np.zeros((nvar, len(reactions)))
    for r, rx in enumerate(reactions):
        for s, m in rx.reactants.items():
            S[new[s] if new else s, r] -= m
        for s, c in rx.products.items():
            S[new[s] if new else s, r] += c
    return S


def stress_reactions(nvar: int = 500, nreact: int = 1500, seed: int = 4,
                     nhub: int = 8):
    """Seeded stiff stress mechanism in the shape of tropospheric chemistry:
    ``nhub`` radical hubs (the OH/HO2/NO/NO2/O3 role) react with the other
    ("leaf") species, which form oxidation chains.  Kinds: 40% bimolecular
    (leaf + hub), 30% unimolecular (leaf), 20% photolysis-like (leaf; its rate
    is set by the input generator), 10% termolecular (leaf + hub + hub).
    Products: the next leaf of the chain (coefficient 1 or 0.5-1.5), a random
    second leaf (probability 0.08) and a hub (probability 0.5).  The fill-reducing order
    puts the hubs last, giving a dense trailing block as in fullchem and an LU
    fill of about 3%."""
    rng = np.random.default_rng(seed)
    nleaf = nvar - nhub
    hubs = list(range(nleaf, nvar))
    reactions, kinds_out = [], []
    kinds = (["bi"] * 40 + ["uni"] * 30 + ["photo"] * 20 + ["ter"] * 10)
    order = rng.permutation(100)

    def coef():
        return 1.0 if rng.uniform() < 0.7 else float(np.round(rng.uniform(0.5, 1.5), 3))

    for r in range(nreact):
        kind = kinds[order[r % 100]]
        leaf = int(rng.integers(nleaf))
        re = {leaf: 1}
        if kind in ("bi", "ter"):
            h = hubs[int(rng.integers(nhub))]
            re[h] = re.get(h, 0) + 1
        if kind == "ter":
            h = hubs[int(rng.integers(nhub))]
            re[h] = re.get(h, 0) + 1
        pr = {}
        nxt = (leaf + int(rng.integers(1, 4))) % nleaf
        pr[nxt] = coef()
        if rng.uniform() < 0.08:
            o = int(rng.integers(nleaf))
            if o not in re:
                pr[o] = pr.get(o, 0.0) + coef()
        if rng.uniform() < 0.5:
            h = hubs[int(rng.integers(nhub))]
            pr[h] = pr.get(h, 0.0) + 1.0
        reactions.append(Reaction(re, pr))
        kinds_out.append(kind)
    return reactions, kinds_out
