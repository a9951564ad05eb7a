"""Mechanism IR: loading, dimensions and exact operation counts.

The IR (``mechanisms/<name>.json``) is produced by ``codegen/kpp_ingest.py``
from the KPP-generated Fortran of the reference (fullchem) or by
``synthetic.py`` (stress mechanisms).  The operation counts below are those
of the reference routines, counted from the same expressions the code
generator emits, so they are the algorithmic FP64 work of one call.
"""
from __future__ import annotations

import json
from dataclasses import dataclass
from functools import lru_cache
from pathlib import Path

import numpy as np

from . import PKG_DIR

MECH_DIR = PKG_DIR / "mechanisms"


def _sum_ops(terms) -> int:
    """Multiplies and adds of a left-to-right sum of products."""
    if not terms:
        return 0
    return sum(len(p) - 1 for _, p in terms) + len(terms) - 1


@dataclass
class Mechanism:
    ir: dict

    @property
    def name(self) -> str:
        return self.ir["name"]

    @property
    def nvar(self) -> int:
        return self.ir["nvar"]

    @property
    def nfix(self) -> int:
        return self.ir["nfix"]

    @property
    def nspec(self) -> int:
        return self.ir["nvar"] + self.ir["nfix"]

    @property
    def nreact(self) -> int:
        return self.ir["nreact"]

    @property
    def nnz(self) -> int:
        return self.ir["nnz"]

    @property
    def crow(self):
        return np.asarray(self.ir["crow"], np.int64)

    @property
    def icol(self):
        return np.asarray(self.ir["icol"], np.int64)

    @property
    def diag(self):
        return np.asarray(self.ir["diag"], np.int64)

    # -- operation counts (FP64 flops of one call, reference form) -------
    def flops_fun(self) -> int:
        """Fun_SPLIT: rates, P_VAR, D_VAR, Vdot = P - D*V."""
        ir = self.ir
        n = sum(len(f) - 1 for f in ir["rate"])
        n += sum(_sum_ops(s) for s in ir["prod"])
        n += sum(_sum_ops(s) for s in ir["dest"])
        return n + 2 * ir["nvar"]

    def flops_jac(self) -> int:
        """Jac_SP: B temporaries and JVS sums."""
        ir = self.ir
        n = sum(len(f) - 1 for f in ir["jac_b"] if f is not None)
        return n + sum(_sum_ops(s) for _, _, s in ir["jvs"])

    def flops_decomp(self) -> int:
        """KppDecomp: one divide per L entry, multiply-add per update."""
        crow, icol, diag = self.ir["crow"], self.ir["icol"], self.ir["diag"]
        nU = [crow[j + 1] - diag[j] - 1 for j in range(self.nvar)]
        n = 0
        for k in range(self.nvar):
            for kk in range(crow[k], diag[k]):
                n += 1 + 2 * nU[icol[kk]]
        return n

    def flops_prepare(self) -> int:
        """ros_PrepareMatrix around KppDecomp: ghinv and the diagonal adds
        (the -Jac0 copy is a sign change, not counted)."""
        return 3 + self.nvar

    def flops_solve(self) -> int:
        crow, diag = self.ir["crow"], self.ir["diag"]
        n = 0
        for i in range(self.nvar):
            n += 2 * (diag[i] - crow[i])            # forward (unit L)
            n += 2 * (crow[i + 1] - diag[i] - 1) + 1  # backward + divide
        return n

    def flops_step_vector(self) -> int:
        """Per attempt, outside the four kernels: stage combinations, Ynew,
        Yerr and the error norm (gckpp_BatchIntegrator.F90:595-668)."""
        nv = self.nvar
        stage_axpy = 2 * nv * (1 + 2 + 3)        # HC*K_j terms of stages 2-4
        stage_ynew = 2 * nv * (2 + 3)            # Ynew for stages 3 and 4
        hc = 1 + 2 + 3                           # the HC divides
        ynew = 2 * nv * 4
        yerr = 2 * nv * 4
        norm = nv * (2 + 1 + 2) + 3              # Scal, divide, square-add; sqrt etc.
        control = 8                              # Fac, Hnew, T update
        return stage_axpy + stage_ynew + hc + ynew + yerr + norm + control

    def algorithmic_flops(self, istat: np.ndarray) -> float:
        """FP64 flops the reference Rodas3 performs for cells with counters
        ISTAT[:, 0:8] = Nfun Njac Nstp Nacc Nrej Ndec Nsol Nsng."""
        ist = np.asarray(istat, np.float64).reshape(-1, 8).sum(axis=0)
        nfun, njac, nstp, _, _, ndec, nsol, _ = ist
        return (self.flops_fun() * nfun + self.flops_jac() * njac
                + (self.flops_decomp() + self.flops_prepare()) * ndec
                + self.flops_solve() * nsol + self.flops_step_vector() * nstp)

    def compulsory_bytes_per_cell(self) -> int:
        """Bytes a cell's solve must move to/from HBM: read C(NSPEC), RCONST,
        ICNTRL(20), RCNTRL(20), Perm; write the NVAR part of C (the fixed
        species are not written back), ISTAT(8), RSTAT(3), IERR."""
        rd = 8 * self.nspec + 8 * self.nreact + 4 * 20 + 8 * 20 + 4
        wr = 8 * self.nvar + 4 * 8 + 8 * 3 + 4
        return rd + wr


@lru_cache(maxsize=None)
def load(name: str = "fullchem") -> Mechanism:
    path = Path(name) if name.endswith(".json") else MECH_DIR / f"{name}.json"
    return Mechanism(json.loads(path.read_text()))
