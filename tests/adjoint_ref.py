"""NumPy reference of the Rodas3 discrete adjoint (test infrastructure).

The discrete map is the sequence of accepted Rodas3 steps with their step
sizes held fixed (oracle ``replay``); its adjoint is swept here step by step
in reverse with the oracle's own Fun_SPLIT, Jac_SP, KppDecomp, KppSolve and
KppSolveTR, the stoichiometric matrix of the reference's CalcStoichNum
(gckpp_Function.F90:7271-12790, from the golden fixture) and the rate
monomials of Fun_SPLIT.  For one accepted step (autonomous form)

    M = I/(gamma*Dir*H) - J(Y),   K_s = M^-1 (f(P_s) + sum_j C_sj/(Dir*H) K_j),
    P_1 = P_2 = Y,  P_s = Y + sum_j A_sj K_j,   Y' = Y + sum_s M_s K_s,

so with mu_s = M^-T lambda_Ks, taken for s = 4..1,

    lambda_Kj += (C_sj/(Dir*H)) mu_s + A_sj J(P_s)^T mu_s,
    lambda_Y  += lambda' + sum_s J(P_s)^T mu_s + sum_s d/dY (mu_s^T J(Y) K_s),

and the rate constants and fixed species collect the same terms through
dA_r/dk_r and dA_r/dF.  It is the checker of the HIP adjoint kernel and is
itself checked against central finite differences of the replay.
"""
from __future__ import annotations

import numpy as np

ROS_A = (0.0, 2.0, 0.0, 2.0, 0.0, 1.0)
ROS_C = (4.0, 1.0, -1.0, 1.0, -1.0, -(8.0 / 3.0))
ROS_M = (2.0, 0.0, 1.0, 1.0)
GAMMA = 0.5


class Monomials:
    """Rate monomials A_r = c_r * k_r^e_r * prod z[idx_rf] over the state
    z = [V, F] (Fun_SPLIT, gckpp_Function.F90:2215-3248)."""

    def __init__(self, ir):
        nvar, nfix = ir["nvar"], ir["nfix"]
        self.nz = nvar + nfix
        R = len(ir["rate"])
        self.c = np.ones(R)
        self.has_k = np.zeros(R, bool)
        self.idx = np.full((R, 3), self.nz, dtype=np.int64)  # nz -> a 1.0 pad
        self.nf = np.zeros(R, dtype=np.int64)
        for r, prod in enumerate(ir["rate"]):
            f = 0
            for kind, val in prod:
                if kind == "lit":
                    self.c[r] = float(str(val).replace("d", "e").replace("D", "e"))
                elif kind == "RCT":
                    assert val == r
                    self.has_k[r] = True
                else:
                    self.idx[r, f] = val if kind == "V" else nvar + val
                    f += 1
            self.nf[r] = f

    def _x(self, z):
        zz = np.append(z, 1.0)
        return zz[self.idx]                          # [R, 3]

    def first_order(self, z, k, nu):
        """sum_r nu_r dA_r/dz (len nz) and nu_r dA_r/dk_r (len R)."""
        x = self._x(z)
        kk = np.where(self.has_k, k, 1.0)
        w = nu * self.c * kk
        m = np.stack([x[:, 1] * x[:, 2], x[:, 0] * x[:, 2], x[:, 0] * x[:, 1]], axis=1)
        gz = np.zeros(self.nz + 1)
        for f in range(3):
            np.add.at(gz, self.idx[:, f], np.where(self.nf > f, w * m[:, f], 0.0))
        gk = np.where(self.has_k, nu * self.c * x[:, 0] * x[:, 1] * x[:, 2], 0.0)
        return gz[:self.nz], gk

    def second_order(self, z, k, nu, dirv):
        """Gradient of sum_r nu_r (grad_z A_r . d) with d = [dirv, 0 for F]."""
        x = self._x(z)
        dd = np.append(np.append(dirv, np.zeros(self.nz - dirv.size)), 0.0)[self.idx]
        kk = np.where(self.has_k, k, 1.0)
        w = nu * self.c * kk
        x1, x2, x3 = x[:, 0], x[:, 1], x[:, 2]
        d1, d2, d3 = dd[:, 0], dd[:, 1], dd[:, 2]
        nf = self.nf
        # d/dx_f of sum_g (prod_{h != g} x_h) d_g, for nf = 1, 2, 3
        g1 = np.where(nf == 2, d2, np.where(nf == 3, x3 * d2 + x2 * d3, 0.0))
        g2 = np.where(nf == 2, d1, np.where(nf == 3, x3 * d1 + x1 * d3, 0.0))
        g3 = np.where(nf == 3, x2 * d1 + x1 * d2, 0.0)
        gz = np.zeros(self.nz + 1)
        for f, g in enumerate((g1, g2, g3)):
            np.add.at(gz, self.idx[:, f], np.where(nf > f, w * g, 0.0))
        dirder = np.where(nf == 1, d1,
                          np.where(nf == 2, x2 * d1 + x1 * d2,
                                   x2 * x3 * d1 + x1 * x3 * d2 + x1 * x2 * d3))
        gk = np.where(self.has_k, nu * self.c * dirder, 0.0)
        return gz[:self.nz], gk


def stoich_matrix(checkforms, nvar, nreact):
    S = np.zeros((nvar, nreact))
    for i, r, v in checkforms["stoich"]:
        S[i, r] = float(v)
    return S


def adjoint_cell(orc, mono, S, V0, F, RCT, Hs, lam_V, lam_F, Dir=1.0):
    """Gradients of a loss with dL/dC_final = (lam_V, lam_F) with respect to
    (V0, F, RCT) through the accepted steps of sizes Hs."""
    ir = orc.ir
    nvar = ir["nvar"]
    diag = np.asarray(ir["diag"])
    Ys, Y = [], np.asarray(V0, np.float64).copy()
    for H in Hs:
        Ys.append(Y)
        Y = orc.replay(Y, F, RCT, [H], Dir)
    lam = np.asarray(lam_V, np.float64).copy()
    gF = np.asarray(lam_F, np.float64).copy()
    gk = np.zeros(len(RCT))
    for Y, H in zip(reversed(Ys), reversed(list(Hs))):
        dh = Dir * H
        F0 = orc.fun_split(Y, F, RCT)[0]
        Mx = -orc.jac_sp(Y, F, RCT)
        Mx[diag] = Mx[diag] + 1.0 / (dh * GAMMA)
        LU, ier = orc.decomp(Mx)
        assert ier == 0
        K, P = [], []
        for s in range(4):
            if s < 2:
                Ps, fs = Y, F0
            else:
                Ps = Y.copy()
                for j in range(s):
                    Ps = Ps + ROS_A[s * (s - 1) // 2 + j] * K[j]
                fs = orc.fun_split(Ps, F, RCT)[0]
            rhs = fs.copy()
            for j in range(s):
                rhs = rhs + ROS_C[s * (s - 1) // 2 + j] / dh * K[j]
            K.append(orc.solve(LU, rhs))
            P.append(Ps)
        lK = [ROS_M[s] * lam for s in range(4)]
        gY = lam.copy()
        for s in (3, 2, 1, 0):
            mu = orc.solve_tr(LU, lK[s])
            nu = S.T @ mu
            gz, gk1 = mono.first_order(np.concatenate([P[s], F]), RCT, nu)
            lp = gz[:nvar]
            gF += gz[nvar:]
            gk += gk1
            gY += lp
            for j in range(s):
                lK[j] = lK[j] + ROS_C[s * (s - 1) // 2 + j] / dh * mu
                if s >= 2:
                    lK[j] = lK[j] + ROS_A[s * (s - 1) // 2 + j] * lp
            hz, hk = mono.second_order(np.concatenate([Y, F]), RCT, nu, K[s])
            gY += hz[:nvar]
            gF += hz[nvar:]
            gk += hk
        lam = gY
    return lam, gF, gk


def log_loss(C_var, w, atol):
    """sum_i w_i * log(C_i) regularised at the ATOL scale: 1/2 log(C^2 + ATOL^2),
    and its gradient."""
    L = float(np.sum(w * 0.5 * np.log(C_var * C_var + atol * atol)))
    return L, w * C_var / (C_var * C_var + atol * atol)
