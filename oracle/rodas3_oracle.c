/*
 * rodas3_oracle.c -- CPU oracle (test infrastructure) for the batched KPP
 * Rosenbrock Rodas3 gas-phase chemistry solve.
 *
 * This file restates, in plain C, the reference's CPU algorithm for the hot
 * path.  It is the checker for the HIP product path and the timed CPU
 * baseline; it is never linked into the product library.  Every floating
 * point expression follows the Fortran evaluation order of the reference:
 * products and sums left to right, no contraction (build with
 * -ffp-contract=off, no -ffast-math), libm pow for Err**(1/3).
 *
 * Reference (paths relative to the reference checkout):
 *   Fun_SPLIT      src/geos-chem/KPP/fullchem/gckpp_Function.F90:2170-5060
 *   Jac_SP         src/geos-chem/KPP/fullchem/gckpp_Jacobian.F90:47-20887
 *   KppDecomp      src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:46-84
 *   KppSolve       src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:645-2311
 *                  (same order as KppSolveIndirect, :175-200)
 *   KppSolveTR     src/geos-chem/KPP/fullchem/gckpp_LinearAlgebra.F90:2327-3701
 *   Integrate_Cell_S src/geos-chem/KPP/fullchem/gckpp_BatchIntegrator.F90:408-707
 *   AddGhinvDiag   src/geos-chem/KPP/fullchem/gckpp_BatchIntegrator.F90:784-795
 *   option merge   gpulib/src/gc_gpu_batch_api.F90:54-110
 *   batch loop     gpulib/src/gc_gpu_batch_api.F90:198-215
 *   OpenMP driver  standalone/kpp_batch.F90:212-227 (schedule(dynamic,16))
 */
#include "rodas3_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

/* gckpp_BatchIntegrator.F90:66 -- 2**-52 */
#define ROUNDOFF 2.220446049250313080847263336181640625e-16
#define DELTAMIN 1.0e-5

/* Evaluate one sum program output.  The first term is negated (exact) when
 * its sign is -1; later terms are added or subtracted in order.  An empty sum
 * is 0. */
static inline double eval_out(const or_sumprog *p, const double *const base[6],
                              int o)
{
    int t0 = p->out_off[o], t1 = p->out_off[o + 1];
    if (t0 == t1) return 0.0;
    double acc = 0.0;
    for (int t = t0; t < t1; ++t) {
        int f0 = p->term_off[t], f1 = p->term_off[t + 1];
        double v = base[p->fac_kind[f0]][p->fac_idx[f0]];
        for (int f = f0 + 1; f < f1; ++f) v = v * base[p->fac_kind[f]][p->fac_idx[f]];
        if (t == t0) acc = (p->term_sign[t] < 0) ? -v : v;
        else if (p->term_sign[t] < 0) acc = acc - v;
        else acc = acc + v;
    }
    return acc;
}

long or_work_doubles(const or_mech *m)
{
    /* Kw(4*NVAR) Ghimj Jac0 (nnz each) Ynew Fcn0 Fcn dFdT Yerr wP wD (NVAR)
     * A(NREACT) B(nb) W(NVAR) Yl(NVAR) RCONSTl(NREACT) FIXl(NFIX) */
    return 4L * m->nvar + 2L * m->nnz + 7L * m->nvar + m->nreact + m->nb +
           m->nvar + m->nvar + m->nreact + m->nfix + 8;
}

/* Flag k of an auto-reduce mask (32 flags per word, flag k in bit k%32 of
 * word k/32); a null mask has every flag on. */
static inline int flag_on(const unsigned *mask, int k)
{
    return mask == NULL || ((mask[k >> 5] >> (k & 31)) & 1u) != 0u;
}

/* Fun_SPLIT (gckpp_Function.F90:2170-5060): rates A(r), then P_VAR and D_VAR
 * sums, then Vdot = P_VAR - D_VAR*V elementwise (:5058).  do_fun: the
 * IF (DO_FUN(i)) guards of both sums (:3280-5056); a species whose flag is
 * off keeps the P[i], D[i] passed in. */
void or_fun_split_m(const or_mech *m, const double *V, const double *F,
                    const double *RCT, double *Vdot, double *P, double *D,
                    double *A, const unsigned *do_fun)
{
    const double *base[6] = {m->lit, A, RCT, V, F, NULL};
    for (int r = 0; r < m->nreact; ++r) A[r] = eval_out(&m->rate, base, r);
    for (int i = 0; i < m->nvar; ++i)
        if (flag_on(do_fun, i)) P[i] = eval_out(&m->prod, base, i);
    for (int i = 0; i < m->nvar; ++i)
        if (flag_on(do_fun, i)) D[i] = eval_out(&m->dest, base, i);
    for (int i = 0; i < m->nvar; ++i) Vdot[i] = P[i] - D[i] * V[i];
}

void or_fun_split(const or_mech *m, const double *V, const double *F,
                  const double *RCT, double *Vdot, double *P, double *D,
                  double *A)
{
    or_fun_split_m(m, V, F, RCT, Vdot, P, D, A, NULL);
}

/* Jac_SP (gckpp_Jacobian.F90:47-20887): B temporaries then JVS sums.
 * do_jvs: the IF (DO_JVS(k)) guards (:3676-20887); an entry whose flag is
 * off keeps the JVS[k] passed in. */
void or_jac_sp_m(const or_mech *m, const double *V, const double *F,
                 const double *RCT, double *JVS, double *B, const unsigned *do_jvs)
{
    const double *base[6] = {m->lit, NULL, RCT, V, F, B};
    for (int k = 0; k < m->nb; ++k) B[k] = eval_out(&m->jacb, base, k);
    for (int k = 0; k < m->nnz; ++k)
        if (flag_on(do_jvs, k)) JVS[k] = eval_out(&m->jvs, base, k);
}

void or_jac_sp(const or_mech *m, const double *V, const double *F,
               const double *RCT, double *JVS, double *B)
{
    or_jac_sp_m(m, V, F, RCT, JVS, B, NULL);
}

/* KppDecomp (gckpp_LinearAlgebra.F90:46-84): in-place unpivoted sparse LU in
 * IKJ order.  The singularity test looks at the row's diagonal entry before
 * that row is eliminated, exactly as the reference does (:61). */
void or_decomp(const or_mech *m, double *JVS, double *W, int *ier)
{
    const int *crow = m->crow, *icol = m->icol, *diag = m->diag;
    *ier = 0;
    for (int k = 0; k < m->nvar; ++k) {
        if (fabs(JVS[diag[k]]) < DBL_MIN) { *ier = k + 1; return; }
        for (int kk = crow[k]; kk < crow[k + 1]; ++kk) W[icol[kk]] = JVS[kk];
        for (int kk = crow[k]; kk < diag[k]; ++kk) {
            int j = icol[kk];
            double a = -W[j] / JVS[diag[j]];
            W[j] = -a;
            for (int jj = diag[j] + 1; jj < crow[j + 1]; ++jj)
                W[icol[jj]] = W[icol[jj]] + a * JVS[jj];
        }
        for (int kk = crow[k]; kk < crow[k + 1]; ++kk) JVS[kk] = W[icol[kk]];
    }
}

/* KppSolve (gckpp_LinearAlgebra.F90:645-2311): forward L (unit diagonal) in
 * row order, then backward U rows NVAR..1 with one divide per row.  do_slv:
 * the IF (DO_SLV(i)) guard of each row's statement (:653-2309); a row whose
 * flag is off leaves X[i] as it is. */
void or_solve_m(const or_mech *m, const double *JVS, double *X, const unsigned *do_slv)
{
    const int *crow = m->crow, *icol = m->icol, *diag = m->diag;
    for (int i = 0; i < m->nvar; ++i) {
        if (!flag_on(do_slv, i)) continue;
        double x = X[i];
        for (int j = crow[i]; j < diag[i]; ++j) x = x - JVS[j] * X[icol[j]];
        X[i] = x;
    }
    for (int i = m->nvar - 1; i >= 0; --i) {
        if (!flag_on(do_slv, i)) continue;
        double x = X[i];
        for (int j = diag[i] + 1; j < crow[i + 1]; ++j) x = x - JVS[j] * X[icol[j]];
        X[i] = x / JVS[diag[i]];
    }
}

void or_solve(const or_mech *m, const double *JVS, double *X)
{
    or_solve_m(m, JVS, X, NULL);
}

/* KppSolveTR (gckpp_LinearAlgebra.F90:2327-3701): (LU)^T XX = X.  Phase 1
 * solves U^T in ascending rows, each sum over U(j,i) for ascending j; phase 2
 * solves L^T in descending rows, each sum over L(j,i) for ascending j. */
void or_solve_tr(const or_mech *m, const double *JVS, const double *X,
                 double *XX)
{
    const int n = m->nvar;
    const int *crow = m->crow, *icol = m->icol, *diag = m->diag;
    /* column lists built on the fly (oracle: clarity over speed) */
    for (int i = 0; i < n; ++i) {
        double x = X[i];
        for (int j = 0; j < i; ++j)
            for (int kk = diag[j] + 1; kk < crow[j + 1]; ++kk)
                if (icol[kk] == i) x = x - JVS[kk] * XX[j];
        XX[i] = x / JVS[diag[i]];
    }
    for (int i = n - 1; i >= 0; --i) {
        double x = XX[i];
        for (int j = i + 1; j < n; ++j)
            for (int kk = crow[j]; kk < diag[j]; ++kk)
                if (icol[kk] == i) x = x - JVS[kk] * XX[j];
        XX[i] = x;
    }
}

/* x**(1/3) as the reference writes it: Err**(ONE/ros_ELO) with the exponent
 * the double nearest 1/3.  Mode 0 is libm pow, as the reference CPU build
 * (gfortran) calls it; mode 1 is the correctly rounded value, from a
 * 113-bit quad-precision powq rounded once to double. */
static inline double pow_third(const or_mech *m, double x)
{
    if (m->pow_mode == 1) return (double)powq((__float128)x, (__float128)(1.0 / 3.0));
    return pow(x, 1.0 / 3.0);
}

static inline double fmin_f(double a, double b) { return (a < b) ? a : b; }
static inline double fmax_f(double a, double b) { return (a > b) ? a : b; }

/* Integrate_Cell_S (gckpp_BatchIntegrator.F90:408-707), called with the
 * option merge of gpu_solve_cell (gc_gpu_batch_api.F90:79-88). */
void or_solve_cell(const or_mech *m, double *Y, const double *FIX,
                   const double *RCONST, double Tstart, double Tend,
                   const double *AbsTol, const double *RelTol,
                   const int *ICNTRL_U, const double *RCNTRL_U,
                   int *ISTAT, double *RSTAT, int *IERR, double *work)
{
    or_solve_cell_rec(m, Y, FIX, RCONST, Tstart, Tend, AbsTol, RelTol, ICNTRL_U,
                      RCNTRL_U, ISTAT, RSTAT, IERR, work, NULL, 0, NULL);
}

/* As or_solve_cell; additionally the step sizes H of the accepted steps (the
 * stepper's own H, sign included), in order, go to Hrec[0..maxrec) and their
 * number to *nrec (when Hrec is not NULL). */
void or_solve_cell_rec(const or_mech *m, double *Y, const double *FIX,
                       const double *RCONST, double Tstart, double Tend,
                       const double *AbsTol, const double *RelTol,
                       const int *ICNTRL_U, const double *RCNTRL_U,
                       int *ISTAT, double *RSTAT, int *IERR, double *work,
                       double *Hrec, int maxrec, int *nrec)
{
    or_solve_cell_m(m, Y, FIX, RCONST, Tstart, Tend, AbsTol, RelTol, ICNTRL_U, RCNTRL_U,
                    ISTAT, RSTAT, IERR, work, Hrec, maxrec, nrec, NULL, NULL, NULL);
}

/* The cell solve with its auto-reduce masks (null: every flag on).  The
 * workspace the guarded statements leave alone -- P_VAR, D_VAR (wP, wD) and
 * Jac0 -- holds zeros from the cell's start. */
void or_solve_cell_m(const or_mech *m, double *Y, const double *FIX,
                     const double *RCONST, double Tstart, double Tend,
                     const double *AbsTol, const double *RelTol,
                     const int *ICNTRL_U, const double *RCNTRL_U,
                     int *ISTAT, double *RSTAT, int *IERR, double *work,
                     double *Hrec, int maxrec, int *nrec, const unsigned *do_fun,
                     const unsigned *do_jvs, const unsigned *do_slv)
{
    if (Hrec) *nrec = 0;
    static const double ros_A[6] = {0.0, 2.0, 0.0, 2.0, 0.0, 1.0};
    static const double ros_C[6] = {4.0, 1.0, -1.0, 1.0, -1.0, -(8.0 / 3.0)};
    static const int ros_NewF[4] = {1, 0, 1, 1};
    static const double ros_M[4] = {2.0, 0.0, 1.0, 1.0};
    static const double ros_E[4] = {0.0, 0.0, 0.0, 1.0};
    static const double ros_Gamma[4] = {0.5, 1.5, 0.0, 0.0};
    const int NVAR = m->nvar, NNZ = m->nnz;

    int ICNTRL[20];
    double RCNTRL[20];
    for (int k = 0; k < 20; ++k) { ICNTRL[k] = 0; RCNTRL[k] = 0.0; }
    ICNTRL[14] = 5;
    for (int k = 0; k < 20; ++k) {
        if (ICNTRL_U[k] != 0) ICNTRL[k] = ICNTRL_U[k];
        if (RCNTRL_U[k] > 0.0) RCNTRL[k] = RCNTRL_U[k];
    }

    double *Kw = work;
    double *Ghimj = Kw + 4 * NVAR;
    double *Jac0 = Ghimj + NNZ;
    double *Ynew = Jac0 + NNZ;
    double *Fcn0 = Ynew + NVAR;
    double *Fcn = Fcn0 + NVAR;
    double *dFdT = Fcn + NVAR;
    double *Yerr = dFdT + NVAR;
    double *wP = Yerr + NVAR;
    double *wD = wP + NVAR;
    double *A = wD + NVAR;
    double *B = A + m->nreact;
    double *W = B + m->nb;
    memset(Jac0, 0, sizeof(double) * (size_t)NNZ);
    memset(wP, 0, sizeof(double) * (size_t)NVAR);
    memset(wD, 0, sizeof(double) * (size_t)NVAR);

    int Nsteps = 0;
    *IERR = 0;
    for (int k = 0; k < 8; ++k) ISTAT[k] = 0;
    for (int k = 0; k < 3; ++k) RSTAT[k] = 0.0;

    if (ICNTRL[11] == 1) { *IERR = -90; return; }
    const int Autonomous = !(ICNTRL[0] == 0);
    const int VectorTol = (ICNTRL[1] == 0);
    if (!(ICNTRL[2] == 0 || ICNTRL[2] == 4)) { *IERR = -2; return; }
    const int Max_no_steps = (ICNTRL[3] == 0) ? 200000 : ICNTRL[3];

    const double Hmin = (RCNTRL[0] > 0.0) ? RCNTRL[0] : 0.0;
    const double Hmax = (RCNTRL[1] > 0.0) ? fmin_f(fabs(RCNTRL[1]), fabs(Tend - Tstart))
                                          : fabs(Tend - Tstart);
    const double Hstart = (RCNTRL[2] > 0.0) ? fmin_f(fabs(RCNTRL[2]), fabs(Tend - Tstart))
                                            : fmax_f(Hmin, DELTAMIN);
    const double FacMin = (RCNTRL[3] > 0.0) ? RCNTRL[3] : 0.2;
    const double FacMax = (RCNTRL[4] > 0.0) ? RCNTRL[4] : 6.0;
    const double FacRej = (RCNTRL[5] > 0.0) ? RCNTRL[5] : 0.1;
    const double FacSafe = (RCNTRL[6] > 0.0) ? RCNTRL[6] : 0.9;

    double T = Tstart;
    double H = fmin_f(fmax_f(fabs(Hmin), fabs(Hstart)), fabs(Hmax));
    if (fabs(H) <= 10.0 * ROUNDOFF) H = DELTAMIN;
    const int Direction = (Tend >= Tstart) ? +1 : -1;
    const double Dir = (double)Direction;
    H = Dir * H;
    int RejectLastH = 0, RejectMoreH = 0;

    while ((Direction > 0 && ((T - Tend) + ROUNDOFF <= 0.0)) ||
           (Direction < 0 && ((Tend - T) + ROUNDOFF <= 0.0))) {
        if (Nsteps > Max_no_steps) { *IERR = -6; return; }
        if (((T + 0.1 * H) == T) || (H <= ROUNDOFF)) { *IERR = -7; return; }
        H = fmin_f(H, fabs(Tend - T));

        or_fun_split_m(m, Y, FIX, RCONST, Fcn0, wP, wD, A, do_fun);
        ISTAT[0] += 1;
        if (!Autonomous) {
            double Delta = sqrt(ROUNDOFF) * fmax_f(1.0e-6, fabs(T));
            or_fun_split_m(m, Y, FIX, RCONST, dFdT, wP, wD, A, do_fun);
            ISTAT[0] += 1;
            for (int k = 0; k < NVAR; ++k) dFdT[k] = (dFdT[k] - Fcn0[k]) / Delta;
        }
        or_jac_sp_m(m, Y, FIX, RCONST, Jac0, B, do_jvs);
        ISTAT[1] += 1;

        for (;;) { /* UntilAccepted */
            int Nconsecutive = 0;
            int Singular = 1;
            while (Singular) {
                for (int k = 0; k < NNZ; ++k) Ghimj[k] = -Jac0[k];
                double ghinv = 1.0 / ((Dir * H) * ros_Gamma[0]);
                for (int i = 0; i < NVAR; ++i)
                    Ghimj[m->diag[i]] = Ghimj[m->diag[i]] + ghinv;
                int ising;
                or_decomp(m, Ghimj, W, &ising);
                ISTAT[5] += 1;
                if (ising == 0) {
                    Singular = 0;
                } else {
                    ISTAT[7] += 1;
                    Nconsecutive += 1;
                    if (Nconsecutive <= 5) H = H * 0.5;
                    else { *IERR = -8; return; }
                }
            }

            for (int istage = 1; istage <= 4; ++istage) {
                double *Ks = Kw + NVAR * (istage - 1);
                if (istage == 1) {
                    for (int k = 0; k < NVAR; ++k) Fcn[k] = Fcn0[k];
                } else if (ros_NewF[istage - 1]) {
                    for (int k = 0; k < NVAR; ++k) Ynew[k] = Y[k];
                    for (int j = 1; j <= istage - 1; ++j) {
                        double aa = ros_A[(istage - 1) * (istage - 2) / 2 + j - 1];
                        for (int k = 0; k < NVAR; ++k)
                            Ynew[k] = Ynew[k] + aa * Kw[NVAR * (j - 1) + k];
                    }
                    or_fun_split_m(m, Ynew, FIX, RCONST, Fcn, wP, wD, A, do_fun);
                    ISTAT[0] += 1;
                }
                for (int k = 0; k < NVAR; ++k) Ks[k] = Fcn[k];
                for (int j = 1; j <= istage - 1; ++j) {
                    double HC = ros_C[(istage - 1) * (istage - 2) / 2 + j - 1] / (Dir * H);
                    for (int k = 0; k < NVAR; ++k)
                        Ks[k] = Ks[k] + HC * Kw[NVAR * (j - 1) + k];
                }
                if (!Autonomous && ros_Gamma[istage - 1] != 0.0) {
                    double HG = (Dir * H) * ros_Gamma[istage - 1];
                    for (int k = 0; k < NVAR; ++k) Ks[k] = Ks[k] + HG * dFdT[k];
                }
                or_solve_m(m, Ghimj, Ks, do_slv);
                ISTAT[6] += 1;
            }

            for (int k = 0; k < NVAR; ++k) Ynew[k] = Y[k];
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < NVAR; ++k)
                    Ynew[k] = Ynew[k] + ros_M[j] * Kw[NVAR * j + k];
            for (int k = 0; k < NVAR; ++k) Yerr[k] = 0.0;
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < NVAR; ++k)
                    Yerr[k] = Yerr[k] + ros_E[j] * Kw[NVAR * j + k];

            double Err = 0.0;
            for (int i = 0; i < NVAR; ++i) {
                double Ymax = fmax_f(fabs(Y[i]), fabs(Ynew[i]));
                double Scal = VectorTol ? AbsTol[i] + RelTol[i] * Ymax
                                        : AbsTol[0] + RelTol[0] * Ymax;
                double q = Yerr[i] / Scal;
                Err = Err + q * q;
            }
            Err = sqrt(Err / (double)NVAR);
            Err = fmax_f(Err, 1.0e-10);

            double Fac = fmin_f(FacMax, fmax_f(FacMin, FacSafe / pow_third(m, Err)));
            double Hnew = H * Fac;

            Nsteps += 1;
            ISTAT[2] += 1;
            if ((Err <= 1.0) || (H <= Hmin)) {
                ISTAT[3] += 1;
                if (Hrec) {
                    if (*nrec < maxrec) Hrec[*nrec] = H;
                    *nrec += 1;
                }
                if (ICNTRL[15] == 1) {
                    for (int k = 0; k < NVAR; ++k) Y[k] = fmax_f(Ynew[k], 0.0);
                } else {
                    for (int k = 0; k < NVAR; ++k) Y[k] = Ynew[k];
                }
                T = T + Dir * H;
                Hnew = fmax_f(Hmin, fmin_f(Hnew, Hmax));
                if (RejectLastH) Hnew = fmin_f(Hnew, H);
                RSTAT[1] = H;
                RSTAT[2] = Hnew;
                RSTAT[0] = T;
                RejectLastH = 0;
                RejectMoreH = 0;
                H = Hnew;
                break;
            } else {
                if (RejectMoreH) Hnew = H * FacRej;
                RejectMoreH = RejectLastH;
                RejectLastH = 1;
                H = Hnew;
                if (ISTAT[3] >= 1) ISTAT[4] += 1;
            }
        }
    }
    *IERR = 1;
}

/* The accepted steps of Integrate_Cell_S with their step sizes given: the
 * discrete map that the Rodas3 discrete adjoint differentiates (step sizes and
 * the accept/reject sequence held fixed).  Autonomous form (ICNTRL(1) /= 0,
 * the GEOS-Chem setting); each step is the stage sequence, the prepared matrix
 * and the update of gckpp_BatchIntegrator.F90:565-690 with the stepper's H =
 * Hs[k] (as recorded by or_solve_cell_rec), clamped at 0 when clamp is set.
 * *ier = 0, or k+1 if the matrix of step k is singular. */
void or_replay_cell(const or_mech *m, double *Y, const double *FIX,
                    const double *RCONST, double Dir, int nsteps,
                    const double *Hs, int clamp, double *work, int *ier)
{
    static const double ros_A[6] = {0.0, 2.0, 0.0, 2.0, 0.0, 1.0};
    static const double ros_C[6] = {4.0, 1.0, -1.0, 1.0, -1.0, -(8.0 / 3.0)};
    static const double ros_M[4] = {2.0, 0.0, 1.0, 1.0};
    static const double ros_Gamma[4] = {0.5, 1.5, 0.0, 0.0};
    const int NVAR = m->nvar, NNZ = m->nnz;
    double *Kw = work;
    double *Ghimj = Kw + 4 * NVAR;
    double *Ynew = Ghimj + NNZ;
    double *Fcn0 = Ynew + NVAR;
    double *Fcn = Fcn0 + NVAR;
    double *wP = Fcn + NVAR;
    double *wD = wP + NVAR;
    double *A = wD + NVAR;
    double *B = A + m->nreact;
    double *W = B + m->nb;
    *ier = 0;
    for (int st = 0; st < nsteps; ++st) {
        const double H = Hs[st];
        or_fun_split(m, Y, FIX, RCONST, Fcn0, wP, wD, A);
        or_jac_sp(m, Y, FIX, RCONST, Ghimj, B);
        for (int k = 0; k < NNZ; ++k) Ghimj[k] = -Ghimj[k];
        const double ghinv = 1.0 / ((Dir * H) * ros_Gamma[0]);
        for (int i = 0; i < NVAR; ++i) Ghimj[m->diag[i]] = Ghimj[m->diag[i]] + ghinv;
        int ising;
        or_decomp(m, Ghimj, W, &ising);
        if (ising) { *ier = st + 1; return; }
        for (int istage = 1; istage <= 4; ++istage) {
            double *Ks = Kw + NVAR * (istage - 1);
            if (istage == 1 || istage == 2) {
                for (int k = 0; k < NVAR; ++k) Fcn[k] = Fcn0[k];
            } else {
                for (int k = 0; k < NVAR; ++k) Ynew[k] = Y[k];
                for (int j = 1; j <= istage - 1; ++j) {
                    double aa = ros_A[(istage - 1) * (istage - 2) / 2 + j - 1];
                    for (int k = 0; k < NVAR; ++k) Ynew[k] = Ynew[k] + aa * Kw[NVAR * (j - 1) + k];
                }
                or_fun_split(m, Ynew, FIX, RCONST, Fcn, wP, wD, A);
            }
            for (int k = 0; k < NVAR; ++k) Ks[k] = Fcn[k];
            for (int j = 1; j <= istage - 1; ++j) {
                double HC = ros_C[(istage - 1) * (istage - 2) / 2 + j - 1] / (Dir * H);
                for (int k = 0; k < NVAR; ++k) Ks[k] = Ks[k] + HC * Kw[NVAR * (j - 1) + k];
            }
            or_solve(m, Ghimj, Ks);
        }
        for (int k = 0; k < NVAR; ++k) Ynew[k] = Y[k];
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < NVAR; ++k) Ynew[k] = Ynew[k] + ros_M[j] * Kw[NVAR * j + k];
        for (int k = 0; k < NVAR; ++k) Y[k] = clamp ? fmax_f(Ynew[k], 0.0) : Ynew[k];
    }
}

/* gc_gpu_integrate_batch_dev semantics on the CPU (gc_gpu_batch_api.F90:
 * 198-215): cell ic solves column Perm(ic) (1-based) of the passed arrays,
 * with local copies of Y, FIX and RCONST (gpu_solve_cell, :90-109).  The loop
 * is OpenMP schedule(dynamic,16) as in standalone/kpp_batch.F90:212-227. */
double or_integrate_batch(const or_mech *m, int nCell, double Tstart,
                          double Tend, double *C, const double *RCONST,
                          const double *ATOL, const double *RTOL,
                          const int *ICNTRL, const double *RCNTRL,
                          const int *Perm, int *ISTAT, double *RSTAT,
                          int *IERR, int nthreads)
{
    return or_integrate_batch_m(m, nCell, Tstart, Tend, C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL,
                                Perm, ISTAT, RSTAT, IERR, nthreads, NULL, NULL, NULL);
}

/* As or_integrate_batch with per-cell auto-reduce masks (column idx of the
 * caller's arrays has its DO_FUN words at do_fun + idx*ceil(NVAR/32), DO_JVS
 * at do_jvs + idx*ceil(LU_NONZERO/32), DO_SLV at do_slv + idx*ceil(NVAR/32);
 * a null array: every flag on). */
double or_integrate_batch_m(const or_mech *m, int nCell, double Tstart,
                            double Tend, double *C, const double *RCONST,
                            const double *ATOL, const double *RTOL,
                            const int *ICNTRL, const double *RCNTRL,
                            const int *Perm, int *ISTAT, double *RSTAT,
                            int *IERR, int nthreads, const unsigned *do_fun,
                            const unsigned *do_jvs, const unsigned *do_slv)
{
    const long wv = (m->nvar + 31) / 32, wj = (m->nnz + 31) / 32;
    if (nCell <= 0) return 0.0;
    const int nspec = m->nvar + m->nfix;
    const long nw = or_work_doubles(m);
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    double t0 = omp_get_wtime();
#pragma omp parallel num_threads(nthreads)
    {
        double *work = (double *)malloc(sizeof(double) * (size_t)nw);
        double *Yl = work + (nw - (m->nvar + m->nreact + m->nfix + 8));
        double *RCl = Yl + m->nvar;
        double *FIXl = RCl + m->nreact;
#pragma omp for schedule(dynamic, 16)
        for (int ic = 0; ic < nCell; ++ic) {
            long idx = (long)Perm[ic] - 1;
            double *Cc = C + idx * nspec;
            memcpy(Yl, Cc, sizeof(double) * m->nvar);
            memcpy(FIXl, Cc + m->nvar, sizeof(double) * m->nfix);
            memcpy(RCl, RCONST + idx * m->nreact, sizeof(double) * m->nreact);
            or_solve_cell_m(m, Yl, FIXl, RCl, Tstart, Tend, ATOL, RTOL,
                            ICNTRL + idx * 20, RCNTRL + idx * 20,
                            ISTAT + idx * 8, RSTAT + idx * 3, IERR + idx, work,
                            NULL, 0, NULL, do_fun ? do_fun + idx * wv : NULL,
                            do_jvs ? do_jvs + idx * wj : NULL,
                            do_slv ? do_slv + idx * wv : NULL);
            memcpy(Cc, Yl, sizeof(double) * m->nvar);
        }
        free(work);
    }
    return omp_get_wtime() - t0;
}
