// gc_rodas3_kernels.h -- launch interface between the C-ABI host layer and the
// Rodas3 kernels (one translation unit per register-budget variant).
#pragma once
#include <hip/hip_runtime.h>

// Doubles of slot workspace per cell slot: Y, Ynew, Fcn0, 4 stage vectors,
// dFdT (8*NVAR); RCONST copy and the rate scratch (2*NREACT); FIX; LU.
#define GC_WS_DOUBLES_PER_SLOT(nvar, nfix, nreact, nnz) \
  (8L * (nvar) + 2L * (nreact) + (long)(nfix) + (long)(nnz))

// Device pointers, caller (Fortran) layout: C(NSPEC,nCell), RCONST(NREACT,nCell),
// ICNTRL(20,nCell), RCNTRL(20,nCell), Perm(nCell) 1-based, ISTAT(8,nCell),
// RSTAT(3,nCell), IERR(nCell); ATOL/RTOL(NVAR) shared by all cells.
struct GcBatchArgs {
  int nCell;
  double Tstart, Tend;
  double* C;
  const double* RCONST;
  const double* ATOL;
  const double* RTOL;
  const int* ICNTRL;
  const double* RCNTRL;
  const int* Perm;
  int* ISTAT;
  double* RSTAT;
  int* IERR;
};

struct GcMechDims {
  int nvar, nfix, nreact, nnz;
};

#define GC_DECLARE_VARIANT(W)                                                         \
  extern "C" hipError_t gc_launch_plain_w##W(const GcBatchArgs* b, double* ws,       \
                                              unsigned S, int block, hipStream_t st); \
  extern "C" hipError_t gc_launch_persistent_w##W(const GcBatchArgs* b, double* ws,  \
                                                   unsigned S, unsigned* queue,      \
                                                   int block, int cells_per_wave,    \
                                                   int grain, hipStream_t st);       \
  extern "C" void gc_mech_dims_w##W(GcMechDims* d);

GC_DECLARE_VARIANT(1)
GC_DECLARE_VARIANT(2)
