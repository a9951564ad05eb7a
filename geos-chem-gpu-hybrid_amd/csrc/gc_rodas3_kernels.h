// gc_rodas3_kernels.h -- launch interface between the C-ABI host layer and the
// Rodas3 kernels (one translation unit per register-budget variant).
#pragma once
#include <hip/hip_runtime.h>

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

// What a compiled variant is: mechanism dimensions, doubles of workspace per
// cell slot (the workspace is slots/64 wave blocks of 64*ws_doubles_per_slot
// doubles), its occupancy target and its LU row-segment size in LDS.
struct GcVariantInfo {
  int nvar, nfix, nreact, nnz;
  long ws_doubles_per_slot;
  int waves_per_eu, seg;
};

// Internal to the library (not exported): the C-ABI in include/gcchemgpu.h is
// the only public surface.
#define GC_INTERNAL __attribute__((visibility("hidden")))
#define GC_DECLARE_VARIANT(V)                                                          \
  extern "C" GC_INTERNAL hipError_t gc_launch_plain_v##V(const GcBatchArgs* a, double* ws,        \
                                              int block, hipStream_t st);             \
  extern "C" GC_INTERNAL hipError_t gc_launch_persistent_v##V(const GcBatchArgs* a, double* ws,   \
                                                   unsigned S, unsigned* queue,       \
                                                   int block, int cells_per_wave,     \
                                                   int grain, hipStream_t st);        \
  extern "C" GC_INTERNAL void gc_variant_info_v##V(GcVariantInfo* d);

GC_DECLARE_VARIANT(1)
GC_DECLARE_VARIANT(2)
