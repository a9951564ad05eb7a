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
  // auto-reduce masks, per cell (NULL: every flag on): DO_FUN(NVAR),
  // DO_JVS(LU_NONZERO), DO_SLV(NVAR) as 32-bit words, flag k in bit k%32 of
  // word k/32, ceil(n/32) words per cell; any non-null selects the masked kernels
  const unsigned* doFun = nullptr;
  const unsigned* doJvs = nullptr;
  const unsigned* doSlv = nullptr;
};

// One batch of a multi-batch launch: its arguments and its first position in
// the launch's queue (the number of cells of the batches before it).
struct GcBatchDesc {
  GcBatchArgs a;
  unsigned q0;
  unsigned* done;  // cells of the batch finished so far (host-visible), or null
};

// Input gate of the overlapped upload of the host-array entry point: the
// batch's inputs arrive in HBM in chunks of `cells` consecutive launch
// positions while the persistent kernel already runs; flags[k] turns nonzero
// (written by the copy engine after chunk k's data) once chunk k has landed.
// A wave hands out a position only once its chunk's flag is seen, waiting at
// most `spin` sleeps of ~3.4 us per chunk; past that it sets *err and takes
// no more cells (every wait ends).
struct GcGate {
  const unsigned* flags;
  unsigned cells;
  unsigned long long spin;
  unsigned* err;
};

// What a compiled variant is: mechanism dimensions, doubles of workspace per
// cell slot (the workspace is slots/64 wave blocks of 64*ws_doubles_per_slot
// doubles), its occupancy target and its LU row-segment size in LDS.
struct GcVariantInfo {
  int nvar, nfix, nreact, nnz;
  long ws_doubles_per_slot;
  long adj_doubles_per_slot;  // workspace of one cell of the adjoint sweep
  int waves_per_eu, seg;
};

// Pool of accepted-step records kept by the recording form of the persistent
// kernel for the discrete adjoint.  Every accepted step gets a record (its H
// and the previous record of the same cell); every `every`-th accepted step
// of a cell (its 1st, (every+1)-th, ...) is a checkpoint whose start state
// Y(NVAR) also goes to the state pool.  The adjoint sweep replays the steps
// between checkpoints.
struct GcRecord {
  double* H;          // [cap]
  int* prev;          // [cap], -1 for a cell's first step
  int* yslot;         // [cap], state-pool slot of a checkpoint record, else -1
  int* head;          // [nCell] (caller's cell index): last record, -1 if none
  double* Y;          // [ycap][NVAR]
  double* K;          // [ycap][4][NVAR] stage vectors of the checkpoint steps, or null
  unsigned* count;    // records taken; > cap means the pool overflowed
  unsigned* ycount;   // states taken; > ycap means the state pool overflowed
  unsigned cap, ycap;
  int every;          // checkpoint interval (>= 1)
};

// Arguments of the adjoint sweep (arrays in the caller's Fortran layout).
struct GcAdjArgs {
  int nCell;
  double Tstart, Tend;
  const double* C;       // (NSPEC, nCell): fixed species are read
  const double* RCONST;  // (NREACT, nCell)
  const int* ICNTRL;     // (20, nCell)
  const int* Perm;       // (nCell), 1-based launch order
  const double* recY;
  const double* recK;    // stage vectors of the checkpoint steps, or null: recomputed
  const double* recH;
  const int* recPrev;
  const int* recYslot;
  const int* head;
  unsigned cap;          // records in the pool (every slot followed is checked)
  unsigned ycap;         // states in the state pool
  int every;             // checkpoint interval of the recording
  long slot_doubles;     // workspace per lane (forward + adjoint + segment)
  const double* lamC;    // (NSPEC, nCell): dL/dC at Tend
  double* gC;            // (NSPEC, nCell): dL/dC at Tstart
  double* gRCONST;       // (NREACT, nCell): dL/dRCONST
  int* IERR;             // (nCell): 1, -8 singular step, -3 non-autonomous,
                         // -5 a record chain that does not fit the pools
};

// One batch of an adjoint sweep: its arguments, its first launch position
// (the number of cells of the batches before it) and its finished-cell
// counter (host-visible, or null).
struct GcAdjDesc {
  GcAdjArgs a;
  unsigned q0;
  unsigned* done;
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
  extern "C" GC_INTERNAL hipError_t gc_launch_persistent_gated_v##V(                  \
      const GcBatchArgs* a, double* ws, unsigned S, unsigned* queue, int block,       \
      int cells_per_wave, int grain, const GcGate* gate, hipStream_t st);             \
  extern "C" GC_INTERNAL hipError_t gc_launch_record_v##V(                            \
      const GcBatchArgs* a, double* ws, unsigned S, unsigned* queue, int block,       \
      int cells_per_wave, int grain, const GcRecord* rec, hipStream_t st);            \
  extern "C" GC_INTERNAL hipError_t gc_launch_persistent_multi_v##V(                  \
      const GcBatchDesc* descs, int nb, unsigned qn, double* ws, unsigned S,           \
      unsigned* queue, int block, int cells_per_wave, int grain, const GcRecord* rec,  \
      hipStream_t st);                                                                 \
  extern "C" GC_INTERNAL hipError_t gc_launch_adjoint_v##V(                            \
      const GcAdjDesc* descs, int nb, unsigned qn, double* ws, long slot_doubles,      \
      unsigned waves, unsigned* queue, int block, hipStream_t st);                     \
  extern "C" GC_INTERNAL void gc_variant_info_v##V(GcVariantInfo* d);

GC_DECLARE_VARIANT(1)
GC_DECLARE_VARIANT(2)
GC_DECLARE_VARIANT(3)
