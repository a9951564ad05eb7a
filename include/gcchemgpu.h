/*
 * gcchemgpu.h -- C ABI of libgcchemgpu.so, the MI355X (gfx950) batched KPP
 * Rosenbrock Rodas3 gas-phase chemistry solver.
 *
 * The first five symbols are drop-in replacements for the entry points that
 * GEOS-Chem's fullchem driver binds through ISO_C_BINDING
 * (src/geos-chem/GeosCore/fullchem_mod.F90:153-212); their argument meaning,
 * layouts and error behaviour are those of the reference library.  Arrays are
 * Fortran-contiguous (the per-cell vector is the fastest index); every array
 * argument is passed by reference, the scalars marked "by value" by value.
 *
 * The gc_amd_* symbols are additions of this library (no counterpart in the
 * reference): a device-resident form of the batch solve for callers that keep
 * the batch in HBM (the PyTorch binding, benchmarks), launch tuning, the
 * discrete adjoint, the cost order and hybrid CPU/GPU split that the
 * reference keeps in its caller, and a query of the compiled mechanism's
 * dimensions.
 */
#ifndef GCCHEMGPU_H
#define GCCHEMGPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Batch solve on the calling host thread's current device.
 * Replaces gc_gpu_integrate_batch (gpulib/src/gc_gpu_batch_api.F90:115-136).
 *   nCell, Tstart, Tend      by value
 *   C(NSPEC,nCell)           inout: VAR = 1:NVAR, FIX = NVAR+1:NSPEC
 *   RCONST(NREACT,nCell)     in
 *   ATOL(NVAR), RTOL(NVAR)   in, shared by all cells
 *   ICNTRL(20,nCell)         in, merged per cell as the stock INTEGRATE does
 *   RCNTRL(20,nCell)         in
 *   Perm(nCell)              in, 1-based: cell ic of the launch solves column
 *                            Perm(ic) of the passed arrays (launch/cost order)
 *   ISTAT(8,nCell)           out: Nfun Njac Nstp Nacc Nrej Ndec Nsol Nsng
 *   RSTAT(3,nCell)           out: Texit Hexit Hnew
 *   IERR(nCell)              out: 1 ok, -2 method, -6 max steps, -7 step too
 *                            small, -8 singular six times, -90 auto-reduce,
 *                            -99 device selection failed (all cells), -98 the
 *                            device solve could not run (all cells; see
 *                            gc_amd_last_error)
 *   tTotal, tKernel          out: seconds with and without host<->device copies
 * nCell <= 0 returns at once with both timers 0.  Blocking. */
void gc_gpu_integrate_batch(int nCell, double Tstart, double Tend, double *C,
                            const double *RCONST, const double *ATOL,
                            const double *RTOL, const int *ICNTRL,
                            const double *RCNTRL, const int *Perm, int *ISTAT,
                            double *RSTAT, int *IERR, double *tTotal,
                            double *tKernel);

/* As gc_gpu_integrate_batch on device `dev` (dev < 0 keeps the calling
 * thread's current device); if selecting `dev` fails every IERR is -99.
 * Replaces gc_gpu_integrate_batch_dev (gpulib/src/gc_gpu_batch_api.F90:147-216).
 * Host threads may call it concurrently with different devices. */
void gc_gpu_integrate_batch_dev(int dev, int nCell, double Tstart, double Tend,
                                double *C, const double *RCONST,
                                const double *ATOL, const double *RTOL,
                                const int *ICNTRL, const double *RCNTRL,
                                const int *Perm, int *ISTAT, double *RSTAT,
                                int *IERR, double *tTotal, double *tKernel);

/* Number of visible devices and the name of device 0 (NUL-terminated,
 * truncated to nameLen-1).  Replaces gc_gpu_info
 * (gpulib/src/gc_gpu_devinfo.c:23-50); nDev = 0 when no device is usable. */
void gc_gpu_info(int *nDev, char *devName, int nameLen);

/* Make `dev` the calling host thread's device; returns the device now current
 * or -1 on failure.  Replaces gc_gpu_set_device (gpulib/src/gc_gpu_devinfo.c:67-86). */
int gc_gpu_set_device(int dev);

/* Solve one dummy cell (C = 1e6, RCONST = 0, ATOL 1e-2, RTOL 0.5e-2,
 * ICNTRL(1,2,3,7,15) = (1,0,4,1,-1), [0,1200] s) on the current device:
 * rc = its IERR (1 on success), nsteps = ISTAT(3).  The reference answer is
 * rc = 1, nsteps = 12.  Replaces gc_gpu_selftest
 * (gpulib/src/gc_gpu_batch_api.F90:349-376). */
void gc_gpu_selftest(int *rc, int *nsteps, double *tTotal);

/* ---- additions (no counterpart in the reference) ------------------------ */

/* Kernel forms of the device solve. */
#define GC_AMD_KERNEL_PLAIN 0      /* one thread per cell, nested loops     */
#define GC_AMD_KERNEL_PERSISTENT 1 /* work queue + flattened state machine  */

/* Batch solve with every array already in device memory (same layouts and
 * meaning as gc_gpu_integrate_batch).  `stream` is the caller's hipStream_t
 * (NULL: the null stream): the solve starts after all work already queued on
 * it, on the library's own stream of the current device.  Returns 0 on success or a
 * negative code: -1 bad argument (nCell < 0, Perm entry outside 1..nCell,
 * unknown kernel), -2 device memory, -3 launch or runtime error (details in
 * gc_amd_last_error).  *tKernel (may be NULL) receives the device time of the
 * solve in seconds.  Blocking: returns after the solve has finished.
 * chunk > 0 solves the launch order in chunks of that many cells, each chunk
 * with workspace for `chunk` slots only (the low-memory mode of the reference,
 * gpulib/src/gc_gpu_batch_api.F90:218-343); the chunk loop is replayed as one
 * captured HIP graph. */
int gc_amd_integrate_batch_device(int kernel, int chunk, int nCell,
                                  double Tstart, double Tend, double *C,
                                  const double *RCONST, const double *ATOL,
                                  const double *RTOL, const int *ICNTRL,
                                  const double *RCNTRL, const int *Perm,
                                  int *ISTAT, double *RSTAT, int *IERR,
                                  void *stream, double *tKernel);

/* Allocation-free form of gc_amd_integrate_batch_device (chunk = 0): the
 * same arguments, results and return codes, with every device buffer the
 * solve needs beyond the caller's inputs and outputs -- the work-queue
 * counter, the launch-order check flag and the workspace slots -- taken from
 * a 256 MiB scratch array that each device holds from the library's first
 * call on it until the process ends.  The call itself allocates nothing (no
 * hipMalloc, hipFree or allocating library routine); the first call of the
 * library on a device, of any entry point, sets up the device's stream and
 * events and finds its scratch address.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).
 * Limits: the persistent kernel solves any nCell with the resident grid the
 * scratch holds (floor((256 MiB - 4 KiB) / (512 * W)) waves of 64 slots,
 * W = gc_amd_workspace_doubles_per_slot(): 40 waves for fullchem), so it is
 * far slower than gc_amd_integrate_batch_device on large batches; the plain
 * kernel needs one slot per cell and returns -1 when nCell exceeds
 * 64 x that many waves (2,560 cells for fullchem).  Every call on a device
 * uses the same scratch, which keeps the last call's contents (each call
 * clears the counter and the flag it needs); two calls that overlapped on one
 * device would race on it, so the library serialises the calls of a device:
 * a call made while another is running on the same device waits for it.
 * Returns 0, -1 (bad argument, Perm entry outside 1..nCell, nCell beyond the
 * plain kernel's limit), -3 (runtime error; gc_amd_last_error).  Blocking;
 * ordered after `stream`. */
int gc_amd_integrate_batch_noalloc(int kernel, int nCell, double Tstart,
                                   double Tend, double *C, const double *RCONST,
                                   const double *ATOL, const double *RTOL,
                                   const int *ICNTRL, const double *RCNTRL,
                                   const int *Perm, int *ISTAT, double *RSTAT,
                                   int *IERR, void *stream, double *tKernel);

/* Batch solve with the cells' auto-reduce flags, arrays in device memory as
 * for gc_amd_integrate_batch_device (chunk = 0).  The reference's generated
 * mechanism code guards each species' production and loss sums with
 * DO_FUN(i) (gckpp_Function.F90:29, 3280-5056), each Jacobian entry with
 * DO_JVS(k) (gckpp_Jacobian.F90:31, 3676-20887) and each KppSolve row with
 * DO_SLV(i) (gckpp_LinearAlgebra.F90:31, 653-2309), all .TRUE. in the batch
 * build; here they are per-cell data.  A flag that is off skips its statement
 * exactly as the guard does: P_VAR(i)/D_VAR(i) and JVS(k) keep the value the
 * cell's workspace holds, which is zero from the cell's start (so a masked
 * species has Vdot = 0 - 0*V and a masked entry Ghimj = -0.0, + ghinv on the
 * diagonal), and a masked row leaves X(i) as it is.  Masks are 32-bit words,
 * flag k (0-based) in bit k%32 of word k/32, per cell (cell-major, the
 * caller's column index): doFun ceil(NVAR/32) words, doJvs
 * ceil(LU_NONZERO/32), doSlv ceil(NVAR/32).  A null mask means every flag
 * on.  The flags are the caller's choice: choosing them (the stock
 * integrator's auto-reduce threshold rule) stays with the caller, and
 * ICNTRL(12) = 1 still returns IERR -90.  No counterpart in the reference's
 * batch library, whose path has no auto-reduce.  Returns as
 * gc_amd_integrate_batch_device. */
int gc_amd_integrate_batch_masked(int kernel, int nCell, double Tstart,
                                  double Tend, double *C, const double *RCONST,
                                  const double *ATOL, const double *RTOL,
                                  const int *ICNTRL, const double *RCNTRL,
                                  const int *Perm, const unsigned *doFun,
                                  const unsigned *doJvs, const unsigned *doSlv,
                                  int *ISTAT, double *RSTAT, int *IERR,
                                  void *stream, double *tKernel);

/* ---- several batches in one launch ---------------------------------------
 * Independent batches -- each with its own arrays, size, launch order and
 * time span, each with the meaning of one gc_amd_integrate_batch_device call
 * (the per-call semantics of gpulib/src/gc_gpu_batch_api.F90:147-216) --
 * solved by one persistent launch whose work queue runs through batch 0's
 * launch order, then batch 1's, and so on, so lanes freed in one batch take
 * cells of the next.  No counterpart in the reference (one batch per call). */
typedef struct GcAmdBatchDesc {
  int nCell;
  double Tstart, Tend;
  double *C;                      /* (NSPEC, nCell) inout, device memory */
  const double *RCONST;           /* (NREACT, nCell) */
  const double *ATOL, *RTOL;      /* (NVAR) */
  const int *ICNTRL;              /* (20, nCell) */
  const double *RCNTRL;           /* (20, nCell) */
  const int *Perm;                /* (nCell), 1-based launch order */
  int *ISTAT;                     /* (8, nCell) out */
  double *RSTAT;                  /* (3, nCell) out */
  int *IERR;                      /* (nCell) out */
} GcAmdBatchDesc;

/* Progress of a multi-batch launch: batches whose cells have all finished,
 * cells finished, cells in all batches. */
typedef void (*gc_amd_progress_fn)(int batchesDone, long long cellsDone,
                                   long long cellsTotal, void *user);

/* Solve nBatch (1..4096) batches in one persistent launch.  progress (may be
 * NULL) is called on the calling thread whenever the number of finished
 * cells has changed, polled every 0.2 s while the launch runs, and once at
 * its end.  Returns 0 or the codes of gc_amd_integrate_batch_device (-1 also
 * for a bad descriptor). */
int gc_amd_integrate_batches_device(int nBatch, const GcAmdBatchDesc *batch,
                                    gc_amd_progress_fn progress, void *user,
                                    void *stream, double *tKernel);

/* The recording form of gc_amd_integrate_batches_device for the discrete
 * adjoint: one record pool for all batches (arguments as in
 * gc_amd_record_batch_device, progress as above); head has one entry per
 * cell of all batches, batch k's cell c (0-based column) at
 * head[n_0 + ... + n_{k-1} + c]. */
int gc_amd_record_batches_device(int nBatch, const GcAmdBatchDesc *batch,
                                 double *recH, int *recPrev, int *recYslot,
                                 int *head, long cap, double *recY, long ycap,
                                 int every, double *recK, long *used, long *yused,
                                 gc_amd_progress_fn progress, void *user,
                                 void *stream, double *tKernel);

/* One batch of a multi-batch adjoint sweep (as gc_amd_adjoint_batch_device). */
typedef struct GcAmdAdjBatch {
  int nCell;
  double Tstart, Tend;
  const double *C;
  const double *RCONST;
  const int *ICNTRL;
  const int *Perm;
  const double *lamC;             /* (NSPEC, nCell) dL/dC at Tend */
  double *gC;                     /* (NSPEC, nCell) dL/dC at Tstart */
  double *gRCONST;                /* (NREACT, nCell) dL/dRCONST */
  int *IERR;                      /* (nCell) */
} GcAmdAdjBatch;

/* Reverse sweep over the batches of one gc_amd_record_batches_device call in
 * one launch (same pools, head and `every`, 1..64); progress (may be NULL) as
 * in gc_amd_integrate_batches_device.  Returns as
 * gc_amd_adjoint_batch_device. */
int gc_amd_adjoint_batches_device(int nBatch, const GcAmdAdjBatch *batch,
                                  const double *recH, const int *recPrev,
                                  const int *recYslot, const int *head, long cap,
                                  const double *recY, long ycap, int every,
                                  const double *recK, gc_amd_progress_fn progress,
                                  void *user, void *stream, double *tKernel);

/* ---- device-resident batch across chemistry steps ------------------------
 * The reference stages the whole batch host->device on every call and keeps
 * nothing on the device (gpulib/src/gc_gpu_batch_api.F90:199-211); its caller
 * carries each cell's step size Hnew (RSTAT(3) -> State_Chm%KPPHvalue -> next
 * step's RCNTRL(3), src/geos-chem/GeosCore/fullchem_mod.F90:1775, 2044) and
 * internal step count (ISTAT(3) -> BatchPrevNsteps, the next step's cost
 * sort key, :2522-2535 and :905-981) across steps on the host.  A resident
 * batch keeps all of it in HBM instead: the batch arrays, the controls, the
 * warm-start step sizes and the cost keys; the next step's launch order (the
 * stable descending sort by the previous step's ISTAT(3)) is computed on the
 * device, so the caller keeps its cells in its own (grid) order.  Host
 * arrays move through pinned buffers, host copies overlapping the transfers.
 * No counterpart in the reference. */
typedef struct GcAmdResident GcAmdResident;

/* A resident batch of nCell cells on the calling thread's current device,
 * with the controls ICNTRL(20,nCell), RCNTRL(20,nCell), ATOL(NVAR) and
 * RTOL(NVAR) uploaded once (layouts of gc_gpu_integrate_batch).  Returns 0,
 * -1 bad argument, -2 device memory, -3 runtime error. */
int gc_amd_resident_create(int nCell, const int *ICNTRL, const double *RCNTRL,
                           const double *ATOL, const double *RTOL,
                           GcAmdResident **out);

/* One chemistry step of the resident batch over [Tstart, Tend] with the
 * persistent kernel.  C(NSPEC,nCell) and RCONST(NREACT,nCell) are host arrays
 * uploaded for this step, or NULL to keep what the device holds (C: the
 * previous step's result; both are required on the first step).  From the
 * second step on, the launch order is the stable descending sort of the
 * cells by the previous step's ISTAT(3) (identity on the first step), and
 * with warm != 0 each cell starts from the previous step's Hnew:
 * RCNTRL(3) = RSTAT(3).  Cout, ISTAT, RSTAT, IERR (host, NULL to skip) receive
 * the results as gc_gpu_integrate_batch returns them, PermUsed (NULL to skip)
 * the 1-based launch order the step used.  tTotal / tKernel: seconds with and
 * without the host copies.  Returns 0 or the codes of
 * gc_amd_integrate_batch_device.  Blocking. */
int gc_amd_resident_step(GcAmdResident *r, double Tstart, double Tend,
                         const double *C, const double *RCONST, int warm,
                         double *Cout, int *ISTAT, double *RSTAT, int *IERR,
                         int *PermUsed, double *tTotal, double *tKernel);

/* Free a resident batch's device memory. */
void gc_amd_resident_destroy(GcAmdResident *r);

/* Launch tuning of the persistent kernel (values <= 0 keep the current one):
 *   block           work-group size (64..256, multiple of 64)
 *   blocks_per_cu   resident work-groups per compute unit (grid = CUs * this;
 *                   the adjoint sweep runs at most block/64 * this waves per CU)
 *   cells_per_wave  lanes of each 64-wide wave that take cells (16..64)
 *   grain           queue positions a wave claims per atomic (1..4096)
 *   variant         LU row-segment variant: 1 = 32-entry segments in LDS
 *                   (2 waves per SIMD), 2 = 16-entry segments (4 waves per
 *                   SIMD), 3 = 24-entry segments (3 waves per SIMD)
 * The defaults are the launch the autotuner stored for the library's
 * mechanism, compiled in; GC_AMD_TUNE (environment) overrides them at the
 * first call.  Returns 0, or -1 if a value is out of range (nothing
 * changed). */
int gc_amd_set_tuning(int block, int blocks_per_cu, int cells_per_wave,
                      int grain, int variant);
void gc_amd_get_tuning(int *block, int *blocks_per_cu, int *cells_per_wave,
                       int *grain, int *variant);

/* Overlapped upload of the host-array entry points (no counterpart in the
 * reference, whose gc_gpu_batch_api.F90:199-211 stages the whole batch before
 * the launch).  With mode 1 (the default) a gc_gpu_integrate_batch[_dev] call
 * whose launch order is the identity -- the model gathers its batch in cost
 * order and passes Perm = 1..nCell (fullchem_mod.F90:2359-2360) -- and that
 * runs the persistent kernel starts the solve first and uploads C, RCONST,
 * ICNTRL and RCNTRL in chunks of `cells` launch positions (0: what fits one
 * 32 MB staging buffer) while it runs; a wave hands out a position only once
 * its chunk has landed, waiting at most `spin` sleeps of ~3.4 us per chunk
 * (default 600,000, ~2 s; the 4x5 grid's whole upload takes ~70 ms).  A call
 * whose chunks do not land within that bound (copies that cannot run beside
 * the kernel) switches the overlap off for the rest of the process (mode 0)
 * and is solved again with the upload first.  Mode 0 uploads everything
 * first.  Mode 2 is mode 1 with the chunks' arrival never signalled, so every
 * wait runs to its bound and the call fails with IERR -98 for every cell:
 * the failure path, for tests.  Results are identical in modes 0 and 1.
 * The default is mode 0 when the process starts under rocprofv3 counter
 * collection (ROCPROF_COUNTER_COLLECTION) or a launch-serializing debug mode
 * (HIP_LAUNCH_BLOCKING, AMD_SERIALIZE_KERNEL, AMD_SERIALIZE_COPY), where the
 * copies could not land while the kernel runs.
 * Returns 0, or -1 for a value out of range (nothing changed). */
int gc_amd_set_gate(int mode, int cells, unsigned long long spin);
void gc_amd_get_gate(int *mode, int *cells, unsigned long long *spin);

/* ---- discrete adjoint (frozen step sizes) -------------------------------- */

/* Forward batch solve (persistent kernel, same arguments and results as
 * gc_amd_integrate_batch_device with kernel = persistent and chunk = 0) that
 * also records every accepted step for the adjoint.  Record s holds the
 * step's size recH[s], the previous record of the same cell recPrev[s] (-1 for
 * the first) and recYslot[s]: for every `every`-th accepted step of a cell
 * (its 1st, (every+1)-th, ...) a checkpoint slot of the state pool, whose
 * recY[slot*NVAR .. +NVAR) is the state at the start of the step, else -1.
 * head[c] (c = 0-based column of the caller's arrays) is the cell's last
 * record or -1.  recK, when not null, holds ycap x 4 x NVAR doubles: a
 * checkpoint step's stage vectors K_1..K_4 go to recK[slot*4*NVAR ..), so
 * that the sweep need not recompute them.  The pools hold `cap` records and
 * `ycap` states; *used and *yused receive what the solve needed.  Returns 0,
 * -4 if a pool was too small (the solve results are valid, the records are
 * not: call again with pools of at least *used / *yused), or the codes of
 * gc_amd_integrate_batch_device.
 * Blocking; ordered after `stream`. */
int gc_amd_record_batch_device(int nCell, double Tstart, double Tend, double *C,
                               const double *RCONST, const double *ATOL,
                               const double *RTOL, const int *ICNTRL,
                               const double *RCNTRL, const int *Perm, int *ISTAT,
                               double *RSTAT, int *IERR, double *recH, int *recPrev,
                               int *recYslot, int *head, long cap, double *recY,
                               long ycap, int every, double *recK, long *used,
                               long *yused, void *stream, double *tKernel);

/* Reverse sweep of the discrete adjoint of the recorded accepted Rodas3 steps
 * (step sizes held fixed).  Given lamC(NSPEC,nCell) = dL/dC at Tend, writes
 * gC(NSPEC,nCell) = dL/dC at Tstart (VAR and FIX) and gRCONST(NREACT,nCell) =
 * dL/dRCONST.  The record arguments are those of a successful
 * gc_amd_record_batch_device call (same pools and `every`, 1..64).  Each
 * segment of a cell's steps is replayed forward from its checkpoint (bit for
 * bit the recorded trajectory), keeping each step's stage vectors (a
 * checkpoint step's are taken from recK when it is not null: the recK of the
 * recording); then, last step first, each step's factors are recomputed and
 * the adjoint is swept through it with transposed solves on the same factors
 * (KppSolveTR order).  C supplies the fixed species;
 * ICNTRL(1) must be non-zero (autonomous, as GEOS-Chem sets it) and
 * ICNTRL(16) = 1 (clamp) is honoured.  IERR(c) = 1, -3 for a non-autonomous
 * cell (gradients 0), -8 if a recorded step's matrix is singular, -5 if the
 * cell's chain does not fit the pools.  The sweep runs as a persistent grid
 * (64 cells in Perm order at a time per wave), so its workspace is one slot
 * per resident lane whatever nCell.  Returns 0 or a negative code as above.
 * Blocking; ordered after `stream`. */
int gc_amd_adjoint_batch_device(int nCell, double Tstart, double Tend,
                                const double *C, const double *RCONST,
                                const int *ICNTRL, const int *Perm,
                                const double *recH, const int *recPrev,
                                const int *recYslot, const int *head, long cap,
                                const double *recY, long ycap, int every,
                                const double *recK, const double *lamC, double *gC,
                                double *gRCONST, int *IERR, void *stream,
                                double *tKernel);

/* ---- cost order and the hybrid CPU/GPU split ------------------------------
 * The reference runs these steps inside its caller
 * (src/geos-chem/GeosCore/fullchem_mod.F90, batch mode 3); here they are
 * library entry points with the same arithmetic.  Slots and cells are
 * 1-based, as in the caller. */

#define GC_AMD_MAX_DEV 16

/* Stable counting sort of a batch's cells by the previous chemistry step's
 * internal step count, most expensive first (the slot assignment of
 * fullchem_mod.F90:905-981).  Cell c (1-based, in the caller's stock L,J,I
 * order) gets slot slotOf[c-1]; cellAt[s-1] = c is the inverse map.  sort = 0
 * keeps the stock order (GC_BATCH_SORT=0).  Returns 0, -1 for nCell < 0, a
 * null array or a negative step count, -2 if host memory ran out. */
int gc_amd_cost_slots(int nCell, const int *prevNsteps, int sort, int *slotOf,
                      int *cellAt);

/* Work model and split of a cost-ordered batch (fullchem_mod.F90:997-1061).
 *   slotNsteps(nCell)   in: previous internal step count of the cell in each slot
 *   fixedFrac           by value: nonzero keeps *workFrac (GC_GPU_FRAC_FIXED=1)
 *   workFrac            inout: GPU share of the work (initially 0.5 or
 *                       GC_GPU_FRAC); with both rates > 0 and fixedFrac = 0 it
 *                       becomes rateGpu/(rateGpu+rateCpu) clipped to [0.01,0.99]
 *   rateGpu, rateCpu    by value: measured work units per second (<= 0: none yet)
 *   nDev                by value: devices to drive (1..GC_AMD_MAX_DEV)
 *   devRate(nDev)       in (may be NULL): per-device rates; used as slice
 *                       weights once every used device has one, else equal
 *   work(0:nCell)       out: prefix sums of max(step count, 1)
 *   nGpu                out: slots 1..nGpu go to the devices, the rest to the CPU
 *   nDevUse             out: devices used (min(nDev, max(nGpu, 1)))
 *   devLo, devHi(nDev)  out: device d solves slots devLo[d]..devHi[d]
 *                       (contiguous, work-balanced; 1 and 0 when unused)
 * Returns 0, or -1 for a bad argument. */
int gc_amd_hybrid_plan(int nCell, const int *slotNsteps, int fixedFrac,
                       double *workFrac, double rateGpu, double rateCpu, int nDev,
                       const double *devRate, long long *work, int *nGpu,
                       int *nDevUse, int *devLo, int *devHi);

/* Measured-rate update after a split solve (fullchem_mod.F90:2672-2700):
 * device d's rate from work[devHi]-work[devLo-1] over devTw[d], the GPU rate
 * from work[nGpu] over gpuTw (the slowest device's wall time), the CPU rate
 * from work[nCell]-work[nGpu] over tCpu; each is an exponential moving
 * average (the first value as is, then half old, half new).  Measurements
 * that are absent (no slots, time <= 0) leave their rate unchanged. */
void gc_amd_hybrid_update(int nCell, int nGpu, const long long *work, int nDevUse,
                          const int *devLo, const int *devHi, const double *devTw,
                          double gpuTw, double tCpu, double *rateGpu,
                          double *rateCpu, double *devRate);

/* State of the hybrid split across chemistry steps (the caller's SAVE
 * variables in the reference, fullchem_mod.F90:139-148).  Set work_frac
 * (0.5), fixed_frac (0), ndev (1..16), rate_gpu, rate_cpu and dev_rate[]
 * (-1: not measured) before the first step; the rest is written by
 * gc_amd_hybrid_step. */
typedef struct GcAmdHybrid {
  double work_frac;
  int fixed_frac;
  int ndev;
  double rate_gpu;
  double rate_cpu;
  double dev_rate[GC_AMD_MAX_DEV];
  int n_gpu;                    /* out: slots 1..n_gpu went to the devices */
  int n_dev_use;                /* out */
  int dev_lo[GC_AMD_MAX_DEV];   /* out: device slices, 1-based inclusive */
  int dev_hi[GC_AMD_MAX_DEV];
  double t_gpu;                 /* out: slowest device's wall time, s */
  double t_kernel;              /* out: kernel seconds summed over devices */
  double t_cpu;                 /* out: wall time of the CPU tail, s */
  double t_region;              /* out: wall time of the whole split solve, s */
} GcAmdHybrid;

/* The caller's CPU solve of slots first..last (1-based, inclusive). */
typedef void (*gc_amd_cpu_tail_fn)(int first, int last, void *user);

/* One hybrid chemistry step over a batch already in slot (cost) order: plan
 * the split with gc_amd_hybrid_plan, solve each device's slice with
 * gc_gpu_integrate_batch_dev from its own host thread, concurrently run
 * cpuTail(nGpu+1, nCell, user) on the calling thread (not called when every
 * slot went to the devices), then update the rates with
 * gc_amd_hybrid_update.  The batch arrays have the layouts of
 * gc_gpu_integrate_batch with nCell columns in slot order; only columns
 * 1..nGpu are read and written.  The CPU rate is measured over cpuTail's own
 * wall time (the reference, whose CPU threads also issue the launches,
 * measures from the first launch to the end of its solve loop).  Returns the
 * number of device slots with IERR < 0 (the caller retries them on the CPU,
 * as fullchem_mod.F90:2603-2650 does), or -1 for a bad argument, -2 if host
 * memory ran out. */
int gc_amd_hybrid_step(int nCell, double Tstart, double Tend, double *C,
                       const double *RCONST, const double *ATOL, const double *RTOL,
                       const int *ICNTRL, const double *RCNTRL,
                       const int *slotNsteps, int *ISTAT, double *RSTAT, int *IERR,
                       GcAmdHybrid *state, gc_amd_cpu_tail_fn cpuTail, void *user);

/* Dimensions of the mechanism compiled into this library. */
void gc_amd_mech_info(int *nvar, int *nfix, int *nreact, int *lu_nonzero);

/* Doubles of device workspace one cell slot needs. */
long gc_amd_workspace_doubles_per_slot(void);

/* Text of the last error on the calling thread ("" if none). */
const char *gc_amd_last_error(void);

/* Release the library's device buffers on the current device. */
void gc_amd_release(void);

#ifdef __cplusplus
}
#endif
#endif
