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
 * the batch in HBM (the PyTorch binding, benchmarks), launch tuning, and a
 * query of the compiled mechanism's dimensions.
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

/* Launch tuning of the persistent kernel (values <= 0 keep the current one):
 *   block           work-group size (64..256, multiple of 64)
 *   blocks_per_cu   resident work-groups per compute unit (grid = CUs * this)
 *   cells_per_wave  lanes of each 64-wide wave that take cells (16..64)
 *   grain           queue positions a wave claims per atomic (1..4096)
 *   variant         LU row-segment variant: 1 = 32-entry segments in LDS
 *                   (2 waves per SIMD), 2 = 16-entry segments (4 waves per
 *                   SIMD), 3 = 24-entry segments (3 waves per SIMD)
 * Returns 0, or -1 if a value is out of range (nothing changed). */
int gc_amd_set_tuning(int block, int blocks_per_cu, int cells_per_wave,
                      int grain, int variant);
void gc_amd_get_tuning(int *block, int *blocks_per_cu, int *cells_per_wave,
                       int *grain, int *variant);

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
 * bit the recorded trajectory), then each step's factors and stages are
 * recomputed (a checkpoint step's stages are taken from recK when it is not
 * null: the recK of the recording) and the adjoint is swept through it with
 * transposed solves on the same factors (KppSolveTR order).  C supplies the fixed species;
 * ICNTRL(1) must be non-zero (autonomous, as GEOS-Chem sets it) and
 * ICNTRL(16) = 1 (clamp) is honoured.  IERR(c) = 1, -3 for a non-autonomous
 * cell (gradients 0), -8 if a recorded step's matrix is singular, -5 if the
 * cell's chain does not fit the pools.  Returns 0 or a negative code as
 * above.  Blocking; ordered after `stream`. */
int gc_amd_adjoint_batch_device(int nCell, double Tstart, double Tend,
                                const double *C, const double *RCONST,
                                const int *ICNTRL, const int *Perm,
                                const double *recH, const int *recPrev,
                                const int *recYslot, const int *head, long cap,
                                const double *recY, long ycap, int every,
                                const double *recK, const double *lamC, double *gC,
                                double *gRCONST, int *IERR, void *stream,
                                double *tKernel);

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
