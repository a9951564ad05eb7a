// gc_rodas3.hip -- batched KPP Rosenbrock Rodas3 on gfx950, one cell per lane.
//
// Two launch forms share the generated mechanism kernels (Fun, Jac_SP + matrix
// preparation, KppDecomp, KppSolve; see codegen/gen_hip.py):
//
//  * plain      one thread per cell in Perm order, the nested time / reject /
//               singular loops of Integrate_Cell_S per thread, exactly the
//               reference's batched launch (gpulib/src/gc_gpu_batch_api.F90:
//               198-215, one `gang vector` thread per cell).
//  * persistent a fixed grid of lanes that pull cells from a device work queue
//               in Perm (cost) order.  The per-cell loops are flattened into a
//               per-lane state machine whose unit of work is one matrix
//               factorisation attempt, so lanes of a wave at different steps,
//               retries or cells still share every attempt; a lane whose cell
//               finishes refills from the queue at once.
//
// Numerics are those of Integrate_Cell_S
// (src/geos-chem/KPP/fullchem/gckpp_BatchIntegrator.F90:408-707) with the option
// merge of gpu_solve_cell (gpulib/src/gc_gpu_batch_api.F90:54-110).  Two
// bit-neutral changes: the Jacobian is re-evaluated at every factorisation
// attempt instead of being kept as Jac0 (Y is unchanged between a step's
// attempts, so the values are identical and 45 KB of state per cell is saved),
// and KppDecomp's singularity test is applied to the prepared diagonal when it
// is stored (the reference tests each row's diagonal before that row is
// eliminated, i.e. the same stored value).  Err**(1/3) uses the correctly
// rounded gc_pow_third (gc_math.h).
//
// Cell state lives in a slot-major SoA workspace: array a of slot s, element k,
// at ws[(off_a + k) * S + s].  Caller arrays keep the reference's Fortran
// layout (cell-major: C(NSPEC,nCell), RCONST(NREACT,nCell), ...).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gc_math.h"
#include "mech_gen.inc"
#include "gc_rodas3_kernels.h"

#ifndef GC_WPE
#define GC_WPE 1
#endif
#define GC_CAT2(a, b) a##b
#define GC_CAT(a, b) GC_CAT2(a, b)
#define GC_SFX(name) GC_CAT(name, GC_CAT(_w, GC_WPE))

namespace {

constexpr int NVAR = MECH_NVAR;
constexpr int NFIX = MECH_NFIX;
constexpr int NSPEC = MECH_NVAR + MECH_NFIX;
constexpr int NREACT = MECH_NREACT;
constexpr int NNZ = MECH_NNZ;

// gckpp_BatchIntegrator.F90:66-69
constexpr double ROUNDOFF = 2.220446049250313080847263336181640625e-16;
constexpr double DELTAMIN = 1.0e-5;
constexpr double TINYD = 2.2250738585072014e-308;

// Rodas3 tableau (gckpp_BatchIntegrator.F90:443-452)
__constant__ double ros_A[6] = {0.0, 2.0, 0.0, 2.0, 0.0, 1.0};
__constant__ double ros_C[6] = {4.0, 1.0, -1.0, 1.0, -1.0, -(8.0 / 3.0)};
__constant__ double ros_M[4] = {2.0, 0.0, 1.0, 1.0};
__constant__ double ros_E[4] = {0.0, 0.0, 0.0, 1.0};
__constant__ double ros_Gamma[4] = {0.5, 1.5, 0.0, 0.0};

// workspace layout (element offsets; multiply by S)
constexpr long W_Y = 0;
constexpr long W_YNEW = W_Y + NVAR;
constexpr long W_FCN0 = W_YNEW + NVAR;
constexpr long W_K = W_FCN0 + NVAR;
constexpr long W_DFDT = W_K + 4L * NVAR;
constexpr long W_RCT = W_DFDT + NVAR;
constexpr long W_FIX = W_RCT + NREACT;
constexpr long W_A = W_FIX + NFIX;
constexpr long W_LU = W_A + NREACT;
constexpr long W_TOTAL = W_LU + NNZ;
static_assert(W_TOTAL == GC_WS_DOUBLES_PER_SLOT(NVAR, NFIX, NREACT, NNZ), "layout");

__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return (a > b) ? a : b; }

struct Ws {
  double* base;
  unsigned S;
  unsigned slot;
  __device__ double* arr(long off) const { return base + off * (long)S; }
  __device__ double ld(long off, int k) const { return base[(off + k) * (long)S + slot]; }
  __device__ void st(long off, int k, double v) const { base[(off + k) * (long)S + slot] = v; }
};

// Options after the merge of gc_gpu_batch_api.F90:79-88 and the decode of
// gckpp_BatchIntegrator.F90:468-518.
struct Opts {
  int autonomous, vector_tol, max_steps, clamp;
  double Hmin, Hmax, Hstart, FacMin, FacMax, FacRej, FacSafe;
};

// Returns 0 or the early IERR (-90 auto-reduce, -2 method).
__device__ int decode_opts(const int* icu, const double* rcu, double Tstart, double Tend,
                           Opts& o) {
  int ic[20];
  double rc[20];
  for (int k = 0; k < 20; ++k) { ic[k] = 0; rc[k] = 0.0; }
  ic[14] = 5;
  for (int k = 0; k < 20; ++k) {
    if (icu[k] != 0) ic[k] = icu[k];
    if (rcu[k] > 0.0) rc[k] = rcu[k];
  }
  if (ic[11] == 1) return -90;
  o.autonomous = !(ic[0] == 0);
  o.vector_tol = (ic[1] == 0);
  if (!(ic[2] == 0 || ic[2] == 4)) return -2;
  o.max_steps = (ic[3] == 0) ? 200000 : ic[3];
  o.clamp = (ic[15] == 1);
  const double span = fabs(Tend - Tstart);
  o.Hmin = (rc[0] > 0.0) ? rc[0] : 0.0;
  o.Hmax = (rc[1] > 0.0) ? dmin(fabs(rc[1]), span) : span;
  o.Hstart = (rc[2] > 0.0) ? dmin(fabs(rc[2]), span) : dmax(o.Hmin, DELTAMIN);
  o.FacMin = (rc[3] > 0.0) ? rc[3] : 0.2;
  o.FacMax = (rc[4] > 0.0) ? rc[4] : 6.0;
  o.FacRej = (rc[5] > 0.0) ? rc[5] : 0.1;
  o.FacSafe = (rc[6] > 0.0) ? rc[6] : 0.9;
  return 0;
}

// Copy cell `cell` (0-based column of the caller arrays) into the slot.
__device__ void gather(const GcBatchArgs& b, long cell, const Ws& w) {
  const double* c = b.C + cell * NSPEC;
  const double* r = b.RCONST + cell * NREACT;
  for (int i = 0; i < NVAR; ++i) w.st(W_Y, i, c[i]);
  for (int j = 0; j < NFIX; ++j) w.st(W_FIX, j, c[NVAR + j]);
  for (int k = 0; k < NREACT; ++k) w.st(W_RCT, k, r[k]);
}

__device__ void scatter(const GcBatchArgs& b, long cell, const Ws& w, const int* ist,
                        const double* rst, int ierr) {
  double* c = b.C + cell * NSPEC;
  for (int i = 0; i < NVAR; ++i) c[i] = w.ld(W_Y, i);
  for (int k = 0; k < 8; ++k) b.ISTAT[cell * 8 + k] = ist[k];
  for (int k = 0; k < 3; ++k) b.RSTAT[cell * 3 + k] = rst[k];
  b.IERR[cell] = ierr;
}

// Fcn0 <- Fun(Y); for non-autonomous problems also the dFdT replica
// (gckpp_BatchIntegrator.F90:549-561).
__device__ void step_functions(const Ws& w, const Opts& o, double T, int* ist) {
  mech_fun(w.arr(W_Y), w.arr(W_FIX), w.arr(W_RCT), w.arr(W_A), w.arr(W_FCN0), w.S, w.slot);
  ist[0] += 1;
  if (!o.autonomous) {
    const double Delta = sqrt(ROUNDOFF) * dmax(1.0e-6, fabs(T));
    mech_fun(w.arr(W_Y), w.arr(W_FIX), w.arr(W_RCT), w.arr(W_A), w.arr(W_DFDT), w.S, w.slot);
    ist[0] += 1;
    for (int k = 0; k < NVAR; ++k)
      w.st(W_DFDT, k, (w.ld(W_DFDT, k) - w.ld(W_FCN0, k)) / Delta);
  }
}

// The four stages after a successful factorisation, then Ynew, Yerr and the
// error norm (gckpp_BatchIntegrator.F90:594-665).  Returns Err.
__device__ double stages_and_error(const Ws& w, const Opts& o, double Dir, double H,
                                   const double* atol, const double* rtol, int* ist) {
  double* const K = w.arr(W_K);
  const long KS = (long)NVAR * w.S;  // stride between stage vectors
  for (int s = 0; s < 4; ++s) {
    double* Ks = K + s * KS;
    if (s == 0 || s == 1) {
      // stage 1: Fcn = Fcn0; stage 2 has ros_NewF = .FALSE. so Fcn is still Fcn0
      for (int k = 0; k < NVAR; ++k) w.st(W_K + s * NVAR, k, w.ld(W_FCN0, k));
    } else {
      for (int k = 0; k < NVAR; ++k) {
        double y = w.ld(W_Y, k);
        for (int j = 0; j < s; ++j) {
          const double aa = ros_A[s * (s - 1) / 2 + j];
          y = y + aa * w.ld(W_K + j * NVAR, k);
        }
        w.st(W_YNEW, k, y);
      }
      mech_fun(w.arr(W_YNEW), w.arr(W_FIX), w.arr(W_RCT), w.arr(W_A), Ks, w.S, w.slot);
      ist[0] += 1;
    }
    for (int j = 0; j < s; ++j) {
      const double HC = ros_C[s * (s - 1) / 2 + j] / (Dir * H);
      for (int k = 0; k < NVAR; ++k)
        w.st(W_K + s * NVAR, k, w.ld(W_K + s * NVAR, k) + HC * w.ld(W_K + j * NVAR, k));
    }
    if (!o.autonomous && ros_Gamma[s] != 0.0) {
      const double HG = (Dir * H) * ros_Gamma[s];
      for (int k = 0; k < NVAR; ++k)
        w.st(W_K + s * NVAR, k, w.ld(W_K + s * NVAR, k) + HG * w.ld(W_DFDT, k));
    }
    mech_solve(w.arr(W_LU), Ks, w.S, w.slot);
    ist[6] += 1;
  }
  double Err = 0.0;
  for (int k = 0; k < NVAR; ++k) {
    const double y = w.ld(W_Y, k);
    double yn = y, ye = 0.0;
    for (int j = 0; j < 4; ++j) yn = yn + ros_M[j] * w.ld(W_K + j * NVAR, k);
    for (int j = 0; j < 4; ++j) ye = ye + ros_E[j] * w.ld(W_K + j * NVAR, k);
    w.st(W_YNEW, k, yn);
    const double Ymax = dmax(fabs(y), fabs(yn));
    const double Scal = o.vector_tol ? atol[k] + rtol[k] * Ymax : atol[0] + rtol[0] * Ymax;
    const double q = ye / Scal;
    Err = Err + q * q;
  }
  Err = sqrt(Err / (double)NVAR);
  return dmax(Err, 1.0e-10);
}

__device__ void accept_y(const Ws& w, const Opts& o) {
  if (o.clamp) {
    for (int k = 0; k < NVAR; ++k) w.st(W_Y, k, dmax(w.ld(W_YNEW, k), 0.0));
  } else {
    for (int k = 0; k < NVAR; ++k) w.st(W_Y, k, w.ld(W_YNEW, k));
  }
}

__device__ __forceinline__ bool time_left(int dir, double T, double Tend) {
  return (dir > 0 && ((T - Tend) + ROUNDOFF <= 0.0)) || (dir < 0 && ((Tend - T) + ROUNDOFF <= 0.0));
}

// ---------------------------------------------------------------------------
// plain: Integrate_Cell_S per thread, with its own nested loops.
// ---------------------------------------------------------------------------
__device__ int integrate_cell(const Ws& w, const Opts& o, double Tstart, double Tend,
                              const double* atol, const double* rtol, int* ist, double* rst) {
  int Nsteps = 0;
  double T = Tstart;
  double H = dmin(dmax(fabs(o.Hmin), fabs(o.Hstart)), fabs(o.Hmax));
  if (fabs(H) <= 10.0 * ROUNDOFF) H = DELTAMIN;
  const int dir = (Tend >= Tstart) ? +1 : -1;
  const double Dir = (double)dir;
  H = Dir * H;
  bool rejLast = false, rejMore = false;
  while (time_left(dir, T, Tend)) {
    if (Nsteps > o.max_steps) return -6;
    if (((T + 0.1 * H) == T) || (H <= ROUNDOFF)) return -7;
    H = dmin(H, fabs(Tend - T));
    step_functions(w, o, T, ist);
    ist[1] += 1;  // Jac_SP of this step (evaluated in each attempt below)
    for (;;) {  // UntilAccepted
      int ncons = 0;
      for (;;) {  // singular retries
        const double ghinv = 1.0 / ((Dir * H) * ros_Gamma[0]);
        const int sing = mech_jac_prep(w.arr(W_Y), w.arr(W_FIX), w.arr(W_RCT), w.arr(W_LU),
                                       ghinv, w.S, w.slot);
        ist[5] += 1;
        if (!sing) break;
        ist[7] += 1;
        ncons += 1;
        if (ncons <= 5) H = H * 0.5;
        else return -8;
      }
      mech_decomp(w.arr(W_LU), w.S, w.slot);
      const double Err = stages_and_error(w, o, Dir, H, atol, rtol, ist);
      const double Fac = dmin(o.FacMax, dmax(o.FacMin, o.FacSafe / gc_pow_third(Err)));
      double Hnew = H * Fac;
      Nsteps += 1;
      ist[2] += 1;
      if ((Err <= 1.0) || (H <= o.Hmin)) {
        ist[3] += 1;
        accept_y(w, o);
        T = T + Dir * H;
        Hnew = dmax(o.Hmin, dmin(Hnew, o.Hmax));
        if (rejLast) Hnew = dmin(Hnew, H);
        rst[1] = H;
        rst[2] = Hnew;
        rst[0] = T;
        rejLast = false;
        rejMore = false;
        H = Hnew;
        break;
      } else {
        if (rejMore) Hnew = H * o.FacRej;
        rejMore = rejLast;
        rejLast = true;
        H = Hnew;
        if (ist[3] >= 1) ist[4] += 1;
      }
    }
  }
  return 1;
}

}  // namespace

extern "C" __global__ void
__attribute__((amdgpu_flat_work_group_size(1, 256), amdgpu_waves_per_eu(GC_WPE, GC_WPE)))
GC_SFX(gc_rodas3_plain)(GcBatchArgs b, double* ws, unsigned S) {
  const unsigned ic = blockIdx.x * blockDim.x + threadIdx.x;
  if (ic >= (unsigned)b.nCell) return;
  const Ws w{ws, S, ic};
  const long cell = (long)b.Perm[ic] - 1;
  int ist[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double rst[3] = {0.0, 0.0, 0.0};
  gather(b, cell, w);
  Opts o;
  int ierr = decode_opts(b.ICNTRL + cell * 20, b.RCNTRL + cell * 20, b.Tstart, b.Tend, o);
  if (ierr == 0) ierr = integrate_cell(w, o, b.Tstart, b.Tend, b.ATOL, b.RTOL, ist, rst);
  scatter(b, cell, w, ist, rst, ierr);
}

// ---------------------------------------------------------------------------
// persistent: work queue + flattened per-lane state machine.
// ---------------------------------------------------------------------------
namespace {
enum : int { PH_FREE = 0, PH_STEP = 1, PH_RETRY = 2, PH_SING = 3 };
}

extern "C" __global__ void
__attribute__((amdgpu_flat_work_group_size(1, 256), amdgpu_waves_per_eu(GC_WPE, GC_WPE)))
GC_SFX(gc_rodas3_persistent)(GcBatchArgs b, double* ws, unsigned S, unsigned* queue,
                             int cells_per_wave, int grain) {
  const unsigned slot = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const Ws w{ws, S, slot};
  const bool lane_on = (lane < cells_per_wave) && (slot < S);

  // per-lane state
  int phase = PH_FREE;
  long cell = -1;
  Opts o;
  int ist[8];
  double rst[3];
  double T = 0.0, H = 0.0, Dir = 1.0;
  int dir = 1, Nsteps = 0, ncons = 0;
  bool rejLast = false, rejMore = false;
  bool exhausted = !lane_on;
  // wave-uniform reservation of queue positions [rsv, rsv_end)
  unsigned rsv = 0, rsv_end = 0;

  for (;;) {
    // ---- refill free lanes from the queue (cost order) -------------------
    for (;;) {
      const bool need = (phase == PH_FREE) && !exhausted;
      const unsigned long long mask = __ballot(need);
      if (mask == 0ull) break;
      const unsigned cnt = (unsigned)__popcll(mask);
      if (rsv >= rsv_end) {
        unsigned base = 0;
        const unsigned g = (unsigned)(grain > 0 ? grain : 1);
        const unsigned take = (g > cnt) ? g : cnt;
        if (lane == __ffsll((long long)mask) - 1) base = atomicAdd(queue, take);
        base = __shfl(base, __ffsll((long long)mask) - 1);
        rsv = base;
        rsv_end = base + take;
      }
      const unsigned avail = rsv_end - rsv;
      const unsigned rank = (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
      if (need && rank < avail) {
        const unsigned q = rsv + rank;
        if (q < (unsigned)b.nCell) {
          cell = (long)b.Perm[q] - 1;
          for (int k = 0; k < 8; ++k) ist[k] = 0;
          for (int k = 0; k < 3; ++k) rst[k] = 0.0;
          gather(b, cell, w);
          const int e = decode_opts(b.ICNTRL + cell * 20, b.RCNTRL + cell * 20, b.Tstart, b.Tend, o);
          if (e != 0) {
            scatter(b, cell, w, ist, rst, e);
            cell = -1;  // stays free: take another
          } else {
            Nsteps = 0;
            T = b.Tstart;
            H = dmin(dmax(fabs(o.Hmin), fabs(o.Hstart)), fabs(o.Hmax));
            if (fabs(H) <= 10.0 * ROUNDOFF) H = DELTAMIN;
            dir = (b.Tend >= b.Tstart) ? +1 : -1;
            Dir = (double)dir;
            H = Dir * H;
            rejLast = rejMore = false;
            phase = PH_STEP;
          }
        } else {
          exhausted = true;
        }
      }
      rsv += (cnt < avail) ? cnt : avail;
      if (rsv >= (unsigned)b.nCell && rsv >= rsv_end) {
        // queue drained for this wave: lanes still free are done
        if (phase == PH_FREE) exhausted = true;
      }
    }
    if (!__any(phase != PH_FREE)) break;

    // ---- top of TimeLoop for lanes starting a step -----------------------
    int ierr_now = 0;
    if (phase == PH_STEP) {
      if (!time_left(dir, T, b.Tend)) ierr_now = 1;
      else if (Nsteps > o.max_steps) ierr_now = -6;
      else if (((T + 0.1 * H) == T) || (H <= ROUNDOFF)) ierr_now = -7;
      if (ierr_now != 0) {
        scatter(b, cell, w, ist, rst, ierr_now);
        phase = PH_FREE;
        cell = -1;
      } else {
        H = dmin(H, fabs(b.Tend - T));
      }
    }
    if (phase == PH_STEP) {
      step_functions(w, o, T, ist);
      ist[1] += 1;
    }
    if (phase == PH_STEP || phase == PH_RETRY) ncons = 0;

    // ---- one factorisation attempt ---------------------------------------
    // PH_STEP: first attempt of a step; PH_RETRY: attempt after a rejection;
    // PH_SING: the previous factorisation of this attempt was singular and H
    // was halved (ros_PrepareMatrix loop, gckpp_BatchIntegrator.F90:572-592).
    bool factored = false;
    if (phase != PH_FREE) {
      const double ghinv = 1.0 / ((Dir * H) * ros_Gamma[0]);
      const int sing = mech_jac_prep(w.arr(W_Y), w.arr(W_FIX), w.arr(W_RCT), w.arr(W_LU),
                                     ghinv, w.S, w.slot);
      ist[5] += 1;
      if (sing) {
        ist[7] += 1;
        ncons += 1;
        if (ncons <= 5) {
          H = H * 0.5;
          phase = PH_SING;
        } else {
          scatter(b, cell, w, ist, rst, -8);
          phase = PH_FREE;
          cell = -1;
        }
      } else {
        factored = true;
      }
    }
    if (factored) {
      mech_decomp(w.arr(W_LU), w.S, w.slot);
      const double Err = stages_and_error(w, o, Dir, H, b.ATOL, b.RTOL, ist);
      const double Fac = dmin(o.FacMax, dmax(o.FacMin, o.FacSafe / gc_pow_third(Err)));
      double Hnew = H * Fac;
      Nsteps += 1;
      ist[2] += 1;
      if ((Err <= 1.0) || (H <= o.Hmin)) {
        ist[3] += 1;
        accept_y(w, o);
        T = T + Dir * H;
        Hnew = dmax(o.Hmin, dmin(Hnew, o.Hmax));
        if (rejLast) Hnew = dmin(Hnew, H);
        rst[1] = H;
        rst[2] = Hnew;
        rst[0] = T;
        rejLast = false;
        rejMore = false;
        H = Hnew;
        phase = PH_STEP;
      } else {
        if (rejMore) Hnew = H * o.FacRej;
        rejMore = rejLast;
        rejLast = true;
        H = Hnew;
        if (ist[3] >= 1) ist[4] += 1;
        phase = PH_RETRY;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// host launch wrappers of this register-budget variant
// ---------------------------------------------------------------------------
extern "C" hipError_t GC_SFX(gc_launch_plain)(const GcBatchArgs* b, double* ws, unsigned S,
                                               int block, hipStream_t st) {
  if (b->nCell <= 0) return hipSuccess;
  const unsigned grid = ((unsigned)b->nCell + (unsigned)block - 1u) / (unsigned)block;
  hipLaunchKernelGGL(GC_SFX(gc_rodas3_plain), dim3(grid), dim3(block), 0, st, *b, ws, S);
  return hipGetLastError();
}

extern "C" hipError_t GC_SFX(gc_launch_persistent)(const GcBatchArgs* b, double* ws, unsigned S,
                                                    unsigned* queue, int block,
                                                    int cells_per_wave, int grain,
                                                    hipStream_t st) {
  if (b->nCell <= 0) return hipSuccess;
  const unsigned grid = S / (unsigned)block;
  hipLaunchKernelGGL(GC_SFX(gc_rodas3_persistent), dim3(grid), dim3(block), 0, st, *b, ws, S,
                     queue, cells_per_wave, grain);
  return hipGetLastError();
}

extern "C" void GC_SFX(gc_mech_dims)(GcMechDims* d) {
  d->nvar = NVAR;
  d->nfix = NFIX;
  d->nreact = NREACT;
  d->nnz = NNZ;
}
