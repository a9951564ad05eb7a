// gc_api.hip -- C ABI of libgcchemgpu.so (declarations: include/gcchemgpu.h).
//
// Host side of the batched Rodas3 solve: device selection and query, per-device
// buffers and stream, host<->device staging for the reference entry points, the
// launch-order validation, the plain / persistent / chunked launch forms, and
// launch tuning.  The library does no I/O (as gpulib/src/gc_gpu_batch_api.F90:
// 11-15 requires of the reference): errors are reported through IERR, return
// codes and gc_amd_last_error().
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gcchemgpu.h"
#include "gc_rodas3_kernels.h"

namespace {

thread_local std::string g_last_error;

void set_error(const char* where, hipError_t e) {
  g_last_error = std::string(where) + ": " + hipGetErrorString(e);
}
void set_error(const char* msg) { g_last_error = msg; }

GcVariantInfo variant_info(int v) {
  GcVariantInfo x;
  if (v == 2) gc_variant_info_v2(&x);
  else if (v == 3) gc_variant_info_v3(&x);
  else gc_variant_info_v1(&x);
  return x;
}

struct MechDims {
  int nvar, nfix, nreact, nnz;
};

MechDims mech_dims() {
  const GcVariantInfo v = variant_info(1);
  return MechDims{v.nvar, v.nfix, v.nreact, v.nnz};
}

long ws_per_slot() { return variant_info(1).ws_doubles_per_slot; }

// ---- launch tuning (process-wide) ------------------------------------------
// blocks_per_cu = 0 means "fill the CU": as many resident work-groups as the
// variant's occupancy target allows (waves_per_eu * 4 SIMDs per CU).
struct Tuning {
  int block = 256;
  int blocks_per_cu = 0;
  int cells_per_wave = 64;
  int grain = 64;
  int variant = 1;
};
std::mutex g_tune_mu;
Tuning g_tune;
bool g_tune_env_read = false;

Tuning current_tuning() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (!g_tune_env_read) {
    g_tune_env_read = true;
    // GC_AMD_TUNE="block,blocks_per_cu,cells_per_wave,grain,variant"
    if (const char* s = getenv("GC_AMD_TUNE")) {
      int v[5] = {0, 0, 0, 0, 0};
      if (sscanf(s, "%d,%d,%d,%d,%d", &v[0], &v[1], &v[2], &v[3], &v[4]) == 5) {
        if (v[0] >= 64 && v[0] <= 256 && v[0] % 64 == 0) g_tune.block = v[0];
        if (v[1] >= 0 && v[1] <= 32) g_tune.blocks_per_cu = v[1];
        if (v[2] >= 1 && v[2] <= 64) g_tune.cells_per_wave = v[2];
        if (v[3] >= 1 && v[3] <= 4096) g_tune.grain = v[3];
        if (v[4] >= 1 && v[4] <= 3) g_tune.variant = v[4];
      }
    }
  }
  return g_tune;
}

// ---- per-device context -----------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      bytes = 0;
      if (e != hipSuccess) return e;
    }
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    else p = nullptr;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// The chunk loop of one launch shape, captured as a graph and instantiated
// once; later calls of the same shape replay it (updated in place when only
// the arguments differ).
struct ChunkGraph {
  hipGraphExec_t exec = nullptr;
  int kernel = -1, span = 0, nCell = 0;
  unsigned S = 0;
  Tuning t;
  GcBatchArgs args{};
  void* ws = nullptr;
  void* queue = nullptr;
  bool same_shape(int k, int sp, int n, unsigned s, const Tuning& u) const {
    return exec && kernel == k && span == sp && nCell == n && S == s && t.block == u.block &&
           t.blocks_per_cu == u.blocks_per_cu && t.cells_per_wave == u.cells_per_wave &&
           t.grain == u.grain && t.variant == u.variant;
  }
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
};

bool same_args(const GcBatchArgs& a, const GcBatchArgs& b) {
  return a.nCell == b.nCell && a.Tstart == b.Tstart && a.Tend == b.Tend && a.C == b.C &&
         a.RCONST == b.RCONST && a.ATOL == b.ATOL && a.RTOL == b.RTOL && a.ICNTRL == b.ICNTRL &&
         a.RCNTRL == b.RCNTRL && a.Perm == b.Perm && a.ISTAT == b.ISTAT && a.RSTAT == b.RSTAT &&
         a.IERR == b.IERR;
}

// Device scratch of gc_amd_integrate_batch_noalloc: the queue counter and
// the launch-order flag in its first 4 KiB, then the persistent kernel's
// workspace slots.  Every device holds its own copy until the process ends;
// host code uses its per-device address (DevCtx::scratch), never this name.
constexpr size_t kNoAllocBytes = size_t(256) << 20;
constexpr size_t kNoAllocHead = 4096;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

struct DevCtx {
  std::mutex mu;
  bool init = false;
  hipStream_t stream = nullptr;
  int n_cu = 0;
  DevBuf ws, queue, flag;
  ChunkGraph graph;
  unsigned char* scratch = nullptr;  // this device's copy of `scratch`
  hipError_t scratch_err = hipSuccess;
  // staging for the host-array entry points
  DevBuf C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL, Perm, ISTAT, RSTAT, IERR;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_in = nullptr;  // orders the library stream after the caller's
};

constexpr int kMaxDev = 64;
DevCtx g_ctx[kMaxDev];

hipError_t ctx_init(DevCtx& c, int dev) {
  if (c.init) return hipSuccess;
  hipError_t e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
  if (e != hipSuccess) return e;
  if ((e = hipEventCreate(&c.ev0)) != hipSuccess) return e;
  if ((e = hipEventCreate(&c.ev1)) != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&c.ev_in, hipEventDisableTiming)) != hipSuccess) return e;
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, dev)) != hipSuccess) return e;
  c.n_cu = prop.multiProcessorCount;
  // the address of this device's scratch array (the allocation-free entry
  // point's workspace); the other entry points do not need it
  void* sp = nullptr;
  c.scratch_err = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
  c.scratch = c.scratch_err == hipSuccess ? (unsigned char*)sp : nullptr;
  (void)hipGetLastError();
  c.init = true;
  return hipSuccess;
}

DevCtx* current_ctx(int* dev_out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return nullptr;
  if (dev_out) *dev_out = dev;
  return &g_ctx[dev];
}

// ---- small utility kernels ---------------------------------------------------
__global__ void k_check_perm(const int* perm, int n, int* bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int p = perm[i];
    if (p < 1 || p > n) atomicOr(bad, 1);
  }
}

__global__ void k_fill_i32(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_zero_u32(unsigned* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---- the device solve ----------------------------------------------------------
// Launches one chunk (a contiguous range of the launch order) and returns the
// runtime status of the launch.
hipError_t launch_chunk(int kernel, const Tuning& t, const GcBatchArgs& b, unsigned S,
                        double* ws, unsigned* q, hipStream_t st) {
  if (kernel == GC_AMD_KERNEL_PLAIN) {
    // the reference's launch: one thread per cell, 128-thread work-groups
    return gc_launch_plain_v1(&b, ws, 128, st);
  }
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, q, 1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // A wave claims `grain` queue positions per atomic.  Never let one claim
  // exceed the wave's fair share of the batch, or a few waves would take every
  // cell while the rest of the grid idles.
  const unsigned waves = S / 64u;
  unsigned fair = ((unsigned)b.nCell + waves - 1u) / waves;
  if (fair < 1u) fair = 1u;
  const int grain = (unsigned)t.grain > fair ? (int)fair : t.grain;
  if (t.variant == 2)
    return gc_launch_persistent_v2(&b, ws, S, q, t.block, t.cells_per_wave, grain, st);
  if (t.variant == 3)
    return gc_launch_persistent_v3(&b, ws, S, q, t.block, t.cells_per_wave, grain, st);
  return gc_launch_persistent_v1(&b, ws, S, q, t.block, t.cells_per_wave, grain, st);
}

// Number of slots (lanes with workspace) for a launch of n cells; a
// persistent grid has at most max_blocks work-groups when that is > 0.
unsigned slots_for(DevCtx& c, int kernel, const Tuning& t, int n, unsigned max_blocks = 0) {
  if (kernel == GC_AMD_KERNEL_PLAIN) return ((unsigned)n + 63u) & ~63u;
  // persistent: one resident grid; no more waves than there are cells to hand out
  int bpc = t.blocks_per_cu;
  if (bpc <= 0) {
    const int waves_per_cu = 4 * variant_info(t.variant).waves_per_eu;
    bpc = waves_per_cu / (t.block / 64);
    if (bpc < 1) bpc = 1;
  }
  unsigned grid_cap = (unsigned)(c.n_cu * bpc);
  if (max_blocks > 0 && max_blocks < grid_cap) grid_cap = max_blocks;
  const unsigned cells_per_block = (unsigned)(t.block / 64 * t.cells_per_wave);
  unsigned need = ((unsigned)n + cells_per_block - 1u) / cells_per_block;
  if (need < 1u) need = 1u;
  const unsigned grid = need < grid_cap ? need : grid_cap;
  return grid * (unsigned)t.block;
}

// Launch order must be a map into 1..nCell (a bad entry would address outside
// the caller's arrays).  Returns 0, -1 (bad entry), -2 or -3.
// `flag` is one int of device memory (the library's own when null).
int check_perm(DevCtx& c, const int* Perm, int nCell, hipStream_t st, int* flag = nullptr) {
  hipError_t e;
  if (!flag) {
    if ((e = c.flag.ensure(sizeof(int))) != hipSuccess) { set_error("hipMalloc(flag)", e); return -2; }
    flag = (int*)c.flag.p;
  }
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, (unsigned*)flag, 1);
  hipLaunchKernelGGL(k_check_perm, dim3((nCell + 255) / 256), dim3(256), 0, st, Perm, nCell,
                     flag);
  int bad = 0;
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess) {
    set_error("perm check", e);
    return -3;
  }
  if (bad) {
    set_error("Perm entry outside 1..nCell");
    return -1;
  }
  return 0;
}

int device_solve(DevCtx& c, int kernel, int chunk, const GcBatchArgs& b, hipStream_t st,
                 double* tKernel) {
  if (tKernel) *tKernel = 0.0;
  if (b.nCell < 0 || (kernel != GC_AMD_KERNEL_PLAIN && kernel != GC_AMD_KERNEL_PERSISTENT) ||
      chunk < 0) {
    set_error("gc_amd_integrate_batch_device: bad argument");
    return -1;
  }
  if (b.nCell == 0) return 0;
  const Tuning t = current_tuning();
  hipError_t e;

  {
    const int rc = check_perm(c, b.Perm, b.nCell, st);
    if (rc) return rc;
  }

  const int span = (chunk > 0 && chunk < b.nCell) ? chunk : b.nCell;
  const unsigned S = slots_for(c, kernel, t, span);
  if ((e = c.ws.ensure((size_t)S * (size_t)ws_per_slot() * sizeof(double))) != hipSuccess) {
    set_error("hipMalloc(workspace)", e);
    return -2;
  }
  if ((e = c.queue.ensure(256)) != hipSuccess) { set_error("hipMalloc(queue)", e); return -2; }

  if (span == b.nCell) {
    if ((e = hipEventRecord(c.ev0, st)) != hipSuccess) { set_error("hipEventRecord", e); return -3; }
    e = launch_chunk(kernel, t, b, S, (double*)c.ws.p, (unsigned*)c.queue.p, st);
    if (e != hipSuccess) { set_error("launch", e); return -3; }
  } else {
    // The chunk loop over the launch order as one graph: captured and
    // instantiated for a new launch shape, updated in place for new
    // arguments, replayed as it is when nothing changed.
    ChunkGraph& cg = c.graph;
    const bool shape = cg.same_shape(kernel, span, b.nCell, S, t);
    if (!(shape && same_args(cg.args, b) && cg.ws == c.ws.p && cg.queue == c.queue.p)) {
      hipGraph_t g = nullptr;
      if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
        set_error("hipStreamBeginCapture", e);
        return -3;
      }
      hipError_t le = hipSuccess;
      for (int c0 = 0; c0 < b.nCell && le == hipSuccess; c0 += span) {
        GcBatchArgs sub = b;
        sub.Perm = b.Perm + c0;
        sub.nCell = (b.nCell - c0 < span) ? (b.nCell - c0) : span;
        le = launch_chunk(kernel, t, sub, S, (double*)c.ws.p, (unsigned*)c.queue.p, st);
      }
      e = hipStreamEndCapture(st, &g);
      if (le != hipSuccess || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        set_error("chunk capture", le != hipSuccess ? le : e);
        return -3;
      }
      bool updated = false;
      if (shape) {
        hipGraphNode_t bad_node = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        updated = hipGraphExecUpdate(cg.exec, g, &bad_node, &res) == hipSuccess &&
                  res == hipGraphExecUpdateSuccess;
        (void)hipGetLastError();  // a refused update leaves no error behind
      }
      if (!updated) {
        cg.release();
        e = hipGraphInstantiate(&cg.exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
          cg.exec = nullptr;
          (void)hipGraphDestroy(g);
          set_error("chunk graph instantiate", e);
          return -3;
        }
      }
      (void)hipGraphDestroy(g);
      cg.kernel = kernel;
      cg.span = span;
      cg.nCell = b.nCell;
      cg.S = S;
      cg.t = t;
      cg.args = b;
      cg.ws = c.ws.p;
      cg.queue = c.queue.p;
    }
    if ((e = hipEventRecord(c.ev0, st)) != hipSuccess) { set_error("hipEventRecord", e); return -3; }
    if ((e = hipGraphLaunch(cg.exec, st)) != hipSuccess) { set_error("chunk graph", e); return -3; }
  }
  if ((e = hipEventRecord(c.ev1, st)) != hipSuccess) { set_error("hipEventRecord", e); return -3; }
  if ((e = hipStreamSynchronize(st)) != hipSuccess) { set_error("solve", e); return -3; }
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess && tKernel) *tKernel = ms * 1e-3;
  return 0;
}

// env helpers mirroring gpulib/src/gc_gpu_batch_api.F90:231-256
int env_kernel() {
  const char* s = getenv("GC_GPU_KERNEL");
  if (s && (strncmp(s, "plain", 5) == 0 || strncmp(s, "PLAIN", 5) == 0)) return GC_AMD_KERNEL_PLAIN;
  return GC_AMD_KERNEL_PERSISTENT;
}
int env_chunk(int nCell) {
  const char* s = getenv("GC_GPU_WORKSPACE");
  if (!s || !(strncmp(s, "global", 6) == 0 || strncmp(s, "GLOBAL", 6) == 0)) return 0;
  int v = 32768;
  if (const char* cs = getenv("GC_GPU_CHUNK")) {
    int val = 0;
    bool ok = *cs != 0;
    for (const char* p = cs; *p; ++p) {
      if (*p < '0' || *p > '9') { ok = false; break; }
      val = val * 10 + (*p - '0');
      if (val > 1 << 30) val = 1 << 30;  // any larger chunk means one launch anyway
    }
    if (ok && val > 0) v = val;
  }
  if (v > nCell) v = nCell;
  if (v < 1024) v = 1024;
  return v;
}

void fill_ierr(int* IERR, int n, int code) {
  for (int i = 0; i < n; ++i) IERR[i] = code;
}

// Recording forward solve (persistent kernel, one launch, no chunking).
int record_solve(DevCtx& c, const GcBatchArgs& b, const GcRecord& rec_in, long* used,
                 long* yused, hipStream_t st, double* tKernel) {
  if (tKernel) *tKernel = 0.0;
  if (used) *used = 0;
  if (yused) *yused = 0;
  if (b.nCell == 0) return 0;
  const Tuning t = current_tuning();
  int rc = check_perm(c, b.Perm, b.nCell, st);
  if (rc) return rc;
  hipError_t e;
  const unsigned S = slots_for(c, GC_AMD_KERNEL_PERSISTENT, t, b.nCell);
  if ((e = c.ws.ensure((size_t)S * (size_t)ws_per_slot() * sizeof(double))) != hipSuccess) {
    set_error("hipMalloc(workspace)", e);
    return -2;
  }
  if ((e = c.queue.ensure(256)) != hipSuccess) { set_error("hipMalloc(queue)", e); return -2; }
  unsigned* q = (unsigned*)c.queue.p;
  GcRecord rec = rec_in;
  rec.count = q + 1;   // the second queue word counts records,
  rec.ycount = q + 2;  // the third one checkpoint states
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, q, 3);
  hipLaunchKernelGGL(k_fill_i32, dim3((b.nCell + 255) / 256), dim3(256), 0, st, rec.head,
                     b.nCell, -1);
  if ((e = hipEventRecord(c.ev0, st)) != hipSuccess) { set_error("hipEventRecord", e); return -3; }
  const unsigned waves = S / 64u;
  unsigned fair = ((unsigned)b.nCell + waves - 1u) / waves;
  if (fair < 1u) fair = 1u;
  const int grain = (unsigned)t.grain > fair ? (int)fair : t.grain;
  e = (t.variant == 2)   ? gc_launch_record_v2(&b, (double*)c.ws.p, S, q, t.block, t.cells_per_wave, grain, &rec, st)
      : (t.variant == 3) ? gc_launch_record_v3(&b, (double*)c.ws.p, S, q, t.block, t.cells_per_wave, grain, &rec, st)
                         : gc_launch_record_v1(&b, (double*)c.ws.p, S, q, t.block, t.cells_per_wave, grain, &rec, st);
  if (e != hipSuccess) { set_error("launch", e); return -3; }
  if ((e = hipEventRecord(c.ev1, st)) != hipSuccess) { set_error("hipEventRecord", e); return -3; }
  unsigned counts[2] = {0, 0};
  if ((e = hipMemcpyAsync(counts, rec.count, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess) {
    set_error("record solve", e);
    return -3;
  }
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess && tKernel) *tKernel = ms * 1e-3;
  if (used) *used = (long)counts[0];
  if (yused) *yused = (long)counts[1];
  if (counts[0] > rec.cap || counts[1] > rec.ycap) {
    set_error("record pools too small: retry with cap >= *used and ycap >= *yused");
    return -4;
  }
  return 0;
}

int adjoint_solve(DevCtx& c, const GcAdjArgs& a, hipStream_t st, double* tKernel) {
  if (tKernel) *tKernel = 0.0;
  if (a.nCell == 0) return 0;
  const Tuning t = current_tuning();
  int rc = check_perm(c, a.Perm, a.nCell, st);
  if (rc) return rc;
  hipError_t e;
  const size_t slots = ((size_t)a.nCell + 63u) & ~(size_t)63u;
  const GcVariantInfo vi = variant_info(t.variant);
  GcAdjArgs aa = a;
  // adjoint block + the segment buffers (every start states, step sizes)
  aa.slot_doubles = vi.adj_doubles_per_slot + (long)a.every * (vi.nvar + 1);
  if ((e = c.ws.ensure(slots * (size_t)aa.slot_doubles * sizeof(double))) != hipSuccess) {
    set_error("hipMalloc(workspace)", e);
    return -2;
  }
  if ((e = hipEventRecord(c.ev0, st)) != hipSuccess) { set_error("hipEventRecord", e); return -3; }
  e = (t.variant == 2)   ? gc_launch_adjoint_v2(&aa, (double*)c.ws.p, 128, st)
      : (t.variant == 3) ? gc_launch_adjoint_v3(&aa, (double*)c.ws.p, 128, st)
                         : gc_launch_adjoint_v1(&aa, (double*)c.ws.p, 128, st);
  if (e != hipSuccess) { set_error("launch", e); return -3; }
  if ((e = hipEventRecord(c.ev1, st)) != hipSuccess) { set_error("hipEventRecord", e); return -3; }
  if ((e = hipStreamSynchronize(st)) != hipSuccess) { set_error("adjoint", e); return -3; }
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess && tKernel) *tKernel = ms * 1e-3;
  return 0;
}

// Device context of the calling thread, initialised, locked by the caller;
// orders the library stream after the caller's stream.
DevCtx* enter(void* stream, int* rc) {
  int d = -1;
  DevCtx* cp = current_ctx(&d);
  if (!cp) {
    set_error("no current device");
    *rc = -3;
    return nullptr;
  }
  cp->mu.lock();
  hipError_t e = ctx_init(*cp, d);
  if (e == hipSuccess) e = hipEventRecord(cp->ev_in, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(cp->stream, cp->ev_in, 0);
  if (e != hipSuccess) {
    set_error("device init / stream ordering", e);
    cp->mu.unlock();
    *rc = -3;
    return nullptr;
  }
  *rc = 0;
  return cp;
}

}  // namespace

extern "C" {

int gc_amd_record_batch_device(int nCell, double Tstart, double Tend, double* C,
                               const double* RCONST, const double* ATOL, const double* RTOL,
                               const int* ICNTRL, const double* RCNTRL, const int* Perm,
                               int* ISTAT, double* RSTAT, int* IERR, double* recH,
                               int* recPrev, int* recYslot, int* head, long cap, double* recY,
                               long ycap, int every, double* recK, long* used, long* yused,
                               void* stream, double* tKernel) {
  if (nCell < 0 || cap < 0 || cap > 0x7fffffffL || ycap < 0 || ycap > 0x7fffffffL ||
      every < 1 || every > 4096) {
    set_error("gc_amd_record_batch_device: bad argument");
    return -1;
  }
  int rc = 0;
  DevCtx* cp = enter(stream, &rc);
  if (!cp) return rc;
  GcBatchArgs b{nCell, Tstart, Tend, C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL, Perm, ISTAT, RSTAT, IERR};
  GcRecord rec{recH, recPrev, recYslot, head, recY, recK, nullptr, nullptr,
               (unsigned)cap, (unsigned)ycap, every};
  rc = record_solve(*cp, b, rec, used, yused, cp->stream, tKernel);
  cp->mu.unlock();
  return rc;
}

int gc_amd_adjoint_batch_device(int nCell, double Tstart, double Tend, const double* C,
                                const double* RCONST, const int* ICNTRL, const int* Perm,
                                const double* recH, const int* recPrev, const int* recYslot,
                                const int* head, long cap, const double* recY, long ycap,
                                int every, const double* recK, const double* lamC, double* gC,
                                double* gRCONST, int* IERR, void* stream, double* tKernel) {
  if (nCell < 0 || cap < 0 || cap > 0x7fffffffL || ycap < 0 || ycap > 0x7fffffffL ||
      every < 1 || every > 64) {
    set_error("gc_amd_adjoint_batch_device: bad argument");
    return -1;
  }
  int rc = 0;
  DevCtx* cp = enter(stream, &rc);
  if (!cp) return rc;
  GcAdjArgs a{nCell, Tstart, Tend, C, RCONST, ICNTRL, Perm, recY, recK, recH, recPrev, recYslot,
              head, (unsigned)cap, (unsigned)ycap, every, 0, lamC, gC, gRCONST, IERR};
  rc = adjoint_solve(*cp, a, cp->stream, tKernel);
  cp->mu.unlock();
  return rc;
}

int gc_gpu_set_device(int dev) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || dev < 0 || dev >= n) return -1;
  if (hipSetDevice(dev) != hipSuccess) return -1;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return -1;
  return cur;
}

void gc_gpu_info(int* nDev, char* devName, int nameLen) {
  *nDev = 0;
  if (nameLen > 0) devName[0] = '\0';
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return;
  *nDev = n;
  if (n > 0 && nameLen > 1) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess) {
      int m = (int)strnlen(prop.name, sizeof(prop.name));
      if (m == 0) m = (int)strnlen(prop.gcnArchName, sizeof(prop.gcnArchName));
      const char* src = prop.name[0] ? prop.name : prop.gcnArchName;
      if (m > nameLen - 1) m = nameLen - 1;
      memcpy(devName, src, (size_t)m);
      devName[m] = '\0';
    }
  }
}

void gc_gpu_integrate_batch_dev(int dev, int nCell, double Tstart, double Tend, double* C,
                                const double* RCONST, const double* ATOL, const double* RTOL,
                                const int* ICNTRL, const double* RCNTRL, const int* Perm,
                                int* ISTAT, double* RSTAT, int* IERR, double* tTotal,
                                double* tKernel) {
  *tTotal = 0.0;
  *tKernel = 0.0;
  if (dev >= 0 && gc_gpu_set_device(dev) != dev) {
    fill_ierr(IERR, nCell, -99);
    return;
  }
  if (nCell <= 0) return;
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < nCell; ++i) {
    if (Perm[i] < 1 || Perm[i] > nCell) {
      set_error("gc_gpu_integrate_batch: Perm entry outside 1..nCell");
      fill_ierr(IERR, nCell, -98);
      return;
    }
  }
  int d = -1;
  DevCtx* cp = current_ctx(&d);
  if (!cp) {
    set_error("no current device");
    fill_ierr(IERR, nCell, -98);
    return;
  }
  DevCtx& c = *cp;
  std::lock_guard<std::mutex> lk(c.mu);
  hipError_t e = ctx_init(c, d);
  const MechDims md = mech_dims();
  const size_t n = (size_t)nCell;
  const size_t nspec = (size_t)(md.nvar + md.nfix);
  struct { DevBuf* b; const void* h; size_t bytes; } in[] = {
      {&c.C, C, n * nspec * 8}, {&c.RCONST, RCONST, n * md.nreact * 8},
      {&c.ATOL, ATOL, (size_t)md.nvar * 8}, {&c.RTOL, RTOL, (size_t)md.nvar * 8},
      {&c.ICNTRL, ICNTRL, n * 20 * 4}, {&c.RCNTRL, RCNTRL, n * 20 * 8}, {&c.Perm, Perm, n * 4}};
  for (auto& x : in) {
    if (e == hipSuccess) e = x.b->ensure(x.bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(x.b->p, x.h, x.bytes, hipMemcpyHostToDevice, c.stream);
  }
  if (e == hipSuccess) e = c.ISTAT.ensure(n * 8 * 4);
  if (e == hipSuccess) e = c.RSTAT.ensure(n * 3 * 8);
  if (e == hipSuccess) e = c.IERR.ensure(n * 4);
  if (e != hipSuccess) {
    set_error("staging", e);
    fill_ierr(IERR, nCell, -98);
    return;
  }
  GcBatchArgs b;
  b.nCell = nCell;
  b.Tstart = Tstart;
  b.Tend = Tend;
  b.C = (double*)c.C.p;
  b.RCONST = (const double*)c.RCONST.p;
  b.ATOL = (const double*)c.ATOL.p;
  b.RTOL = (const double*)c.RTOL.p;
  b.ICNTRL = (const int*)c.ICNTRL.p;
  b.RCNTRL = (const double*)c.RCNTRL.p;
  b.Perm = (const int*)c.Perm.p;
  b.ISTAT = (int*)c.ISTAT.p;
  b.RSTAT = (double*)c.RSTAT.p;
  b.IERR = (int*)c.IERR.p;
  const int kernel = env_kernel();
  const int rc = device_solve(c, kernel, kernel == GC_AMD_KERNEL_PLAIN ? env_chunk(nCell) : 0, b,
                              c.stream, tKernel);
  if (rc != 0) {
    fill_ierr(IERR, nCell, -98);
    return;
  }
  struct { void* h; const DevBuf* b; size_t bytes; } out[] = {
      {C, &c.C, n * nspec * 8}, {ISTAT, &c.ISTAT, n * 8 * 4}, {RSTAT, &c.RSTAT, n * 3 * 8},
      {IERR, &c.IERR, n * 4}};
  for (auto& x : out)
    if (e == hipSuccess) e = hipMemcpyAsync(x.h, x.b->p, x.bytes, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  if (e != hipSuccess) {
    set_error("copy back", e);
    fill_ierr(IERR, nCell, -98);
    return;
  }
  *tTotal = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void gc_gpu_integrate_batch(int nCell, double Tstart, double Tend, double* C,
                            const double* RCONST, const double* ATOL, const double* RTOL,
                            const int* ICNTRL, const double* RCNTRL, const int* Perm, int* ISTAT,
                            double* RSTAT, int* IERR, double* tTotal, double* tKernel) {
  gc_gpu_integrate_batch_dev(-1, nCell, Tstart, Tend, C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL, Perm,
                             ISTAT, RSTAT, IERR, tTotal, tKernel);
}

void gc_gpu_selftest(int* rc, int* nsteps, double* tTotal) {
  const MechDims md = mech_dims();
  std::vector<double> C(md.nvar + md.nfix, 1.0e6), RCONST(md.nreact, 0.0);
  std::vector<double> ATOL(md.nvar, 1.0e-2), RTOL(md.nvar, 0.5e-2), RCNTRL(20, 0.0), RSTAT(3);
  std::vector<int> ICNTRL(20, 0), ISTAT(8), IERR(1), Perm(1, 1);
  ICNTRL[0] = 1;
  ICNTRL[1] = 0;
  ICNTRL[2] = 4;
  ICNTRL[6] = 1;
  ICNTRL[14] = -1;
  double tK = 0.0;
  gc_gpu_integrate_batch(1, 0.0, 1200.0, C.data(), RCONST.data(), ATOL.data(), RTOL.data(),
                         ICNTRL.data(), RCNTRL.data(), Perm.data(), ISTAT.data(), RSTAT.data(),
                         IERR.data(), tTotal, &tK);
  *rc = IERR[0];
  *nsteps = ISTAT[2];
}

int gc_amd_integrate_batch_device(int kernel, int chunk, int nCell, double Tstart, double Tend,
                                  double* C, const double* RCONST, const double* ATOL,
                                  const double* RTOL, const int* ICNTRL, const double* RCNTRL,
                                  const int* Perm, int* ISTAT, double* RSTAT, int* IERR,
                                  void* stream, double* tKernel) {
  int d = -1;
  DevCtx* cp = current_ctx(&d);
  if (!cp) {
    set_error("no current device");
    return -3;
  }
  std::lock_guard<std::mutex> lk(cp->mu);
  hipError_t e = ctx_init(*cp, d);
  if (e != hipSuccess) {
    set_error("device init", e);
    return -3;
  }
  GcBatchArgs b{nCell, Tstart, Tend, C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL, Perm, ISTAT, RSTAT, IERR};
  // The solve runs on the library's stream, ordered after all work already
  // queued on the caller's stream (NULL: the null stream), so inputs the
  // caller has just written on its stream are complete before any kernel
  // reads them.  The call is blocking, so later caller work sees the outputs.
  if ((e = hipEventRecord(cp->ev_in, (hipStream_t)stream)) != hipSuccess ||
      (e = hipStreamWaitEvent(cp->stream, cp->ev_in, 0)) != hipSuccess) {
    set_error("ordering after the caller's stream", e);
    return -3;
  }
  return device_solve(*cp, kernel, chunk, b, cp->stream, tKernel);
}

int gc_amd_integrate_batch_noalloc(int kernel, int nCell, double Tstart, double Tend, double* C,
                                   const double* RCONST, const double* ATOL, const double* RTOL,
                                   const int* ICNTRL, const double* RCNTRL, const int* Perm,
                                   int* ISTAT, double* RSTAT, int* IERR, void* stream,
                                   double* tKernel) {
  if (tKernel) *tKernel = 0.0;
  if (nCell < 0 || (kernel != GC_AMD_KERNEL_PLAIN && kernel != GC_AMD_KERNEL_PERSISTENT)) {
    set_error("gc_amd_integrate_batch_noalloc: bad argument");
    return -1;
  }
  const long wave_bytes = ws_per_slot() * 8L * 64L;
  const unsigned max_waves = (unsigned)((kNoAllocBytes - kNoAllocHead) / (size_t)wave_bytes);
  if (kernel == GC_AMD_KERNEL_PLAIN && (unsigned)nCell > max_waves * 64u) {
    set_error("gc_amd_integrate_batch_noalloc: nCell exceeds the cells the plain kernel can "
              "solve in the 256 MiB scratch (one workspace slot per cell)");
    return -1;
  }
  if (max_waves < 1u) {
    set_error("gc_amd_integrate_batch_noalloc: one wave's workspace exceeds the scratch");
    return -1;
  }
  int rc = 0;
  DevCtx* cp = enter(stream, &rc);
  if (!cp) return rc;
  DevCtx& c = *cp;
  if (!c.scratch) {
    set_error("gc_amd_integrate_batch_noalloc: scratch address", c.scratch_err);
    c.mu.unlock();
    return -3;
  }
  if (nCell == 0) {
    c.mu.unlock();
    return 0;
  }
  hipStream_t st = c.stream;
  unsigned* q = (unsigned*)c.scratch;
  int* flag = (int*)(c.scratch + 256);
  double* ws = (double*)(c.scratch + kNoAllocHead);
  rc = check_perm(c, Perm, nCell, st, flag);
  if (rc != 0) {
    c.mu.unlock();
    return rc;
  }
  Tuning t = current_tuning();
  unsigned S = 0;
  if (kernel == GC_AMD_KERNEL_PLAIN) {
    S = slots_for(c, kernel, t, nCell);
  } else {
    // a resident grid of whole work-groups inside the scratch's slots
    if ((unsigned)(t.block / 64) > max_waves) t.block = 64 * (int)max_waves;
    S = slots_for(c, kernel, t, nCell, max_waves / (unsigned)(t.block / 64));
  }
  GcBatchArgs b{nCell, Tstart, Tend, C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL, Perm, ISTAT, RSTAT, IERR};
  hipError_t e;
  if ((e = hipEventRecord(c.ev0, st)) != hipSuccess ||
      (e = launch_chunk(kernel, t, b, S, ws, q, st)) != hipSuccess ||
      (e = hipEventRecord(c.ev1, st)) != hipSuccess || (e = hipStreamSynchronize(st)) != hipSuccess) {
    set_error("gc_amd_integrate_batch_noalloc: solve", e);
    c.mu.unlock();
    return -3;
  }
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess && tKernel) *tKernel = ms * 1e-3;
  c.mu.unlock();
  return 0;
}

int gc_amd_set_tuning(int block, int blocks_per_cu, int cells_per_wave, int grain, int variant) {
  (void)current_tuning();  // apply GC_AMD_TUNE first so explicit values win
  if ((block > 0 && (block < 64 || block > 256 || block % 64 != 0)) ||
      (blocks_per_cu > 32) || (cells_per_wave > 64) || (grain > 4096) ||
      (variant > 0 && variant > 3))
    return -1;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (block > 0) g_tune.block = block;
  if (blocks_per_cu >= 0) g_tune.blocks_per_cu = blocks_per_cu;
  if (cells_per_wave > 0) g_tune.cells_per_wave = cells_per_wave;
  if (grain > 0) g_tune.grain = grain;
  if (variant > 0) g_tune.variant = variant;
  return 0;
}

void gc_amd_get_tuning(int* block, int* blocks_per_cu, int* cells_per_wave, int* grain,
                       int* variant) {
  const Tuning t = current_tuning();
  *block = t.block;
  *blocks_per_cu = t.blocks_per_cu;
  *cells_per_wave = t.cells_per_wave;
  *grain = t.grain;
  *variant = t.variant;
}

void gc_amd_mech_info(int* nvar, int* nfix, int* nreact, int* lu_nonzero) {
  const MechDims d = mech_dims();
  *nvar = d.nvar;
  *nfix = d.nfix;
  *nreact = d.nreact;
  *lu_nonzero = d.nnz;
}

long gc_amd_workspace_doubles_per_slot(void) { return ws_per_slot(); }

const char* gc_amd_last_error(void) { return g_last_error.c_str(); }

void gc_amd_release(void) {
  int d = -1;
  DevCtx* cp = current_ctx(&d);
  if (!cp) return;
  std::lock_guard<std::mutex> lk(cp->mu);
  cp->graph.release();
  DevBuf* bufs[] = {&cp->ws, &cp->queue, &cp->flag, &cp->C, &cp->RCONST, &cp->ATOL, &cp->RTOL,
                    &cp->ICNTRL, &cp->RCNTRL, &cp->Perm, &cp->ISTAT, &cp->RSTAT, &cp->IERR};
  for (DevBuf* b : bufs) b->release();
}

}  // extern "C"
