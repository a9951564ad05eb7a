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
#include <strings.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <hipcub/hipcub.hpp>

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
// The launch the autotuner stored for this library's mechanism, compiled in by
// the build (-DGC_TUNE_*); these values where none is stored.
#ifndef GC_TUNE_BLOCK
#define GC_TUNE_BLOCK 256
#endif
#ifndef GC_TUNE_BLOCKS_PER_CU
#define GC_TUNE_BLOCKS_PER_CU 0
#endif
#ifndef GC_TUNE_CELLS_PER_WAVE
#define GC_TUNE_CELLS_PER_WAVE 64
#endif
#ifndef GC_TUNE_GRAIN
#define GC_TUNE_GRAIN 64
#endif
#ifndef GC_TUNE_VARIANT
#define GC_TUNE_VARIANT 1
#endif
static_assert(GC_TUNE_BLOCK >= 64 && GC_TUNE_BLOCK <= 256 && GC_TUNE_BLOCK % 64 == 0 &&
                  GC_TUNE_BLOCKS_PER_CU >= 0 && GC_TUNE_BLOCKS_PER_CU <= 32 &&
                  GC_TUNE_CELLS_PER_WAVE >= 1 && GC_TUNE_CELLS_PER_WAVE <= 64 &&
                  GC_TUNE_GRAIN >= 1 && GC_TUNE_GRAIN <= 4096 && GC_TUNE_VARIANT >= 1 &&
                  GC_TUNE_VARIANT <= 3,
              "compiled-in launch outside the ranges gc_amd_set_tuning accepts");

struct Tuning {
  int block = GC_TUNE_BLOCK;
  int blocks_per_cu = GC_TUNE_BLOCKS_PER_CU;
  int cells_per_wave = GC_TUNE_CELLS_PER_WAVE;
  int grain = GC_TUNE_GRAIN;
  int variant = GC_TUNE_VARIANT;
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

// ---- overlapped upload of the host-array entry point (process-wide) -----------
// mode 1: a call whose launch order is the identity (the model gathers its
// batch in cost order and passes Perm = 1..nCell) uploads its inputs in chunks
// of launch positions while the persistent kernel already runs, each wave
// waiting for its positions' chunk (GcGate); mode 0: upload, then solve;
// mode 2: as mode 1, but the flags are never set, so every wait runs to its
// bound (the failure path, for tests).  cells = chunk size in cells (0: what
// fits one 32 MB staging slot); spin = sleeps of ~3.4 us a wave waits for a
// chunk.  A mode-1 call whose chunks do not land within the bound (copies
// that cannot run beside the kernel) switches the overlap off for the rest
// of the process and is solved again with the upload first; in mode 2 the
// call fails (IERR -98).
struct GateConfig {
  int mode = 1;
  int cells = 0;
  unsigned long long spin = 600000ull;  // ~2 s; the whole 4x5 upload takes ~70 ms
};
std::mutex g_gate_mu;
GateConfig g_gate;
bool g_gate_env_read = false;

bool env_set(const char* name) {
  const char* s = getenv(name);
  return s && *s && strcmp(s, "0") != 0 && strcasecmp(s, "false") != 0;
}

GateConfig current_gate() {
  std::lock_guard<std::mutex> lk(g_gate_mu);
  if (!g_gate_env_read) {
    g_gate_env_read = true;
    // Where the device work is serialized around each dispatch (rocprofv3
    // counter collection, the launch-blocking debug modes), the copies cannot
    // land while the gated kernel runs: the overlap starts off there.
    for (const char* v : {"ROCPROF_COUNTER_COLLECTION", "HIP_LAUNCH_BLOCKING",
                          "AMD_SERIALIZE_KERNEL", "AMD_SERIALIZE_COPY"})
      if (env_set(v)) g_gate.mode = 0;
  }
  return g_gate;
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

// ---- pinned staging of host arrays --------------------------------------------
// Host arrays move through pinned buffers in chunks: four host threads copy
// chunks into (or out of) their own pair of pinned buffers while the DMA
// engines move the chunks before, so the host copies and the transfers
// overlap and the transfers run at the pinned rate.  The threads share two
// streams (thread t issues on stream t % 2), so staging adds two streams to
// the process whatever the thread count.  A device without pinned buffers
// (hipHostMalloc failed) falls back to plain pageable copies.
struct Piece {
  void* dev;
  void* host;
  size_t bytes;
};

struct Stager {
  static constexpr int kThreads = 4, kStreams = 2, kSlots = 2;
  static constexpr size_t kChunk = size_t(32) << 20;
  bool tried = false, ok = false;
  unsigned char* pin[kThreads][kSlots] = {};
  unsigned* ones = nullptr;  // pinned 1s: the source of the input-gate flags
  hipStream_t st[kStreams] = {};
  hipEvent_t ev[kThreads][kSlots] = {};
  hipEvent_t done[kStreams] = {};

  void init() {
    if (tried) return;
    tried = true;
    for (int i = 0; i < kStreams; ++i) {
      if (hipStreamCreateWithFlags(&st[i], hipStreamNonBlocking) != hipSuccess) return;
      if (hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess) return;
    }
    for (int t = 0; t < kThreads; ++t) {
      for (int k = 0; k < kSlots; ++k) {
        if (hipHostMalloc((void**)&pin[t][k], kChunk, hipHostMallocDefault) != hipSuccess) {
          pin[t][k] = nullptr;
          (void)hipGetLastError();
          return;
        }
        if (hipEventCreateWithFlags(&ev[t][k], hipEventDisableTiming) != hipSuccess) return;
      }
    }
    if (hipHostMalloc((void**)&ones, sizeof(unsigned), hipHostMallocDefault) != hipSuccess) {
      ones = nullptr;
      (void)hipGetLastError();
      return;
    }
    *ones = 1u;
    ok = true;
  }

  hipStream_t stream_of(int t) const { return st[t % kStreams]; }

  // work(t) on every thread t (thread 0 is the caller's); the first error
  static hipError_t run(const std::function<hipError_t(int)>& work) {
    hipError_t err[kThreads];
    std::vector<std::thread> helpers;
    for (int t = 1; t < kThreads; ++t) helpers.emplace_back([&, t] { err[t] = work(t); });
    err[0] = work(0);
    for (auto& h : helpers) h.join();
    for (int t = 0; t < kThreads; ++t)
      if (err[t] != hipSuccess) return err[t];
    return hipSuccess;
  }

  // the staging streams start once `ev_start` has been reached
  hipError_t wait_for(hipEvent_t ev_start) {
    for (int i = 0; i < kStreams; ++i)
      if (hipStreamWaitEvent(st[i], ev_start, 0) != hipSuccess) return hipErrorUnknown;
    return hipSuccess;
  }

  // chunk c of the pieces -> (piece, offset, length)
  static std::vector<Piece> chunks(const Piece* p, int n) {
    std::vector<Piece> out;
    for (int i = 0; i < n; ++i)
      for (size_t o = 0; o < p[i].bytes; o += kChunk)
        out.push_back({(char*)p[i].dev + o, (char*)p[i].host + o, std::min(kChunk, p[i].bytes - o)});
    return out;
  }

  // Host -> device, ordered after `after` (the copies start once it reaches
  // this point) and before the work queued later on `consumer`.
  hipError_t to_device(int dev, const Piece* p, int n, hipStream_t after, hipStream_t consumer) {
    if (!ok) {
      for (int i = 0; i < n; ++i) {
        hipError_t e = hipMemcpyAsync(p[i].dev, p[i].host, p[i].bytes, hipMemcpyHostToDevice,
                                      consumer);
        if (e != hipSuccess) return e;
      }
      return hipSuccess;
    }
    const std::vector<Piece> c = chunks(p, n);
    if (hipEventRecord(done[0], after) != hipSuccess || wait_for(done[0]) != hipSuccess)
      return hipErrorUnknown;
    hipError_t r = run([&](int t) {
      hipError_t e = hipSetDevice(dev);
      for (size_t k = t, j = 0; k < c.size() && e == hipSuccess; k += kThreads, ++j) {
        const int s = (int)(j % kSlots);
        e = hipEventSynchronize(ev[t][s]);  // the slot's previous transfer is done
        if (e != hipSuccess) break;
        memcpy(pin[t][s], c[k].host, c[k].bytes);
        e = hipMemcpyAsync(c[k].dev, pin[t][s], c[k].bytes, hipMemcpyHostToDevice, stream_of(t));
        if (e == hipSuccess) e = hipEventRecord(ev[t][s], stream_of(t));
      }
      return e;
    });
    if (r != hipSuccess) return r;
    for (int i = 0; i < kStreams; ++i) {
      hipError_t e = hipEventRecord(done[i], st[i]);
      if (e == hipSuccess) e = hipStreamWaitEvent(consumer, done[i], 0);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }

  // Host -> device of a batch in chunks of `cells` consecutive cells (launch
  // positions) behind an input gate: chunk k of every piece (piece i holds
  // stride[i] bytes per cell) goes through one pinned slot, then flags[k] is
  // set to 1 on the same stream, i.e. once the chunk's data has landed
  // (unless `signal` is false: the flags are then never set).  The streams
  // start once `start` has been reached; the call returns when every copy is
  // issued, and sync() waits for them.
  hipError_t to_device_gated(int dev, const Piece* p, const size_t* stride, int n, int nCell,
                             int cells, unsigned* flags, bool signal, hipEvent_t start) {
    const int K = (nCell + cells - 1) / cells;
    if (wait_for(start) != hipSuccess) return hipErrorUnknown;
    return run([&](int t) {
      hipError_t e = hipSetDevice(dev);
      for (int k = t, j = 0; k < K && e == hipSuccess; k += kThreads, ++j) {
        const int s = j % kSlots;
        e = hipEventSynchronize(ev[t][s]);  // the slot's previous transfer is done
        const size_t c0 = (size_t)k * (size_t)cells;
        const size_t m = std::min((size_t)cells, (size_t)nCell - c0);
        size_t off = 0;
        for (int i = 0; i < n && e == hipSuccess; ++i) {
          const size_t bytes = m * stride[i];
          memcpy(pin[t][s] + off, (const char*)p[i].host + c0 * stride[i], bytes);
          e = hipMemcpyAsync((char*)p[i].dev + c0 * stride[i], pin[t][s] + off, bytes,
                             hipMemcpyHostToDevice, stream_of(t));
          off += bytes;
        }
        if (e == hipSuccess && signal)
          e = hipMemcpyAsync(flags + k, ones, sizeof(unsigned), hipMemcpyHostToDevice,
                             stream_of(t));
        if (e == hipSuccess) e = hipEventRecord(ev[t][s], stream_of(t));
      }
      return e;
    });
  }

  hipError_t sync() {
    hipError_t r = hipSuccess;
    for (int i = 0; i < kStreams; ++i) {
      hipError_t e = hipStreamSynchronize(st[i]);
      if (r == hipSuccess) r = e;
    }
    return r;
  }

  // Device -> host after the work queued on `producer`; returns when the
  // host arrays hold the data.
  hipError_t to_host(int dev, const Piece* p, int n, hipStream_t producer) {
    if (!ok) {
      for (int i = 0; i < n; ++i) {
        hipError_t e = hipMemcpyAsync(p[i].host, p[i].dev, p[i].bytes, hipMemcpyDeviceToHost,
                                      producer);
        if (e != hipSuccess) return e;
      }
      return hipStreamSynchronize(producer);
    }
    const std::vector<Piece> c = chunks(p, n);
    if (hipEventRecord(done[0], producer) != hipSuccess || wait_for(done[0]) != hipSuccess)
      return hipErrorUnknown;
    return run([&](int t) {
      hipError_t e = hipSetDevice(dev);
      std::vector<size_t> mine;
      for (size_t k = t; k < c.size(); k += kThreads) mine.push_back(k);
      auto issue = [&](size_t j) {
        const int s = (int)(j % kSlots);
        hipError_t r = hipMemcpyAsync(pin[t][s], c[mine[j]].dev, c[mine[j]].bytes,
                                      hipMemcpyDeviceToHost, stream_of(t));
        if (r == hipSuccess) r = hipEventRecord(ev[t][s], stream_of(t));
        return r;
      };
      if (e == hipSuccess && !mine.empty()) e = issue(0);
      for (size_t j = 0; j < mine.size() && e == hipSuccess; ++j) {
        if (j + 1 < mine.size()) e = issue(j + 1);  // the other slot: its copy-out is done
        if (e == hipSuccess) e = hipEventSynchronize(ev[t][j % kSlots]);
        if (e == hipSuccess) memcpy(c[mine[j]].host, pin[t][j % kSlots], c[mine[j]].bytes);
      }
      return e;
    });
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
  DevBuf ws, queue, flag, descs;
  ChunkGraph graph;
  unsigned char* scratch = nullptr;  // this device's copy of `scratch`
  hipError_t scratch_err = hipSuccess;
  Stager stage;                      // pinned staging of the host-array entry points
  // staging for the host-array entry points
  DevBuf C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL, Perm, ISTAT, RSTAT, IERR;
  DevBuf gate;                 // input-gate flags of the overlapped upload, then its error word
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

// ---- device-resident batch across chemistry steps -----------------------------
__global__ void k_iota1(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i + 1;
}

// The host-array solve with the inputs uploaded behind the input gate (see
// GateConfig): the device's buffers are sized, the flags cleared, ATOL, RTOL
// and the identity launch order set, the gated persistent kernel launched,
// then the chunks of C, RCONST, ICNTRL and RCNTRL (launch positions = cells)
// copied while it runs.  Returns 0, or -2 / -3 (allocation / runtime error)
// or -4 (a chunk's flag was not seen within the bound).
int gated_solve(DevCtx& c, int dev, const GateConfig& g, int nCell, double Tstart, double Tend,
                const double* C, const double* RCONST, const double* ATOL, const double* RTOL,
                const int* ICNTRL, const double* RCNTRL, double* tKernel) {
  *tKernel = 0.0;
  const MechDims md = mech_dims();
  const size_t n = (size_t)nCell;
  const size_t nspec = (size_t)(md.nvar + md.nfix);
  const size_t stride[4] = {nspec * 8, (size_t)md.nreact * 8, 20 * 4, 20 * 8};
  const size_t per_cell = stride[0] + stride[1] + stride[2] + stride[3];
  // chunks of a multiple of 64 cells: every chunk boundary of every array then
  // falls on a 128-byte line boundary (the arrays' per-cell sizes are multiples
  // of 8 bytes), so no line holds both a landed cell and one still in flight
  int cells = g.cells > 0 ? g.cells : (int)(Stager::kChunk / per_cell);
  if ((size_t)cells * per_cell > Stager::kChunk) cells = (int)(Stager::kChunk / per_cell);
  cells -= cells % 64;
  if (cells < 64) cells = 64;
  const int K = (nCell + cells - 1) / cells;
  const Tuning t = current_tuning();
  const unsigned S = slots_for(c, GC_AMD_KERNEL_PERSISTENT, t, nCell);
  hipError_t e = hipSuccess;
  DevBuf* bufs[] = {&c.C, &c.RCONST, &c.ICNTRL, &c.RCNTRL};
  for (int i = 0; i < 4 && e == hipSuccess; ++i) e = bufs[i]->ensure(n * stride[i]);
  if (e == hipSuccess) e = c.ATOL.ensure((size_t)md.nvar * 8);
  if (e == hipSuccess) e = c.RTOL.ensure((size_t)md.nvar * 8);
  if (e == hipSuccess) e = c.Perm.ensure(n * 4);
  if (e == hipSuccess) e = c.ISTAT.ensure(n * 8 * 4);
  if (e == hipSuccess) e = c.RSTAT.ensure(n * 3 * 8);
  if (e == hipSuccess) e = c.IERR.ensure(n * 4);
  if (e == hipSuccess) e = c.gate.ensure(((size_t)K + 1) * sizeof(unsigned));
  if (e == hipSuccess) e = c.ws.ensure((size_t)S * (size_t)ws_per_slot() * sizeof(double));
  if (e == hipSuccess) e = c.queue.ensure(256);
  if (e != hipSuccess) {
    set_error("hipMalloc (overlapped upload)", e);
    return -2;
  }
  unsigned* flags = (unsigned*)c.gate.p;
  unsigned* gerr = flags + K;
  const hipStream_t st = c.stream;
  hipLaunchKernelGGL(k_zero_u32, dim3((K + 1 + 255) / 256), dim3(256), 0, st, flags, K + 1);
  hipLaunchKernelGGL(k_iota1, dim3((nCell + 255) / 256), dim3(256), 0, st, (int*)c.Perm.p, nCell);
  if ((e = hipGetLastError()) == hipSuccess)
    e = hipMemcpyAsync(c.ATOL.p, ATOL, (size_t)md.nvar * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c.RTOL.p, RTOL, (size_t)md.nvar * 8, hipMemcpyHostToDevice, st);
  // the copies may start once the flags are clear; the kernel does not wait for them
  if (e == hipSuccess) e = hipEventRecord(c.stage.done[0], st);
  if (e != hipSuccess) {
    set_error("overlapped upload set-up", e);
    return -3;
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
  const GcGate gate{flags, (unsigned)cells, g.spin, gerr};
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, (unsigned*)c.queue.p, 1);
  if ((e = hipEventRecord(c.ev0, st)) == hipSuccess) {
    const unsigned waves = S / 64u;
    unsigned fair = ((unsigned)nCell + waves - 1u) / waves;
    if (fair < 1u) fair = 1u;
    const int grain = (unsigned)t.grain > fair ? (int)fair : t.grain;
    if (t.variant == 2)
      e = gc_launch_persistent_gated_v2(&b, (double*)c.ws.p, S, (unsigned*)c.queue.p, t.block,
                                        t.cells_per_wave, grain, &gate, st);
    else if (t.variant == 3)
      e = gc_launch_persistent_gated_v3(&b, (double*)c.ws.p, S, (unsigned*)c.queue.p, t.block,
                                        t.cells_per_wave, grain, &gate, st);
    else
      e = gc_launch_persistent_gated_v1(&b, (double*)c.ws.p, S, (unsigned*)c.queue.p, t.block,
                                        t.cells_per_wave, grain, &gate, st);
  }
  if (e == hipSuccess) e = hipEventRecord(c.ev1, st);
  // the inputs follow the launch, chunk by chunk, while it runs
  const Piece pin[4] = {{c.C.p, const_cast<double*>(C), 0}, {c.RCONST.p, const_cast<double*>(RCONST), 0},
                        {c.ICNTRL.p, const_cast<int*>(ICNTRL), 0},
                        {c.RCNTRL.p, const_cast<double*>(RCNTRL), 0}};
  hipError_t ue = hipSuccess;
  if (e == hipSuccess)
    ue = c.stage.to_device_gated(dev, pin, stride, 4, nCell, cells, flags, g.mode != 2,
                                 c.stage.done[0]);
  // the solve ends whatever the copies did: a chunk that never lands ends its
  // waves' waits at the bound
  hipError_t se = hipStreamSynchronize(st);
  hipError_t ce = c.stage.sync();
  if (e != hipSuccess || ue != hipSuccess || se != hipSuccess || ce != hipSuccess) {
    set_error("overlapped upload / solve",
              e != hipSuccess ? e : ue != hipSuccess ? ue : se != hipSuccess ? se : ce);
    return -3;
  }
  unsigned bad = 0;
  if ((e = hipMemcpy(&bad, gerr, sizeof(unsigned), hipMemcpyDeviceToHost)) != hipSuccess) {
    set_error("overlapped upload: gate status", e);
    return -3;
  }
  if (bad) {
    set_error("overlapped upload: an input chunk did not land within the wait bound");
    return -4;
  }
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess) *tKernel = ms * 1e-3;
  return 0;
}

// sort keys of the next launch order: the cell's internal step count of the
// step just solved (ISTAT(3)), and the warm start RCNTRL(3) = RSTAT(3) (Hnew)
__global__ void k_carry_over(const int* ISTAT, const double* RSTAT, double* RCNTRL, int* keys,
                             int n, int warm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = ISTAT[(size_t)i * 8 + 2];
  if (warm) RCNTRL[(size_t)i * 20 + 2] = RSTAT[(size_t)i * 3 + 2];
}

}  // namespace

// ---- several batches in one launch ------------------------------------------------
namespace {

int check_batches(int nBatch, const GcAmdBatchDesc* batch, unsigned* qn) {
  if (nBatch < 1 || nBatch > 4096 || !batch) return -1;
  unsigned long long tot = 0;
  for (int k = 0; k < nBatch; ++k) {
    if (batch[k].nCell < 0) return -1;
    tot += (unsigned)batch[k].nCell;
  }
  if (tot > 0x7fffffffull) return -1;
  *qn = (unsigned)tot;
  return 0;
}

// Device descriptors of the batches (queue offsets = cells of the batches
// before), every launch order checked; done (host-visible, nBatch counters,
// may be null) counts each batch's finished cells.  Returns 0 or a negative
// code.
int upload_descs(DevCtx& c, int nBatch, const GcAmdBatchDesc* batch, hipStream_t st,
                 unsigned* done) {
  std::vector<GcBatchDesc> h((size_t)nBatch);
  unsigned q0 = 0;
  for (int k = 0; k < nBatch; ++k) {
    const GcAmdBatchDesc& x = batch[k];
    h[k].a = GcBatchArgs{x.nCell, x.Tstart, x.Tend, x.C, x.RCONST, x.ATOL, x.RTOL, x.ICNTRL,
                         x.RCNTRL, x.Perm, x.ISTAT, x.RSTAT, x.IERR};
    h[k].q0 = q0;
    h[k].done = done ? done + k : nullptr;
    q0 += (unsigned)x.nCell;
    if (x.nCell > 0) {
      const int rc = check_perm(c, x.Perm, x.nCell, st);
      if (rc) return rc;
    }
  }
  hipError_t e = c.descs.ensure(h.size() * sizeof(GcBatchDesc));
  if (e == hipSuccess)
    e = hipMemcpyAsync(c.descs.p, h.data(), h.size() * sizeof(GcBatchDesc), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    set_error("batch descriptors", e);
    return -2;
  }
  return 0;
}

// Finished-cell counters of a launch with progress reports: one per batch,
// pinned host memory the kernel adds to with system-scope atomics and the
// waiting thread reads.
struct DoneCounters {
  unsigned* p = nullptr;
  ~DoneCounters() {
    if (p) (void)hipHostFree(p);
  }
  int init(int n) {
    if (hipHostMalloc((void**)&p, sizeof(unsigned) * (size_t)n, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      set_error("progress counters: hipHostMalloc");
      return -2;
    }
    memset(p, 0, sizeof(unsigned) * (size_t)n);
    return 0;
  }
};

// Calls progress on each change of the finished-cell counts (polled every
// 0.2 s) until `end` (recorded after the launch) has completed, and once more
// at the end.  size(k) is batch k's cell count.
template <class Size>
void report_progress(hipEvent_t end, const unsigned* done, int nBatch, Size size, unsigned qn,
                     gc_amd_progress_fn progress, void* user) {
  long long last = -1;
  for (;;) {
    const hipError_t q = hipEventQuery(end);
    int nb = 0;
    long long cells = 0;
    for (int k = 0; k < nBatch; ++k) {
      const unsigned v = __atomic_load_n(done + k, __ATOMIC_RELAXED);
      cells += v;
      nb += v >= (unsigned)size(k);
    }
    if (cells != last) {
      progress(nb, cells, (long long)qn, user);
      last = cells;
    }
    if (q != hipErrorNotReady) break;
    std::this_thread::sleep_for(std::chrono::milliseconds(200));
  }
}

int multi_solve(DevCtx& c, int nBatch, const GcAmdBatchDesc* batch, unsigned qn,
                const GcRecord* rec_in, long* used, long* yused, hipStream_t st,
                double* tKernel, gc_amd_progress_fn progress = nullptr, void* user = nullptr) {
  const Tuning t = current_tuning();
  DoneCounters done;
  if (progress && done.init(nBatch)) return -2;
  int rc = upload_descs(c, nBatch, batch, st, done.p);
  if (rc) return rc;
  if (qn == 0) return 0;
  hipError_t e;
  const unsigned S = slots_for(c, GC_AMD_KERNEL_PERSISTENT, t, (int)qn);
  if ((e = c.ws.ensure((size_t)S * (size_t)ws_per_slot() * sizeof(double))) != hipSuccess) {
    set_error("hipMalloc(workspace)", e);
    return -2;
  }
  if ((e = c.queue.ensure(256)) != hipSuccess) { set_error("hipMalloc(queue)", e); return -2; }
  unsigned* q = (unsigned*)c.queue.p;
  GcRecord rec{};
  if (rec_in) {
    rec = *rec_in;
    rec.count = q + 1;
    rec.ycount = q + 2;
    hipLaunchKernelGGL(k_fill_i32, dim3((qn + 255) / 256), dim3(256), 0, st, rec.head, (int)qn, -1);
  }
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, q, 3);
  const unsigned waves = S / 64u;
  unsigned fair = (qn + waves - 1u) / waves;
  if (fair < 1u) fair = 1u;
  const int grain = (unsigned)t.grain > fair ? (int)fair : t.grain;
  const GcBatchDesc* d = (const GcBatchDesc*)c.descs.p;
  double* ws = (double*)c.ws.p;
  const GcRecord* rp = rec_in ? &rec : nullptr;
  if ((e = hipEventRecord(c.ev0, st)) != hipSuccess) { set_error("hipEventRecord", e); return -3; }
  e = (t.variant == 2) ? gc_launch_persistent_multi_v2(d, nBatch, qn, ws, S, q, t.block, t.cells_per_wave, grain, rp, st)
      : (t.variant == 3) ? gc_launch_persistent_multi_v3(d, nBatch, qn, ws, S, q, t.block, t.cells_per_wave, grain, rp, st)
                         : gc_launch_persistent_multi_v1(d, nBatch, qn, ws, S, q, t.block, t.cells_per_wave, grain, rp, st);
  if (e != hipSuccess) { set_error("launch", e); return -3; }
  if ((e = hipEventRecord(c.ev1, st)) != hipSuccess) { set_error("hipEventRecord", e); return -3; }
  if (progress)
    report_progress(c.ev1, done.p, nBatch, [&](int k) { return batch[k].nCell; }, qn, progress,
                    user);
  unsigned counts[2] = {0, 0};
  if (rec_in &&
      (e = hipMemcpyAsync(counts, rec.count, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, st)) != hipSuccess) {
    set_error("record counts", e);
    return -3;
  }
  if ((e = hipStreamSynchronize(st)) != hipSuccess) { set_error("solve", e); return -3; }
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess && tKernel) *tKernel = ms * 1e-3;
  if (rec_in) {
    if (used) *used = (long)counts[0];
    if (yused) *yused = (long)counts[1];
    if (counts[0] > rec.cap || counts[1] > rec.ycap) {
      set_error("record pools too small: retry with cap >= *used and ycap >= *yused");
      return -4;
    }
  }
  return 0;
}

// The reverse sweep over the batches h (arguments filled in but for
// slot_doubles; one record pool and `every` for all) in one persistent launch:
// launch orders checked, one workspace slot per resident lane.
int adjoint_sweep(DevCtx& c, std::vector<GcAdjDesc>& h, hipStream_t st,
                  gc_amd_progress_fn progress, void* user, double* tKernel) {
  if (tKernel) *tKernel = 0.0;
  const int nb = (int)h.size();
  const Tuning t = current_tuning();
  const GcVariantInfo vi = variant_info(t.variant);
  // adjoint block + the segment buffers (`every` start states, step sizes
  // and stage vectors)
  const long slot_doubles = vi.adj_doubles_per_slot + (long)h[0].a.every * (5L * vi.nvar + 1);
  DoneCounters done;
  if (progress && done.init(nb)) return -2;
  unsigned long long q0 = 0;
  for (int k = 0; k < nb; ++k) {
    h[k].a.slot_doubles = slot_doubles;
    h[k].q0 = (unsigned)q0;
    h[k].done = done.p ? done.p + k : nullptr;
    q0 += (unsigned)h[k].a.nCell;
    if (h[k].a.nCell > 0) {
      const int rc = check_perm(c, h[k].a.Perm, h[k].a.nCell, st);
      if (rc) return rc;
    }
  }
  const unsigned qn = (unsigned)q0;
  if (qn == 0) return 0;
  // resident waves of the sweep kernel: its occupancy target (its LDS row
  // segment fits that many per CU), or fewer when the launch tuning asks for
  // fewer waves per CU; never more than there are 64-cell groups
  const int block = 128;
  const unsigned groups = (qn + 63u) / 64u;
  int wpc = 4 * vi.waves_per_eu;
  if (t.blocks_per_cu > 0 && t.blocks_per_cu * (t.block / 64) < wpc)
    wpc = t.blocks_per_cu * (t.block / 64);
  unsigned waves = (unsigned)(c.n_cu * wpc);
  if (waves > groups) waves = groups;
  waves = (waves + 1u) & ~1u;  // whole blocks of two waves
  hipError_t e = c.descs.ensure(h.size() * sizeof(GcAdjDesc));
  if (e == hipSuccess)
    e = hipMemcpyAsync(c.descs.p, h.data(), h.size() * sizeof(GcAdjDesc), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = c.ws.ensure((size_t)waves * 64u * (size_t)slot_doubles * sizeof(double));
  if (e == hipSuccess) e = c.queue.ensure(256);
  if (e != hipSuccess) {
    set_error("adjoint sweep: device memory", e);
    return -2;
  }
  unsigned* q = (unsigned*)c.queue.p;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, q, 1);
  const GcAdjDesc* d = (const GcAdjDesc*)c.descs.p;
  double* ws = (double*)c.ws.p;
  if ((e = hipEventRecord(c.ev0, st)) != hipSuccess) { set_error("hipEventRecord", e); return -3; }
  e = (t.variant == 2)   ? gc_launch_adjoint_v2(d, nb, qn, ws, slot_doubles, waves, q, block, st)
      : (t.variant == 3) ? gc_launch_adjoint_v3(d, nb, qn, ws, slot_doubles, waves, q, block, st)
                         : gc_launch_adjoint_v1(d, nb, qn, ws, slot_doubles, waves, q, block, st);
  if (e != hipSuccess) { set_error("launch", e); return -3; }
  if ((e = hipEventRecord(c.ev1, st)) != hipSuccess) { set_error("hipEventRecord", e); return -3; }
  if (progress)
    report_progress(c.ev1, done.p, nb, [&](int k) { return h[k].a.nCell; }, qn, progress, user);
  if ((e = hipStreamSynchronize(st)) != hipSuccess) { set_error("adjoint", e); return -3; }
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess && tKernel) *tKernel = ms * 1e-3;
  return 0;
}

}  // namespace

struct GcAmdResident {
  int dev = -1;
  int n = 0;
  int steps = 0;
  DevBuf C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL, Perm, ISTAT, RSTAT, IERR;
  DevBuf keys, keys_out, cells, sort_tmp;
  size_t sort_bytes = 0;
};

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
  std::vector<GcAdjDesc> h(1);
  h[0].a = GcAdjArgs{nCell, Tstart, Tend, C, RCONST, ICNTRL, Perm, recY, recK, recH, recPrev,
                     recYslot, head, (unsigned)cap, (unsigned)ycap, every, 0, lamC, gC, gRCONST,
                     IERR};
  rc = adjoint_sweep(*cp, h, cp->stream, nullptr, nullptr, tKernel);
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
  const int kernel = env_kernel();
  const GateConfig gate = current_gate();
  bool identity = true;
  for (int i = 0; i < nCell && identity; ++i) identity = Perm[i] == i + 1;
  if (e == hipSuccess) c.stage.init();
  if (e == hipSuccess && gate.mode != 0 && identity && kernel == GC_AMD_KERNEL_PERSISTENT &&
      c.stage.ok) {
    // the model's call (cells gathered in cost order, Perm = 1..nCell): the
    // inputs stream in while the solve runs
    const int rc = gated_solve(c, d, gate, nCell, Tstart, Tend, C, RCONST, ATOL, RTOL, ICNTRL,
                               RCNTRL, tKernel);
    if (rc == -4 && gate.mode == 1) {
      // the copies did not run beside the kernel: no overlap in this process
      std::lock_guard<std::mutex> glk(g_gate_mu);
      g_gate.mode = 0;
    } else if (rc != 0) {
      fill_ierr(IERR, nCell, -98);
      return;
    } else {
      Piece pout[4] = {{c.C.p, C, n * nspec * 8}, {c.ISTAT.p, ISTAT, n * 8 * 4},
                       {c.RSTAT.p, RSTAT, n * 3 * 8}, {c.IERR.p, IERR, n * 4}};
      e = c.stage.to_host(d, pout, 4, c.stream);
      if (e != hipSuccess) {
        set_error("copy back", e);
        fill_ierr(IERR, nCell, -98);
        return;
      }
      *tTotal = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      return;
    }
  }
  struct { DevBuf* b; const void* h; size_t bytes; } in[] = {
      {&c.C, C, n * nspec * 8}, {&c.RCONST, RCONST, n * md.nreact * 8},
      {&c.ATOL, ATOL, (size_t)md.nvar * 8}, {&c.RTOL, RTOL, (size_t)md.nvar * 8},
      {&c.ICNTRL, ICNTRL, n * 20 * 4}, {&c.RCNTRL, RCNTRL, n * 20 * 8}, {&c.Perm, Perm, n * 4}};
  Piece pin[7];
  int np = 0;
  for (auto& x : in) {
    if (e == hipSuccess) e = x.b->ensure(x.bytes);
    pin[np++] = {x.b->p, const_cast<void*>(x.h), x.bytes};
  }
  if (e == hipSuccess) e = c.stage.to_device(d, pin, np, c.stream, c.stream);
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
  const int rc = device_solve(c, kernel, kernel == GC_AMD_KERNEL_PLAIN ? env_chunk(nCell) : 0, b,
                              c.stream, tKernel);
  if (rc != 0) {
    fill_ierr(IERR, nCell, -98);
    return;
  }
  Piece pout[4] = {{c.C.p, C, n * nspec * 8}, {c.ISTAT.p, ISTAT, n * 8 * 4},
                   {c.RSTAT.p, RSTAT, n * 3 * 8}, {c.IERR.p, IERR, n * 4}};
  if (e == hipSuccess) e = c.stage.to_host(d, pout, 4, c.stream);
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

int gc_amd_integrate_batch_masked(int kernel, int nCell, double Tstart, double Tend, double* C,
                                  const double* RCONST, const double* ATOL, const double* RTOL,
                                  const int* ICNTRL, const double* RCNTRL, const int* Perm,
                                  const unsigned* doFun, const unsigned* doJvs,
                                  const unsigned* doSlv, int* ISTAT, double* RSTAT, int* IERR,
                                  void* stream, double* tKernel) {
  if (tKernel) *tKernel = 0.0;
  int rc = 0;
  DevCtx* cp = enter(stream, &rc);
  if (!cp) return rc;
  GcBatchArgs b{nCell, Tstart, Tend, C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL, Perm, ISTAT, RSTAT, IERR};
  b.doFun = doFun;
  b.doJvs = doJvs;
  b.doSlv = doSlv;
  rc = device_solve(*cp, kernel, 0, b, cp->stream, tKernel);
  cp->mu.unlock();
  return rc;
}

int gc_amd_integrate_batches_device(int nBatch, const GcAmdBatchDesc* batch,
                                    gc_amd_progress_fn progress, void* user, void* stream,
                                    double* tKernel) {
  if (tKernel) *tKernel = 0.0;
  unsigned qn = 0;
  if (check_batches(nBatch, batch, &qn) != 0) {
    set_error("gc_amd_integrate_batches_device: bad argument");
    return -1;
  }
  int rc = 0;
  DevCtx* cp = enter(stream, &rc);
  if (!cp) return rc;
  rc = multi_solve(*cp, nBatch, batch, qn, nullptr, nullptr, nullptr, cp->stream, tKernel,
                   progress, user);
  cp->mu.unlock();
  return rc;
}

int gc_amd_record_batches_device(int nBatch, const GcAmdBatchDesc* batch, double* recH,
                                 int* recPrev, int* recYslot, int* head, long cap, double* recY,
                                 long ycap, int every, double* recK, long* used, long* yused,
                                 gc_amd_progress_fn progress, void* user, void* stream,
                                 double* tKernel) {
  if (tKernel) *tKernel = 0.0;
  if (used) *used = 0;
  if (yused) *yused = 0;
  unsigned qn = 0;
  if (check_batches(nBatch, batch, &qn) != 0 || cap < 0 || cap > 0x7fffffffL || ycap < 0 ||
      ycap > 0x7fffffffL || every < 1 || every > 4096) {
    set_error("gc_amd_record_batches_device: bad argument");
    return -1;
  }
  int rc = 0;
  DevCtx* cp = enter(stream, &rc);
  if (!cp) return rc;
  GcRecord rec{recH, recPrev, recYslot, head, recY, recK, nullptr, nullptr,
               (unsigned)cap, (unsigned)ycap, every};
  rc = multi_solve(*cp, nBatch, batch, qn, &rec, used, yused, cp->stream, tKernel, progress,
                   user);
  cp->mu.unlock();
  return rc;
}

int gc_amd_adjoint_batches_device(int nBatch, const GcAmdAdjBatch* batch, const double* recH,
                                  const int* recPrev, const int* recYslot, const int* head,
                                  long cap, const double* recY, long ycap, int every,
                                  const double* recK, gc_amd_progress_fn progress, void* user,
                                  void* stream, double* tKernel) {
  if (tKernel) *tKernel = 0.0;
  unsigned long long tot = 0;
  bool bad = nBatch < 1 || nBatch > 4096 || !batch || cap < 0 || cap > 0x7fffffffL || ycap < 0 ||
             ycap > 0x7fffffffL || every < 1 || every > 64;
  for (int k = 0; !bad && k < nBatch; ++k) {
    bad = batch[k].nCell < 0;
    tot += bad ? 0u : (unsigned)batch[k].nCell;
  }
  if (bad || tot > 0x7fffffffull) {
    set_error("gc_amd_adjoint_batches_device: bad argument");
    return -1;
  }
  int rc = 0;
  DevCtx* cp = enter(stream, &rc);
  if (!cp) return rc;
  std::vector<GcAdjDesc> h((size_t)nBatch);
  for (int k = 0; k < nBatch; ++k) {
    const GcAmdAdjBatch& x = batch[k];
    h[k].a = GcAdjArgs{x.nCell, x.Tstart, x.Tend, x.C, x.RCONST, x.ICNTRL, x.Perm, recY, recK,
                       recH, recPrev, recYslot, head, (unsigned)cap, (unsigned)ycap, every, 0,
                       x.lamC, x.gC, x.gRCONST, x.IERR};
  }
  rc = adjoint_sweep(*cp, h, cp->stream, progress, user, tKernel);
  cp->mu.unlock();
  return rc;
}

int gc_amd_resident_create(int nCell, const int* ICNTRL, const double* RCNTRL,
                           const double* ATOL, const double* RTOL, GcAmdResident** out) {
  if (!out) return -1;
  *out = nullptr;
  if (nCell < 1 || !ICNTRL || !RCNTRL || !ATOL || !RTOL) {
    set_error("gc_amd_resident_create: bad argument");
    return -1;
  }
  int rc = 0;
  DevCtx* cp = enter(nullptr, &rc);
  if (!cp) return rc;
  DevCtx& c = *cp;
  GcAmdResident* r = new (std::nothrow) GcAmdResident;
  if (!r) {
    c.mu.unlock();
    set_error("gc_amd_resident_create: host memory");
    return -2;
  }
  int d = -1;
  (void)hipGetDevice(&d);
  r->dev = d;
  r->n = nCell;
  const MechDims md = mech_dims();
  const size_t n = (size_t)nCell, nspec = (size_t)(md.nvar + md.nfix);
  hipError_t e = hipSuccess;
  DevBuf* bufs[] = {&r->C, &r->RCONST, &r->ATOL, &r->RTOL, &r->ICNTRL, &r->RCNTRL, &r->Perm,
                    &r->ISTAT, &r->RSTAT, &r->IERR, &r->keys, &r->keys_out, &r->cells};
  const size_t bytes[] = {n * nspec * 8, n * md.nreact * 8, (size_t)md.nvar * 8,
                          (size_t)md.nvar * 8, n * 20 * 4, n * 20 * 8, n * 4, n * 8 * 4,
                          n * 3 * 8, n * 4, n * 4, n * 4, n * 4};
  for (int i = 0; i < 13 && e == hipSuccess; ++i) e = bufs[i]->ensure(bytes[i]);
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, r->sort_bytes, (int*)r->keys.p,
                                                     (int*)r->keys_out.p, (int*)r->cells.p,
                                                     (int*)r->Perm.p, nCell, 0, 32, c.stream);
  if (e == hipSuccess) e = r->sort_tmp.ensure(r->sort_bytes > 0 ? r->sort_bytes : 16);
  Piece pin[4] = {{r->ICNTRL.p, const_cast<int*>(ICNTRL), n * 20 * 4},
                  {r->RCNTRL.p, const_cast<double*>(RCNTRL), n * 20 * 8},
                  {r->ATOL.p, const_cast<double*>(ATOL), (size_t)md.nvar * 8},
                  {r->RTOL.p, const_cast<double*>(RTOL), (size_t)md.nvar * 8}};
  c.stage.init();
  if (e == hipSuccess) e = c.stage.to_device(d, pin, 4, c.stream, c.stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_iota1, dim3((nCell + 255) / 256), dim3(256), 0, c.stream, (int*)r->Perm.p,
                       nCell);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  c.mu.unlock();
  if (e != hipSuccess) {
    set_error("gc_amd_resident_create", e);
    for (DevBuf* b : bufs) b->release();
    r->sort_tmp.release();
    delete r;
    return -2;
  }
  *out = r;
  return 0;
}

int gc_amd_resident_step(GcAmdResident* r, double Tstart, double Tend, const double* C,
                         const double* RCONST, int warm, double* Cout, int* ISTAT,
                         double* RSTAT, int* IERR, int* PermUsed, double* tTotal,
                         double* tKernel) {
  if (tTotal) *tTotal = 0.0;
  if (tKernel) *tKernel = 0.0;
  if (!r || r->n < 1) {
    set_error("gc_amd_resident_step: bad handle");
    return -1;
  }
  if (r->steps == 0 && !C) {
    set_error("gc_amd_resident_step: the first step needs C");
    return -1;
  }
  if (r->steps == 0 && !RCONST) {
    set_error("gc_amd_resident_step: the first step needs RCONST");
    return -1;
  }
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != r->dev) {
    set_error("gc_amd_resident_step: the batch lives on another device");
    return -1;
  }
  const auto t0 = std::chrono::steady_clock::now();
  int rc = 0;
  DevCtx* cp = enter(nullptr, &rc);
  if (!cp) return rc;
  DevCtx& c = *cp;
  const MechDims md = mech_dims();
  const int nCell = r->n;
  const size_t n = (size_t)nCell, nspec = (size_t)(md.nvar + md.nfix);
  hipError_t e = hipSuccess;
  // the launch order and warm start from the step before, on the device
  if (r->steps > 0) {
    hipLaunchKernelGGL(k_carry_over, dim3((nCell + 255) / 256), dim3(256), 0, c.stream,
                       (const int*)r->ISTAT.p, (const double*)r->RSTAT.p, (double*)r->RCNTRL.p,
                       (int*)r->keys.p, nCell, warm);
    hipLaunchKernelGGL(k_iota1, dim3((nCell + 255) / 256), dim3(256), 0, c.stream,
                       (int*)r->cells.p, nCell);
    e = hipGetLastError();
    size_t tb = r->sort_bytes;
    if (e == hipSuccess)
      e = hipcub::DeviceRadixSort::SortPairsDescending(r->sort_tmp.p, tb, (int*)r->keys.p,
                                                       (int*)r->keys_out.p, (int*)r->cells.p,
                                                       (int*)r->Perm.p, nCell, 0, 32, c.stream);
  }
  Piece pin[2];
  int np = 0;
  if (C) pin[np++] = {r->C.p, const_cast<double*>(C), n * nspec * 8};
  if (RCONST) pin[np++] = {r->RCONST.p, const_cast<double*>(RCONST), n * md.nreact * 8};
  if (e == hipSuccess && np) e = c.stage.to_device(r->dev, pin, np, c.stream, c.stream);
  if (e != hipSuccess) {
    set_error("gc_amd_resident_step: staging", e);
    c.mu.unlock();
    return -3;
  }
  GcBatchArgs b{nCell, Tstart, Tend, (double*)r->C.p, (const double*)r->RCONST.p,
                (const double*)r->ATOL.p, (const double*)r->RTOL.p, (const int*)r->ICNTRL.p,
                (const double*)r->RCNTRL.p, (const int*)r->Perm.p, (int*)r->ISTAT.p,
                (double*)r->RSTAT.p, (int*)r->IERR.p};
  rc = device_solve(c, GC_AMD_KERNEL_PERSISTENT, 0, b, c.stream, tKernel);
  if (rc != 0) {
    c.mu.unlock();
    return rc;
  }
  r->steps += 1;
  Piece pout[5];
  np = 0;
  if (Cout) pout[np++] = {r->C.p, Cout, n * nspec * 8};
  if (ISTAT) pout[np++] = {r->ISTAT.p, ISTAT, n * 8 * 4};
  if (RSTAT) pout[np++] = {r->RSTAT.p, RSTAT, n * 3 * 8};
  if (IERR) pout[np++] = {r->IERR.p, IERR, n * 4};
  if (PermUsed) pout[np++] = {r->Perm.p, PermUsed, n * 4};
  e = np ? c.stage.to_host(r->dev, pout, np, c.stream) : hipStreamSynchronize(c.stream);
  c.mu.unlock();
  if (e != hipSuccess) {
    set_error("gc_amd_resident_step: copy back", e);
    return -3;
  }
  if (tTotal) *tTotal = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

void gc_amd_resident_destroy(GcAmdResident* r) {
  if (!r) return;
  int cur = -1;
  if (hipGetDevice(&cur) == hipSuccess && cur != r->dev) (void)hipSetDevice(r->dev);
  DevBuf* bufs[] = {&r->C, &r->RCONST, &r->ATOL, &r->RTOL, &r->ICNTRL, &r->RCNTRL, &r->Perm,
                    &r->ISTAT, &r->RSTAT, &r->IERR, &r->keys, &r->keys_out, &r->cells,
                    &r->sort_tmp};
  for (DevBuf* b : bufs) b->release();
  if (cur >= 0 && cur != r->dev) (void)hipSetDevice(cur);
  delete r;
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

int gc_amd_set_gate(int mode, int cells, unsigned long long spin) {
  (void)current_gate();  // read the environment first so explicit values win
  if (mode < 0 || mode > 2 || cells < 0 || spin == 0ull) return -1;
  std::lock_guard<std::mutex> lk(g_gate_mu);
  g_gate.mode = mode;
  g_gate.cells = cells;
  g_gate.spin = spin;
  return 0;
}

void gc_amd_get_gate(int* mode, int* cells, unsigned long long* spin) {
  const GateConfig g = current_gate();
  *mode = g.mode;
  *cells = g.cells;
  *spin = g.spin;
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
