// gc_membench.hip -- memory-system ceiling for the Rodas3 workspace access shape.
//
// The batched Rodas3 kernels read and write a private workspace block per
// 64-lane wave (element e of lane l at byte e*512 + l*8), in an order set by
// the mechanism's sparsity: every access is one coalesced 512-byte wave
// access at an effectively random element of a multi-megabyte block.  This
// program measures what the memory system delivers for exactly that shape on
// the device it runs on:
//   stream   plain streaming read of the whole buffer (all waves, grid-stride)
//   gather   random 512-byte wave reads, `loads_in_flight` (4, 8, 16 or 32)
//            in flight per wave, with one 512-byte wave write per
//            `loads_per_store` reads; each batch of reads (and the writes
//            after it) falls in a window of `window` consecutive elements at
//            a random place in the block (window = block: fully random), so
//            small windows measure what address locality within a batch buys
//   chain    latency of one 512-byte wave read whose address depends on the
//            previous read (one wave per CU, each in its own block): the
//            round trip a chain of dependent operations pays per link
// for a given block size (doubles per lane) and waves per CU; the working set
// of the gather is CUs * waves_per_cu * block_doubles * 512 bytes, so small
// blocks measure the shape out of the on-chip caches.  It prints one JSON
// line.  Usage:
//   gc_membench [block_doubles waves_per_cu loads_per_store loads_in_flight window]
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>

namespace {

template <int kLoads>
__global__ void __launch_bounds__(256) k_gather(double* buf, long block_elems, int iters,
                                               int loads_per_store, unsigned window,
                                               double* out) {
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  double* base = buf + wave * block_elems * 64;
  unsigned s = 12345u + (unsigned)wave * 7919u;
  const unsigned nel = (unsigned)block_elems;
  const unsigned places = nel - window + 1u;  // 1 <= window <= nel (checked on the host)
  unsigned w0 = 0;
  double acc = 0.0;
  int since = 0;
  for (int it = 0; it < iters; ++it) {
    double v[kLoads];
    s = s * 1664525u + 1013904223u;
    w0 = __builtin_amdgcn_readfirstlane((s >> 8) % places);
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
      s = s * 1664525u + 1013904223u;
      const unsigned e = __builtin_amdgcn_readfirstlane(w0 + (s >> 8) % window);  // < nel
      v[q] = base[(long)e * 64 + lane];
    }
#pragma unroll
    for (int q = 0; q < kLoads; ++q) acc += v[q];
    since += kLoads;
    while (loads_per_store > 0 && since >= loads_per_store) {
      since -= loads_per_store;
      s = s * 1664525u + 1013904223u;
      const unsigned e = __builtin_amdgcn_readfirstlane(w0 + (s >> 8) % window);
      base[(long)e * 64 + lane] = acc;
    }
  }
  out[wave * 64 + lane] = acc;
}

// One wave per block: `iters` dependent reads at random elements of the
// wave's block (the buffer is zero, so the value only feeds the address).
__global__ void __launch_bounds__(64) k_chain(const double* buf, long block_elems, int iters,
                                             double* out) {
  const long wave = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const double* base = buf + wave * block_elems * 64;
  unsigned s = 12345u + (unsigned)wave * 7919u;
  const unsigned nel = (unsigned)block_elems;
  double v = 0.0;
  for (int it = 0; it < iters; ++it) {
    s = s * 1664525u + 1013904223u;
    const unsigned e = __builtin_amdgcn_readfirstlane(((s >> 8) + (v != 0.0 ? 1u : 0u)) % nel);
    v = base[(long)e * 64 + lane];
  }
  out[wave * 64 + lane] = v;
}

__global__ void __launch_bounds__(256) k_stream(const double* buf, long n, double* out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (; i < n; i += stride) acc += buf[i];
  out[(long)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

bool ok(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fprintf(stderr, "gc_membench: %s: %s\n", what, hipGetErrorString(e));
    return false;
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  long block = argc > 1 ? atol(argv[1]) : 12799;
  int wpc = argc > 2 ? atoi(argv[2]) : 8;
  int lps = argc > 3 ? atoi(argv[3]) : 5;
  int kl = argc > 4 ? atoi(argv[4]) : 16;
  long window = argc > 5 ? atol(argv[5]) : block;
  if (block < 64 || block > 65536 || wpc < 4 || wpc > 32 || wpc % 4 || lps < 0 ||
      (kl != 4 && kl != 8 && kl != 16 && kl != 32) || window < 1 || window > block) {
    fprintf(stderr, "gc_membench: bad arguments\n");
    return 2;
  }
  hipDeviceProp_t prop;
  if (!ok(hipGetDeviceProperties(&prop, 0), "device")) return 1;
  const int cus = prop.multiProcessorCount;
  const long waves = (long)cus * wpc;
  const size_t n = (size_t)waves * block * 64;
  const int stream_blocks = cus * 8, stream_threads = 256;
  const size_t nout = (size_t)(waves * 64 > (long)stream_blocks * stream_threads
                                   ? waves * 64 : (long)stream_blocks * stream_threads);
  double *buf = nullptr, *out = nullptr;
  if (!ok(hipMalloc(&buf, n * sizeof(double)), "hipMalloc")) return 1;
  if (!ok(hipMalloc(&out, nout * sizeof(double)), "hipMalloc")) return 1;
  if (!ok(hipMemset(buf, 0, n * sizeof(double)), "hipMemset")) return 1;
  hipEvent_t a, b;
  if (!ok(hipEventCreate(&a), "event") || !ok(hipEventCreate(&b), "event")) return 1;
  float ms = 0.f;

  hipLaunchKernelGGL(k_stream, dim3(stream_blocks), dim3(stream_threads), 0, 0, buf, (long)n, out);
  hipEventRecord(a, 0);
  hipLaunchKernelGGL(k_stream, dim3(stream_blocks), dim3(stream_threads), 0, 0, buf, (long)n, out);
  hipEventRecord(b, 0);
  if (!ok(hipEventSynchronize(b), "stream")) return 1;
  hipEventElapsedTime(&ms, a, b);
  const double stream_tbs = n * 8.0 / (ms * 1e-3) / 1e12;

  const int iters = 8000 / kl;  // 8,000 reads per wave whatever the depth
  const int grid = (int)(waves * 64 / 256);
  auto gather = [&](int it) {
    switch (kl) {
      case 4: hipLaunchKernelGGL(k_gather<4>, dim3(grid), dim3(256), 0, 0, buf, block, it, lps, (unsigned)window, out); break;
      case 8: hipLaunchKernelGGL(k_gather<8>, dim3(grid), dim3(256), 0, 0, buf, block, it, lps, (unsigned)window, out); break;
      case 16: hipLaunchKernelGGL(k_gather<16>, dim3(grid), dim3(256), 0, 0, buf, block, it, lps, (unsigned)window, out); break;
      default: hipLaunchKernelGGL(k_gather<32>, dim3(grid), dim3(256), 0, 0, buf, block, it, lps, (unsigned)window, out); break;
    }
  };
  gather(1);
  hipEventRecord(a, 0);
  gather(iters);
  hipEventRecord(b, 0);
  if (!ok(hipEventSynchronize(b), "gather")) return 1;
  hipEventElapsedTime(&ms, a, b);
  const double loads = (double)waves * iters * kl;
  const double stores = lps > 0 ? (double)waves * ((double)iters * kl / lps) : 0.0;
  const double gather_tbs = (loads + stores) * 512.0 / (ms * 1e-3) / 1e12;

  // the chain: one wave per CU, blocks 0..cus-1 of the buffer (cus <= waves)
  const int links = 20000;
  hipLaunchKernelGGL(k_chain, dim3(cus), dim3(64), 0, 0, buf, block, 100, out);
  hipEventRecord(a, 0);
  hipLaunchKernelGGL(k_chain, dim3(cus), dim3(64), 0, 0, buf, block, links, out);
  hipEventRecord(b, 0);
  if (!ok(hipEventSynchronize(b), "chain")) return 1;
  hipEventElapsedTime(&ms, a, b);
  const double chain_ns = ms * 1e6 / links;
  printf("{\"device\": \"%s\", \"cus\": %d, \"block_doubles\": %ld, \"waves_per_cu\": %d, "
         "\"loads_per_store\": %d, \"loads_in_flight\": %d, \"window\": %ld, \"buffer_gb\": %.2f, "
         "\"stream_read_tbs\": %.3f, "
         "\"gather_tbs\": %.3f, \"chain_ns\": %.1f, \"chain_set_mb\": %.1f}\n",
         prop.gcnArchName, cus, block, wpc, lps, kl, window, n * 8.0 / 1e9, stream_tbs, gather_tbs, chain_ns,
         cus * block * 512.0 / 1e6);
  hipFree(buf);
  hipFree(out);
  return 0;
}
