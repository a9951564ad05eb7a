// hybrid_threads_check.cpp -- host-only check of gc_amd_hybrid_step's
// multi-device path (csrc/gc_hybrid.cpp), built by tests/test_hybrid.py with
// g++ and the address/undefined or thread sanitizer.
//
// The device solve is replaced by a stand-in with the same signature: it
// records which slice each device received, waits inside the call until
// every device call of the step has entered (a rendezvous with a bounded
// wait: the check plans the split first, so it knows how many device calls
// to expect), and writes per-cell outputs that name the device and the slot,
// so the check can prove that the device calls were in progress at the same
// time and that every slot of the head was solved exactly once by the device
// whose slice holds it, at the right offsets of every array.  Device `bad`
// fails the way the library does for a device that cannot be selected
// (IERR -99).
//
// usage: hybrid_threads_check <nCell> <ndev> <bad device or -1>
// Prints one line per check and exits 0 when all pass.
#include <stdio.h>
#include <stdlib.h>

#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>
#include <vector>

#include "gcchemgpu.h"

namespace {
const int kNvar = 5, kNfix = 2, kNreact = 7, kNspec = kNvar + kNfix;
int g_bad = -1;
std::atomic<int> g_active{0}, g_max_active{0};
std::mutex g_mu;
int g_expected = 0;                   // device calls the step's plan makes
std::atomic<int> g_arrived{0};        // device calls that have entered
std::atomic<bool> g_met{true};        // every call saw all of them inside
struct Call {
  int dev, n;
  long c_off, r_off;
};
std::vector<Call> g_calls;
const double* g_C0 = nullptr;
const double* g_R0 = nullptr;
}  // namespace

extern "C" void gc_amd_mech_info(int* nvar, int* nfix, int* nreact, int* nnz) {
  *nvar = kNvar;
  *nfix = kNfix;
  *nreact = kNreact;
  *nnz = 0;
}

extern "C" void gc_gpu_integrate_batch_dev(int dev, int nCell, double, double, double* C,
                                           const double* RCONST, const double*, const double*,
                                           const int* ICNTRL, const double* RCNTRL,
                                           const int* Perm, int* ISTAT, double* RSTAT, int* IERR,
                                           double* tTotal, double* tKernel) {
  const int now = ++g_active;
  int m = g_max_active.load();
  while (now > m && !g_max_active.compare_exchange_weak(m, now)) {
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_calls.push_back({dev, nCell, (long)(C - g_C0), (long)(RCONST - g_R0)});
  }
  // stay inside until every device call of the step has entered, at most 20 s
  ++g_arrived;
  const auto deadline = std::chrono::steady_clock::now() + std::chrono::seconds(20);
  while (g_arrived.load() < g_expected) {
    if (std::chrono::steady_clock::now() > deadline) {
      g_met = false;
      break;
    }
    std::this_thread::yield();
  }
  for (int i = 0; i < nCell; ++i) {
    const int c = Perm[i] - 1;  // identity on the slice
    if (dev == g_bad) {
      IERR[c] = -99;
      continue;
    }
    // the slot this column holds, read back from the inputs the caller set
    const double slot = C[(long)c * kNspec];
    C[(long)c * kNspec + 1] = 1000.0 * (dev + 1) + slot;
    ISTAT[(long)c * 8 + 2] = dev + 1;
    RSTAT[(long)c * 3] = RCONST[(long)c * kNreact] + ICNTRL[(long)c * 20] + RCNTRL[(long)c * 20];
    IERR[c] = 1;
  }
  *tTotal = *tKernel = 1e-4 * nCell;
  --g_active;
}

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  const int n = atoi(argv[1]), ndev = atoi(argv[2]);
  g_bad = atoi(argv[3]);
  std::vector<double> C((size_t)n * kNspec), RCONST((size_t)n * kNreact), RCNTRL((size_t)n * 20),
      RSTAT((size_t)n * 3, -1.0), ATOL(kNvar, 1.0), RTOL(kNvar, 1.0);
  std::vector<int> ICNTRL((size_t)n * 20), ISTAT((size_t)n * 8, -1), IERR(n, 0), nsteps(n);
  for (int s = 0; s < n; ++s) {
    C[(size_t)s * kNspec] = s + 1;  // every column knows its slot
    RCONST[(size_t)s * kNreact] = 2.0 * (s + 1);
    ICNTRL[(size_t)s * 20] = 3;
    RCNTRL[(size_t)s * 20] = 0.5;
    nsteps[s] = 1 + (n - s) % 17;
  }
  g_C0 = C.data();
  g_R0 = RCONST.data();
  GcAmdHybrid st = {};
  st.work_frac = 0.6;
  st.ndev = ndev;
  st.rate_gpu = st.rate_cpu = -1.0;
  for (int d = 0; d < GC_AMD_MAX_DEV; ++d) st.dev_rate[d] = -1.0;
  {
    // the step's own plan (same inputs): how many device calls to expect
    double frac = st.work_frac;
    std::vector<long long> work((size_t)n + 1);
    int ng = 0, nd = 0, lo[GC_AMD_MAX_DEV], hi[GC_AMD_MAX_DEV];
    if (ndev >= 1 && ndev <= GC_AMD_MAX_DEV &&
        gc_amd_hybrid_plan(n, nsteps.data(), st.fixed_frac, &frac, st.rate_gpu, st.rate_cpu,
                           ndev, st.dev_rate, work.data(), &ng, &nd, lo, hi) == 0)
      for (int d = 0; d < nd; ++d) g_expected += hi[d] >= lo[d];
  }
  struct Tail {
    int calls = 0, first = 0, last = 0;
  } tail;
  auto cb = [](int first, int last, void* user) {
    auto* t = static_cast<Tail*>(user);
    ++t->calls;
    t->first = first;
    t->last = last;
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
  };
  const int rc = gc_amd_hybrid_step(n, 0.0, 600.0, C.data(), RCONST.data(), ATOL.data(),
                                    RTOL.data(), ICNTRL.data(), RCNTRL.data(), nsteps.data(),
                                    ISTAT.data(), RSTAT.data(), IERR.data(), &st, cb, &tail);
  int fails = 0;
  const int ng = st.n_gpu;
  printf("plan: nGpu %d of %d, %d devices used, %zu device calls, %d at once, rc %d\n", ng, n,
         st.n_dev_use, g_calls.size(), g_max_active.load(), rc);
  // each device with a non-empty slice was called once, on its slice's base
  // (work-balanced slicing can leave a device nothing when one slot outweighs
  // its share; such a slice is skipped, as the reference skips it)
  int nonempty = 0;
  for (int d = 0; d < st.n_dev_use; ++d) nonempty += st.dev_hi[d] >= st.dev_lo[d];
  fails += (int)g_calls.size() != nonempty;
  for (const Call& c : g_calls) {
    const int lo = st.dev_lo[c.dev], hi = st.dev_hi[c.dev];
    fails += c.n != hi - lo + 1 || c.c_off != (long)(lo - 1) * kNspec ||
             c.r_off != (long)(lo - 1) * kNreact;
  }
  // every head slot solved once by its own device; tail untouched
  int expect_fail = 0;
  for (int d = 0; d < st.n_dev_use; ++d) {
    for (int s = st.dev_lo[d]; s <= st.dev_hi[d]; ++s) {
      const size_t c = (size_t)(s - 1);
      if (d == g_bad) {
        fails += IERR[c] != -99;
        ++expect_fail;
      } else {
        fails += IERR[c] != 1 || ISTAT[c * 8 + 2] != d + 1 ||
                 C[c * kNspec + 1] != 1000.0 * (d + 1) + s ||
                 RSTAT[c * 3] != 2.0 * s + 3 + 0.5;
      }
    }
  }
  for (int s = ng + 1; s <= n; ++s) fails += IERR[s - 1] != 0 || ISTAT[(size_t)(s - 1) * 8 + 2] != -1;
  fails += rc != expect_fail;
  fails += tail.calls != (ng < n) || (ng < n && (tail.first != ng + 1 || tail.last != n));
  // the devices overlapped: every call waited for all of them
  fails += nonempty != g_expected || !g_met.load() || g_max_active.load() != nonempty;
  fails += (ng > 0 && !(st.rate_gpu > 0.0)) || (ng < n && !(st.rate_cpu > 0.0));
  printf("slices:");
  for (int d = 0; d < st.n_dev_use; ++d) printf(" [%d,%d]", st.dev_lo[d], st.dev_hi[d]);
  printf("\ntail: %d call(s), slots %d-%d; failed device slots %d; rendezvous of %d %s\n",
         tail.calls, tail.first, tail.last, expect_fail, g_expected, g_met.load() ? "met" : "timed out");
  printf(fails ? "FAIL\n" : "OK\n");
  return fails ? 1 : 0;
}
