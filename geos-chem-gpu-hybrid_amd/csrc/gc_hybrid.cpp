// gc_hybrid.cpp -- cost order and the hybrid CPU/GPU split of one batch
// (declarations: include/gcchemgpu.h).
//
// Host-only C++.  The reference keeps this logic in its caller
// (src/geos-chem/GeosCore/fullchem_mod.F90, batch mode 3): a stable counting
// sort of the cells by the previous chemistry step's internal step count
// (:905-981), a work model of prefix sums over the cost-ordered slots, the
// GPU's share of that work from measured rates, work-balanced contiguous
// device slices (:997-1061), the device launches issued while the CPU solves
// the tail, and exponential moving averages of the measured rates
// (:2672-2700).  Here the same steps are library entry points, so a caller of
// any language gets the split and the concurrent launch with one call.
//
// No OpenMP: the library must not bring a second OpenMP runtime into the
// caller's process, so device launches run on std::threads and the CPU tail
// runs on the calling thread (inside it, the caller's own OpenMP team).
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

#include "gcchemgpu.h"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// EMA of fullchem_mod.F90:2677-2681: the first measurement is taken as is,
// later ones are averaged half and half with the previous rate.
double ema(double old_rate, double r_new) {
  return old_rate > 0.0 ? 0.5 * old_rate + 0.5 * r_new : r_new;
}

}  // namespace

extern "C" {

int gc_amd_cost_slots(int nCell, const int* prevNsteps, int sort, int* slotOf, int* cellAt) {
  if (nCell < 0 || (nCell > 0 && (!prevNsteps || !slotOf || !cellAt))) return -1;
  int sMax = 0;
  for (int c = 0; c < nCell; ++c) {
    if (prevNsteps[c] < 0) return -1;
    if (sort) sMax = std::max(sMax, prevNsteps[c]);
  }
  // pass 1: cells per key (key 0 = most expensive), then the first slot of
  // every key; pass 2: slots in the caller's order within a key (stable)
  std::vector<int> first;
  try {
    first.assign(static_cast<size_t>(sMax) + 2, 0);
  } catch (...) {
    return -2;
  }
  for (int c = 0; c < nCell; ++c) ++first[sort ? sMax - prevNsteps[c] : 0];
  int pos = 1;
  for (int k = 0; k <= sMax; ++k) {
    const int n = first[k];
    first[k] = pos;
    pos += n;
  }
  for (int c = 0; c < nCell; ++c) {
    const int s = first[sort ? sMax - prevNsteps[c] : 0]++;
    slotOf[c] = s;
    cellAt[s - 1] = c + 1;
  }
  return 0;
}

int gc_amd_hybrid_plan(int nCell, const int* slotNsteps, int fixedFrac, double* workFrac,
                       double rateGpu, double rateCpu, int nDev, const double* devRate,
                       long long* work, int* nGpu, int* nDevUse, int* devLo, int* devHi) {
  if (nCell < 0 || nDev < 1 || nDev > GC_AMD_MAX_DEV || !workFrac || !work || !nGpu ||
      !nDevUse || !devLo || !devHi || (nCell > 0 && !slotNsteps))
    return -1;
  // work model: cost(slot) = max(previous internal step count, 1)
  work[0] = 0;
  for (int ic = 1; ic <= nCell; ++ic)
    work[ic] = work[ic - 1] + std::max(slotNsteps[ic - 1], 1);
  const double wTot = static_cast<double>(work[nCell]);
  double frac = *workFrac;
  if (!fixedFrac && rateGpu > 0.0 && rateCpu > 0.0) {
    frac = rateGpu / (rateGpu + rateCpu);
    frac = std::min(std::max(frac, 0.01), 0.99);
  }
  frac = std::min(std::max(frac, 0.0), 1.0);
  *workFrac = frac;
  // the head: the longest prefix whose work does not exceed the GPU's share
  double wTarget = frac * wTot;
  int ng = 0;
  for (int ic = 1; ic <= nCell; ++ic) {
    if (static_cast<double>(work[ic]) > wTarget) break;
    ng = ic;
  }
  *nGpu = ng;
  int nd = nDev;
  if (ng < nd) nd = std::max(ng, 1);
  *nDevUse = nd;
  for (int d = 0; d < nDev; ++d) {
    devLo[d] = 1;
    devHi[d] = 0;
  }
  // device weights: the measured rates once every used device has one,
  // equal shares before that; contiguous slices balanced on work
  double w[GC_AMD_MAX_DEV];
  bool all_rated = devRate != nullptr;
  for (int d = 0; d < nd && all_rated; ++d) all_rated = devRate[d] > 0.0;
  double wsum = 0.0;
  for (int d = 0; d < nd; ++d) {
    w[d] = all_rated ? devRate[d] : 1.0;
    wsum += w[d];
  }
  int idx = 0;
  double cum = 0.0;
  for (int d = 0; d < nd; ++d) {
    devLo[d] = idx + 1;
    if (d == nd - 1) {
      devHi[d] = ng;
    } else {
      cum += w[d];
      wTarget = static_cast<double>(work[ng]) * cum / wsum;
      for (int ic = idx + 1; ic <= ng; ++ic) {
        if (static_cast<double>(work[ic]) > wTarget) break;
        idx = ic;
      }
      devHi[d] = idx;
    }
    idx = devHi[d];
  }
  return 0;
}

void gc_amd_hybrid_update(int nCell, int nGpu, const long long* work, int nDevUse,
                          const int* devLo, const int* devHi, const double* devTw, double gpuTw,
                          double tCpu, double* rateGpu, double* rateCpu, double* devRate) {
  if (nCell < 0 || nGpu < 0 || nGpu > nCell || !work) return;
  for (int d = 0; d < nDevUse && devRate && devLo && devHi && devTw; ++d) {
    if (devHi[d] >= devLo[d] && devTw[d] > 0.0) {
      const double r = static_cast<double>(work[devHi[d]] - work[devLo[d] - 1]) / devTw[d];
      devRate[d] = ema(devRate[d], r);
    }
  }
  if (nGpu > 0 && gpuTw > 0.0 && rateGpu)
    *rateGpu = ema(*rateGpu, static_cast<double>(work[nGpu]) / gpuTw);
  if (nCell - nGpu > 0 && tCpu > 0.0 && rateCpu)
    *rateCpu = ema(*rateCpu, (static_cast<double>(work[nCell]) -
                              static_cast<double>(work[nGpu])) / tCpu);
}

int gc_amd_hybrid_step(int nCell, double Tstart, double Tend, double* C, const double* RCONST,
                       const double* ATOL, const double* RTOL, const int* ICNTRL,
                       const double* RCNTRL, const int* slotNsteps, int* ISTAT, double* RSTAT,
                       int* IERR, GcAmdHybrid* st, gc_amd_cpu_tail_fn cpuTail, void* user) {
  if (!st || nCell < 0) return -1;
  int nvar = 0, nfix = 0, nreact = 0, nnz = 0;
  gc_amd_mech_info(&nvar, &nfix, &nreact, &nnz);
  const long nspec = nvar + nfix;
  std::vector<long long> work;
  std::vector<int> perm;
  try {
    work.resize(static_cast<size_t>(nCell) + 1);
  } catch (...) {
    return -2;
  }
  if (gc_amd_hybrid_plan(nCell, slotNsteps, st->fixed_frac, &st->work_frac, st->rate_gpu,
                         st->rate_cpu, st->ndev, st->dev_rate, work.data(), &st->n_gpu,
                         &st->n_dev_use, st->dev_lo, st->dev_hi) != 0)
    return -1;
  const int ng = st->n_gpu, nd = st->n_dev_use;
  if (ng > 0 && (!C || !RCONST || !ATOL || !RTOL || !ICNTRL || !RCNTRL || !ISTAT || !RSTAT ||
                 !IERR))
    return -1;
  if (ng < nCell && !cpuTail) return -1;
  // Every device slice is solved in launch order = slot order: Perm is the
  // identity on the slice's base pointers (fullchem_mod.F90:1253-1257).
  int maxSlice = 0;
  for (int d = 0; d < nd; ++d) maxSlice = std::max(maxSlice, st->dev_hi[d] - st->dev_lo[d] + 1);
  try {
    perm.resize(static_cast<size_t>(std::max(maxSlice, 1)));
  } catch (...) {
    return -2;
  }
  for (int i = 0; i < maxSlice; ++i) perm[i] = i + 1;
  double devTw[GC_AMD_MAX_DEV] = {0.0}, devTa[GC_AMD_MAX_DEV] = {0.0},
         devTk[GC_AMD_MAX_DEV] = {0.0};
  auto launch = [&](int d) {
    const int n = st->dev_hi[d] - st->dev_lo[d] + 1;
    const long o = st->dev_lo[d] - 1;
    const double a = now_s();
    gc_gpu_integrate_batch_dev(d, n, Tstart, Tend, C + o * nspec, RCONST + o * nreact, ATOL,
                               RTOL, ICNTRL + o * 20, RCNTRL + o * 20, perm.data(),
                               ISTAT + o * 8, RSTAT + o * 3, IERR + o, &devTa[d], &devTk[d]);
    devTw[d] = now_s() - a;
  };
  const double t0 = now_s();
  std::vector<std::thread> launches;
  try {
    launches.reserve(nd);
  } catch (...) {
  }
  for (int d = 0; d < nd; ++d) {
    if (st->dev_hi[d] < st->dev_lo[d]) continue;
    try {
      launches.emplace_back(launch, d);
    } catch (...) {  // no thread available: launch from the calling thread
      launch(d);
    }
  }
  // the CPU's share, concurrently, on the calling thread
  double tCpu = 0.0;
  if (ng < nCell) {
    const double a = now_s();
    cpuTail(ng + 1, nCell, user);
    tCpu = now_s() - a;
  }
  for (auto& t : launches) t.join();
  double gpuTw = 0.0, gpuTk = 0.0;
  for (int d = 0; d < nd; ++d) {
    gpuTw = std::max(gpuTw, devTw[d]);
    gpuTk += devTk[d];
  }
  st->t_gpu = gpuTw;
  st->t_kernel = gpuTk;
  st->t_cpu = tCpu;
  st->t_region = now_s() - t0;
  gc_amd_hybrid_update(nCell, ng, work.data(), nd, st->dev_lo, st->dev_hi, devTw, gpuTw, tCpu,
                       &st->rate_gpu, &st->rate_cpu, st->dev_rate);
  int nFail = 0;
  for (int ic = 0; ic < ng; ++ic) nFail += IERR[ic] < 0;
  return nFail;
}

}  // extern "C"
