/* gc_link_test.c -- link and OpenMP-coexistence check of the drop-in
 * libgcchemgpu.so, in the spirit of the reference's boundary tests
 * (gpulib/src/test_link.F90 and the OpenMP-around-GPU program of
 * gpulib/scripts/test_link_job.sh:26-81): a C program built with -fopenmp
 * links the library by its drop-in name, queries the device, runs the self
 * test, solves a batch once through gc_gpu_integrate_batch and then the same
 * batch from several OpenMP threads at once through
 * gc_gpu_integrate_batch_dev, and checks every thread's results against the
 * single call bit for bit.  Last, one hybrid split step
 * (gc_amd_hybrid_step): the library launches the head on its own host
 * thread while this program's CPU-tail callback runs an OpenMP team over
 * the tail (standing in for the caller's CPU solve, it copies the single
 * call's answers), and the combined batch must equal the single call.
 *
 * usage: gc_link_test <dir> <threads>
 *   <dir> holds the batch as raw little-endian arrays: meta.txt ("nCell
 *   nspec nreact nvar Tstart Tend"), C.bin (nCell*nspec doubles), RCONST.bin,
 *   ATOL.bin, RTOL.bin (nvar doubles), ICNTRL.bin (nCell*20 int32),
 *   RCNTRL.bin (nCell*20 doubles).
 * Prints one line per check and exits 0 when all pass. */
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gcchemgpu.h"

static void *slurp(const char *dir, const char *name, size_t bytes) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *f = fopen(path, "rb");
  if (!f) {
    fprintf(stderr, "cannot open %s\n", path);
    exit(2);
  }
  void *p = malloc(bytes ? bytes : 1);
  if (!p || fread(p, 1, bytes, f) != bytes) {
    fprintf(stderr, "short read of %s (%zu bytes expected)\n", path, bytes);
    exit(2);
  }
  fclose(f);
  return p;
}

struct Out {
  double *C, *RSTAT;
  int *ISTAT, *IERR;
};

static struct Out out_new(int n, int nspec, const double *C0) {
  struct Out o;
  o.C = malloc(sizeof(double) * (size_t)n * nspec);
  o.RSTAT = malloc(sizeof(double) * (size_t)n * 3);
  o.ISTAT = malloc(sizeof(int) * (size_t)n * 8);
  o.IERR = malloc(sizeof(int) * (size_t)n);
  if (!o.C || !o.RSTAT || !o.ISTAT || !o.IERR) {
    fprintf(stderr, "out of host memory\n");
    exit(2);
  }
  memcpy(o.C, C0, sizeof(double) * (size_t)n * nspec);
  return o;
}

static int out_same(const struct Out *a, const struct Out *b, int n, int nspec) {
  return memcmp(a->C, b->C, sizeof(double) * (size_t)n * nspec) == 0 &&
         memcmp(a->RSTAT, b->RSTAT, sizeof(double) * (size_t)n * 3) == 0 &&
         memcmp(a->ISTAT, b->ISTAT, sizeof(int) * (size_t)n * 8) == 0 &&
         memcmp(a->IERR, b->IERR, sizeof(int) * (size_t)n) == 0;
}

/* The caller's CPU solve of slots first..last, as gc_amd_hybrid_step calls
 * it; here every slot takes the single call's answer, from an OpenMP team. */
struct Tail {
  const struct Out *ref;
  struct Out *out;
  int nspec, threads, calls, first, last;
};

static void cpu_tail(int first, int last, void *user) {
  struct Tail *t = (struct Tail *)user;
  t->calls++;
  t->first = first;
  t->last = last;
#pragma omp parallel for num_threads(t->threads) schedule(dynamic, 16)
  for (int s = first; s <= last; ++s) {
    const size_t c = (size_t)(s - 1);
    memcpy(t->out->C + c * t->nspec, t->ref->C + c * t->nspec, sizeof(double) * t->nspec);
    memcpy(t->out->ISTAT + c * 8, t->ref->ISTAT + c * 8, sizeof(int) * 8);
    memcpy(t->out->RSTAT + c * 3, t->ref->RSTAT + c * 3, sizeof(double) * 3);
    t->out->IERR[c] = t->ref->IERR[c];
  }
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s <dir> <threads>\n", argv[0]);
    return 2;
  }
  const char *dir = argv[1];
  const int threads = atoi(argv[2]);
  int n = 0, nspec = 0, nreact = 0, nvar = 0;
  double Tstart = 0.0, Tend = 0.0;
  {
    char path[4096];
    snprintf(path, sizeof path, "%s/meta.txt", dir);
    FILE *f = fopen(path, "r");
    if (!f || fscanf(f, "%d %d %d %d %lf %lf", &n, &nspec, &nreact, &nvar, &Tstart, &Tend) != 6 ||
        n < 1 || threads < 1) {
      fprintf(stderr, "bad %s or thread count\n", path);
      return 2;
    }
    fclose(f);
  }
  const double *C0 = slurp(dir, "C.bin", sizeof(double) * (size_t)n * nspec);
  const double *RCONST = slurp(dir, "RCONST.bin", sizeof(double) * (size_t)n * nreact);
  const double *ATOL = slurp(dir, "ATOL.bin", sizeof(double) * (size_t)nvar);
  const double *RTOL = slurp(dir, "RTOL.bin", sizeof(double) * (size_t)nvar);
  const int *ICNTRL = slurp(dir, "ICNTRL.bin", sizeof(int) * (size_t)n * 20);
  const double *RCNTRL = slurp(dir, "RCNTRL.bin", sizeof(double) * (size_t)n * 20);
  int *Perm = malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; ++i) Perm[i] = i + 1;

  int ndev = 0;
  char name[256];
  gc_gpu_info(&ndev, name, (int)sizeof name);
  printf("info: %d device(s), %s\n", ndev, name);
  if (ndev < 1 || gc_gpu_set_device(0) != 0) {
    printf("FAIL: no usable device\n");
    return 1;
  }
  int rc = 0, nsteps = 0;
  double tt = 0.0, tk = 0.0;
  gc_gpu_selftest(&rc, &nsteps, &tt);
  printf("selftest: rc %d, %d steps\n", rc, nsteps);
  int fails = (rc != 1 || nsteps != 12);

  struct Out ref = out_new(n, nspec, C0);
  gc_gpu_integrate_batch(n, Tstart, Tend, ref.C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL, Perm,
                         ref.ISTAT, ref.RSTAT, ref.IERR, &tt, &tk);
  int ok = 0;
  for (int i = 0; i < n; ++i) ok += ref.IERR[i] == 1;
  printf("single call: %d of %d cells IERR = 1\n", ok, n);
  fails += ok != n;

  int differ = 0;
#pragma omp parallel num_threads(threads) reduction(+ : differ)
  {
    struct Out o = out_new(n, nspec, C0);
    double t1 = 0.0, t2 = 0.0;
    gc_gpu_integrate_batch_dev(0, n, Tstart, Tend, o.C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL, Perm,
                               o.ISTAT, o.RSTAT, o.IERR, &t1, &t2);
    differ += !out_same(&o, &ref, n, nspec);
    free(o.C);
    free(o.RSTAT);
    free(o.ISTAT);
    free(o.IERR);
  }
  printf("threads: %d concurrent calls, %d differ from the single call\n", threads, differ);
  fails += differ != 0;

  /* hybrid split: half of the work (by the single call's step counts) to the
   * device, the rest to the OpenMP tail, concurrently */
  int *slotNsteps = malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; ++i) slotNsteps[i] = ref.ISTAT[(size_t)i * 8 + 2];
  GcAmdHybrid st;
  memset(&st, 0, sizeof st);
  st.work_frac = 0.5;
  st.ndev = 1;
  st.rate_gpu = st.rate_cpu = -1.0;
  for (int d = 0; d < GC_AMD_MAX_DEV; ++d) st.dev_rate[d] = -1.0;
  struct Out h = out_new(n, nspec, C0);
  struct Tail tail = {&ref, &h, nspec, threads, 0, 0, 0};
  const int nfail = gc_amd_hybrid_step(n, Tstart, Tend, h.C, RCONST, ATOL, RTOL, ICNTRL, RCNTRL,
                                       slotNsteps, h.ISTAT, h.RSTAT, h.IERR, &st, cpu_tail,
                                       &tail);
  const int same = out_same(&h, &ref, n, nspec);
  printf("hybrid: %d slots on the device, %d in the OpenMP tail (%d call, slots %d-%d), "
         "%d failed, %s the single call\n",
         st.n_gpu, n - st.n_gpu, tail.calls, tail.first, tail.last, nfail,
         same ? "same as" : "DIFFERENT from");
  fails += !(nfail == 0 && same && st.n_gpu > 0 && st.n_gpu < n && tail.calls == 1 &&
             tail.first == st.n_gpu + 1 && tail.last == n && st.rate_gpu > 0.0);
  printf(fails ? "FAIL\n" : "OK\n");
  return fails ? 1 : 0;
}
