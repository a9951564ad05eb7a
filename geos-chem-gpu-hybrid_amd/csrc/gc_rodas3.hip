// gc_rodas3.hip -- batched KPP Rosenbrock Rodas3 on gfx950, one cell per lane.
//
// Mechanism kernels (Fun_SPLIT, Jac_SP fused with the matrix preparation and
// KppDecomp, KppSolve) are small loops that walk operand tables generated from
// the mechanism (codegen/gen_tables.py).  Every table word is wave-uniform and
// arrives through scalar loads; every data operand is one coalesced
// buffer_load_dwordx2 with the table word as soffset.  The code is a few KB
// whatever the mechanism, so it stays resident in the instruction cache.
//
// Two launch forms share these routines:
//  * plain      one thread per cell in Perm order, the nested time / reject /
//               singular loops of Integrate_Cell_S per thread -- the
//               reference's batched launch (gpulib/src/gc_gpu_batch_api.F90:
//               198-215, one `gang vector` thread per cell).
//  * persistent a fixed grid of lanes pulling cells from a device work queue
//               in Perm (cost) order; the per-cell loops are flattened into a
//               per-lane state machine whose unit of work is one factorisation
//               attempt, so lanes at different steps, retries or cells share
//               every attempt, and a lane whose cell finishes refills at once.
//
// Numerics are those of Integrate_Cell_S
// (src/geos-chem/KPP/fullchem/gckpp_BatchIntegrator.F90:408-707) with the option
// merge of gpu_solve_cell (gpulib/src/gc_gpu_batch_api.F90:54-110).  Two
// bit-neutral changes: the Jacobian is re-evaluated at every factorisation
// attempt instead of being kept as Jac0 (Y is unchanged between a step's
// attempts, so the values are identical and 45 KB of state per cell is
// saved), and KppDecomp's singularity test is applied to each prepared
// diagonal entry when it is formed (the reference tests each row's diagonal
// before that row is eliminated, i.e. the same value; a singular matrix is
// discarded in both).  Err**(1/3) uses the correctly rounded gc_pow_third.
//
// Workspace: one block per 64-lane wave; element e of lane l at byte
// e*512 + l*8 of the block (MECH_E_* element offsets).  Caller arrays keep the
// reference's Fortran layout (cell-major C(NSPEC,nCell), RCONST(NREACT,nCell)).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gc_math.h"
#include "mech_tables.inc"
#include "gc_rodas3_kernels.h"

#ifndef GC_VARIANT
#define GC_VARIANT 1
#endif
#if GC_VARIANT == 1
#define GC_WPE 2
#define GC_SEG 32
#define GC_LU_TABLE gc_lu32
#elif GC_VARIANT == 2
#define GC_WPE 4
#define GC_SEG 16
#define GC_LU_TABLE gc_lu16
#elif GC_VARIANT == 3
#define GC_WPE 3
#define GC_SEG 24
#define GC_LU_TABLE gc_lu24
#else
#error "unknown GC_VARIANT"
#endif
#define GC_CAT2(a, b) a##b
#define GC_CAT(a, b) GC_CAT2(a, b)
#define GC_SFX(name) GC_CAT(name, GC_CAT(_v, GC_VARIANT))

namespace {

constexpr int NVAR = MECH_NVAR;
constexpr int NFIX = MECH_NFIX;
constexpr int NSPEC = MECH_NVAR + MECH_NFIX;
constexpr int NREACT = MECH_NREACT;
constexpr int E_TOTAL = MECH_E_TOTAL;

// element byte offsets within a wave block
constexpr int O_Y = MECH_E_Y * 512;
constexpr int O_YNEW = MECH_E_YNEW * 512;
constexpr int O_FCN0 = MECH_E_FCN0 * 512;
constexpr int O_K = MECH_E_K * 512;
constexpr int O_DFDT = MECH_E_DFDT * 512;
constexpr int O_RCT = MECH_E_RCT * 512;
constexpr int O_FIX = MECH_E_FIX * 512;
constexpr int O_VEC = NVAR * 512;  // byte stride between consecutive NVAR vectors

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

typedef __attribute__((address_space(4))) const int ctab_t;
typedef __attribute__((address_space(4))) const double clit_t;
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return (a > b) ? a : b; }

__device__ __forceinline__ ctab_t* tab(const int* p) { return (ctab_t*)p; }

// The calling wave's workspace block and this lane's place in it.
struct Blk {
  __amdgpu_buffer_rsrc_t r;
  int voff;     // lane*8
  double* lds;  // this lane's slot-0 address in the wave's LDS row segment
  __device__ __forceinline__ double ld(int soff) const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  }
  __device__ __forceinline__ void st(int soff, double v) const {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
  }
  // the low 32 bits of an element (the mask words)
  __device__ __forceinline__ unsigned ld32(int soff) const {
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
  }
  __device__ __forceinline__ void st32(int soff, unsigned v) const {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, soff, 0);
  }
  __device__ __forceinline__ double lds_ld(int off) const {
    return *(double*)((char*)lds + off);
  }
  __device__ __forceinline__ void lds_st(int off, double v) const { *(double*)((char*)lds + off) = v; }
};

__device__ __forceinline__ Blk make_blk(double* ws, unsigned wave, int lane, double* lds_wave,
                                        int total = E_TOTAL) {
  const unsigned long long base =
      (unsigned long long)(ws + (size_t)wave * (size_t)total * 64ull);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)base);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(base >> 32));
  void* p = (void*)(((unsigned long long)hi << 32) | lo);
  Blk b;
  b.r = __builtin_amdgcn_make_buffer_rsrc(p, (short)0, total * 512, 0x00020000);
  b.voff = lane * 8;
  b.lds = lds_wave + lane;
  return b;
}

// One auto-reduce flag of the lane's cell (masked kernels): flag k of the
// mask whose first element is `elem0` is bit k%32 of element elem0 + k/32.
// Statements are visited in ascending (or descending) flag order, so the last
// word read is kept; k is wave-uniform (a table word), so is the reload test.
struct MaskWord {
  int w = -1;
  unsigned bits = 0u;
  __device__ __forceinline__ bool on(const Blk& b, int elem0, int k) {
    const int wi = k >> 5;
    if (wi != w) {
      w = wi;
      bits = b.ld32((elem0 + wi) * 512);
    }
    return ((bits >> (k & 31)) & 1u) != 0u;
  }
};

// Element passes: val(k) loads what element k needs (loads only), put(k, v)
// combines and stores it.  VB consecutive elements' loads are issued together
// before any of them is combined or stored, so a pass keeps VB elements' loads
// in flight instead of one round trip per element (the stores of a pass would
// otherwise order every later load behind them).  The scheduling fence keeps
// the compiler from sinking later elements' loads below earlier elements'
// arithmetic, which would put a wait for memory between every two elements.
// Valid for passes in which element k reads nothing that the pass stores at
// another index.
constexpr int VB = 8;

template <int N, class V, class P>
__device__ __forceinline__ void vpass(V val, P put) {
  int k = 0;
  for (; k + VB <= N; k += VB) {
    decltype(val(0)) v[VB];
#pragma unroll
    for (int u = 0; u < VB; ++u) v[u] = val(k + u);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < VB; ++u) put(k + u, v[u]);
  }
  for (; k < N; ++k) put(k, val(k));
}

// Table words arrive in 16-byte groups (one s_load_dwordx4 each): every
// record starts on a 4-word boundary of a 64-byte aligned table, and every
// table has at least 32 words of zero padding after its last record.
typedef int i4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(4))) const i4 ci4_t;

__device__ __forceinline__ i4 ld4(ctab_t* p) { return *(ci4_t*)p; }

__device__ __forceinline__ double dbl(int lo, int hi) {
  return __builtin_bit_cast(double, (unsigned long long)(unsigned)lo |
                                        ((unsigned long long)(unsigned)hi << 32));
}


// R product records {dst, lit, f0, f1, f2, f3, 0, 0} of shape (NF, LIT), each
// lit*f0*f1*... left to right: every record's factors (and literal) are
// fetched before any product is formed -- the fence stops the compiler from
// interleaving one record's loads with the previous record's multiplies, which
// would cost a memory round trip per record -- then the products are stored.
// Record words a = {dst, lit, f0, f1}, c = {f2, f3, 0, 0}.
template <int NF, bool LIT, int R>
__device__ __forceinline__ void prod_batch(const Blk& b, ctab_t* rec) {
  i4 a[R], c[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    a[q] = ld4(rec + 8 * q);
    if (NF > 2) c[q] = ld4(rec + 8 * q + 4);
  }
  double f[R][NF], lit[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (LIT) lit[q] = ((clit_t*)gc_lits)[a[q].y];
    f[q][0] = b.ld(a[q].z);
    if (NF > 1) f[q][1] = b.ld(a[q].w);
    if (NF > 2) f[q][2] = b.ld(c[q].x);
    if (NF > 3) f[q][3] = b.ld(c[q].y);
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    double v = f[q][0];
    if (LIT) v = lit[q] * v;
#pragma unroll
    for (int m = 1; m < NF; ++m) v = v * f[q][m];
    b.st(a[q].x, v);
  }
}

template <int NF, bool LIT>
__device__ __forceinline__ ctab_t* prod_group(const Blk& b, ctab_t* rec, int n) {
  int r = 0;
  for (; r + 8 <= n; r += 8, rec += 64) prod_batch<NF, LIT, 8>(b, rec);
  if (r + 4 <= n) {
    prod_batch<NF, LIT, 4>(b, rec);
    r += 4;
    rec += 32;
  }
  if (r + 2 <= n) {
    prod_batch<NF, LIT, 2>(b, rec);
    r += 2;
    rec += 16;
  }
  if (r < n) {
    prod_batch<NF, LIT, 1>(b, rec);
    rec += 8;
  }
  return rec;
}

// Product records grouped by shape (counts in grp[0..7], shape =
// (nf-1)*2 + has_lit): rates A(r), D_VAR term products, B temporaries.
__device__ void run_products(const Blk& b, const int* table, const int* grp_) {
  ctab_t* rec = tab(table);
  ctab_t* grp = tab(grp_);
  rec = prod_group<1, false>(b, rec, grp[0]);
  rec = prod_group<1, true>(b, rec, grp[1]);
  rec = prod_group<2, false>(b, rec, grp[2]);
  rec = prod_group<2, true>(b, rec, grp[3]);
  rec = prod_group<3, false>(b, rec, grp[4]);
  rec = prod_group<3, true>(b, rec, grp[5]);
  rec = prod_group<4, false>(b, rec, grp[6]);
  prod_group<4, true>(b, rec, grp[7]);
}

// One batch of a species' sums: PC P terms {soff, 0, coef} and DC D terms
// {f1, f2, f3, ci} at t, all their loads issued before either sum is
// advanced; each sum goes on in term order, its first term (fp / fd) setting
// the accumulator.
template <int PC, int DC>
__device__ __forceinline__ void pd_batch(const Blk& b, ctab_t* t, bool fp, bool fd, double& P,
                                         double& D) {
  i4 qp[PC > 0 ? PC : 1], qd[DC > 0 ? DC : 1];
#pragma unroll
  for (int r = 0; r < PC; ++r) qp[r] = ld4(t + 4 * r);
#pragma unroll
  for (int r = 0; r < DC; ++r) qd[r] = ld4(t + 4 * (PC + r));
  double x[PC > 0 ? PC : 1], f[DC > 0 ? 3 * DC : 1];
#pragma unroll
  for (int r = 0; r < PC; ++r) x[r] = b.ld(qp[r].x);
#pragma unroll
  for (int r = 0; r < DC; ++r) {
    f[3 * r] = b.ld(qd[r].x);
    f[3 * r + 1] = b.ld(qd[r].y);
    f[3 * r + 2] = b.ld(qd[r].z);
  }
#pragma unroll
  for (int r = 0; r < PC; ++r) {
    const double v = dbl(qp[r].z, qp[r].w) * x[r];
    P = (r == 0 && fp) ? v : P + v;
  }
#pragma unroll
  for (int r = 0; r < DC; ++r) {
    const double v = ((((clit_t*)gc_lits)[qd[r].w] * f[3 * r]) * f[3 * r + 1]) * f[3 * r + 2];
    D = (r == 0 && fd) ? v : D + v;
  }
}

// Fun_SPLIT (gckpp_Function.F90:2170-5060): rates, then per species P_VAR,
// D_VAR (term products evaluated in place) and Vdot = P_VAR - D_VAR*V,
// stored at out + i*512.  A species' terms come in batches of up to 8 P and
// 4 D terms (both sums padded to multiples of 4 with -0.0 terms), one round
// trip per batch.
// Masked (M): IF (DO_FUN(i)) guards both sums of species i
// (gckpp_Function.F90:3280-5056); a species whose flag is off keeps the P_VAR
// and D_VAR the cell's workspace holds, zero from the cell's start, so its
// Vdot is 0 - 0*V.
template <bool M>
__device__ void fun(const Blk& b, bool ynew, int out) {
  if (ynew) run_products(b, gc_prod_rate_yn, gc_prodgrp_rate_yn);
  else run_products(b, gc_prod_rate_y, gc_prodgrp_rate_y);
  ctab_t* t = tab(ynew ? gc_pd_yn : gc_pd_y);
  MaskWord mf;
  for (int i = 0; i < NVAR; ++i) {
    const i4 h = ld4(t);
    const int nP = h.x, nD = h.y, vo = h.z, oo = h.w;
    t += 4;
    const double v = b.ld(vo);
    double P = 0.0, D = 0.0;
    for (int kp = 0, kd = 0; kp < nP || kd < nD;) {
      const int pc = (nP - kp < 8) ? nP - kp : 8;  // 0, 4 or 8
      const int dc = (nD - kd < 4) ? nD - kd : 4;  // 0 or 4
      const bool fp = kp == 0, fd = kd == 0;
      if (pc == 8) {
        if (dc == 4) pd_batch<8, 4>(b, t, fp, fd, P, D);
        else pd_batch<8, 0>(b, t, fp, fd, P, D);
      } else if (pc == 4) {
        if (dc == 4) pd_batch<4, 4>(b, t, fp, fd, P, D);
        else pd_batch<4, 0>(b, t, fp, fd, P, D);
      } else {
        pd_batch<0, 4>(b, t, fp, fd, P, D);
      }
      t += 4 * (pc + dc);
      kp += pc;
      kd += dc;
    }
    if (M && !mf.on(b, MECH_E_MFUN, oo >> 9)) {
      P = 0.0;
      D = 0.0;
    }
    b.st(out + oo, P - D * v);
  }
}

// One term {soff, flags, coef} of an ENTRYG record with its operand x: the
// entry's sum restarts at a first term; at its last term the prepared entry
// -sum (+ ghinv on the diagonal) goes to its LDS slot.
// Masked (M): IF (DO_JVS(k)) guards the entry's sum (gckpp_Jacobian.F90:
// 3676-20887); an entry whose flag is off keeps the Jac0 value of the cell's
// workspace, zero from the cell's start, so Ghimj(k) = -0.0 (+ ghinv on the
// diagonal).
template <bool M>
__device__ __forceinline__ void entry_term(const Blk& b, i4 t, double x, double ghinv,
                                           double& acc, int& sing, MaskWord& mj) {
  const int fl = t.y;
  const double cx = dbl(t.z, t.w) * x;
  acc = (fl & 1) ? cx : acc + cx;
  if (fl & 2) {
    if (M && !mj.on(b, MECH_E_MJVS, fl >> 14)) acc = 0.0;
    double g = -acc;
    if (fl & 4) {
      g = g + ghinv;
      sing |= (__builtin_fabs(g) < TINYD) ? 1 : 0;
    }
    b.lds_st(((fl >> 8) & 63) * 512, g);
  }
}

// A PIVB batch of G pivot sub-records of four updates each: header {8, G, 0,
// 0}, G packed words {kind | sj << 2 | t0 << 8 | t1 << 14 | t2 << 20 |
// t3 << 26} (LDS slot indices; padded to a multiple of 4 words), then per
// sub-record {operand, U0, U1, U2, U3} (padded to a multiple of 4 words).
// Every data load of the batch is issued before the first sub-record is
// applied; sub-records then go in order, each reading its multiplier's inputs
// from the LDS row segment after the previous one's updates: kind 1 (PIVL)
// a = -W(j)/U(j,j) with W(j) = -a stored back, kind 2 (PIVS) a = -L(k,j),
// kind 0 continues the previous sub-record's multiplier (a record of more
// than four updates).  W(t_k) += a*U_k; padding updates hit the spare slot
// with U = 0.0.
template <int G>
__device__ __forceinline__ ctab_t* piv_batch(const Blk& b, ctab_t* op, double& a) {
  constexpr int NPK = (G + 3) / 4, ND = (5 * G + 3) / 4;
  i4 pk[NPK], dw[ND];
#pragma unroll
  for (int i = 0; i < NPK; ++i) pk[i] = ld4(op + 4 + 4 * i);
#pragma unroll
  for (int i = 0; i < ND; ++i) dw[i] = ld4(op + 4 + 4 * NPK + 4 * i);
  auto word = [&](const i4* q, int k) {
    const i4 v = q[k >> 2];
    return (k & 3) == 0 ? v.x : (k & 3) == 1 ? v.y : (k & 3) == 2 ? v.z : v.w;
  };
  double v[5 * G];
#pragma unroll
  for (int k = 0; k < 5 * G; ++k) v[k] = b.ld(word(dw, k));
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const unsigned p = (unsigned)word(pk, g);
    const unsigned kind = p & 3u;
    if (kind == 1u) {
      const int sj = (int)((p >> 2) & 63u) * 512;
      a = -b.lds_ld(sj) / v[5 * g];
      b.lds_st(sj, -a);
    } else if (kind == 2u) {
      a = -v[5 * g];
    }
    const int t0 = (int)((p >> 8) & 63u) * 512, t1 = (int)((p >> 14) & 63u) * 512;
    const int t2 = (int)((p >> 20) & 63u) * 512, t3 = (int)(p >> 26) * 512;
    const double w0 = b.lds_ld(t0), w1 = b.lds_ld(t1), w2 = b.lds_ld(t2), w3 = b.lds_ld(t3);
    b.lds_st(t0, w0 + a * v[5 * g + 1]);
    b.lds_st(t1, w1 + a * v[5 * g + 2]);
    b.lds_st(t2, w2 + a * v[5 * g + 3]);
    b.lds_st(t3, w3 + a * v[5 * g + 4]);
  }
  return op + 4 + 4 * NPK + 4 * ND;
}

// Jac_SP fused with Ghimj = -Jac0, diag += ghinv and KppDecomp, row segments
// in LDS (the lu table format is described in codegen/gen_tables.py).
// Returns 1 when a prepared diagonal entry is below TINY.
template <bool M>
__device__ int jac_lu(const Blk& b, double ghinv) {
  run_products(b, gc_prod_bjac, gc_prodgrp_bjac);
  ctab_t* op = tab(GC_LU_TABLE);
  int sing = 0;
  double a = 0.0;  // the current pivot multiplier (carried across PIVB batches)
  MaskWord mj;
  for (;;) {
    const i4 h = ld4(op), f = ld4(op + 4);  // header and the first 4 words after it
    const int code = h.x;
    if (code == 0) break;
    if (code == 5) {
      // every entry of the segment: T terms (a multiple of 4), sixteen, eight
      // or four loads in flight at a time; an entry may span batches
      const int T = h.y;
      op += 4;
      double acc = 0.0;
      int q = 0;
      for (; q + 16 <= T; q += 16, op += 64) {
        i4 tt[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) tt[r] = ld4(op + 4 * r);
        double x[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = b.ld(tt[r].x);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < 16; ++r) entry_term<M>(b, tt[r], x[r], ghinv, acc, sing, mj);
      }
      for (; q + 8 <= T; q += 8, op += 32) {
        i4 tt[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) tt[r] = ld4(op + 4 * r);
        double x[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) x[r] = b.ld(tt[r].x);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < 8; ++r) entry_term<M>(b, tt[r], x[r], ghinv, acc, sing, mj);
      }
      for (; q < T; q += 4, op += 16) {
        i4 tt[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) tt[r] = ld4(op + 4 * r);
        double x[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) x[r] = b.ld(tt[r].x);
#pragma unroll
        for (int r = 0; r < 4; ++r) entry_term<M>(b, tt[r], x[r], ghinv, acc, sing, mj);
      }
    } else if (code == 8) {
      switch (h.y) {
        case 1: op = piv_batch<1>(b, op, a); break;
        case 2: op = piv_batch<2>(b, op, a); break;
        case 3: op = piv_batch<3>(b, op, a); break;
        case 4: op = piv_batch<4>(b, op, a); break;
        case 5: op = piv_batch<5>(b, op, a); break;
        default: op = piv_batch<6>(b, op, a); break;
      }
    } else {  // code 4: store the segment, pairs {slot, LU} (padding writes the sink)
      const int n = h.y;
      op += 4;
      i4 q = f;
      for (int u = 0; u < n; u += 2, op += 4) {
        if (u > 0) q = ld4(op);
        const double v0 = b.lds_ld(q.x), v1 = b.lds_ld(q.z);
        b.st(q.y, v0);
        b.st(q.w, v1);
      }
    }
  }
  return sing;
}

// Rows {n, x_i, diag, 0} + n/2 {LU, x_j, LU, x_j} (n a multiple of 4):
// x_i = x_i - LU*x_j ... in order, then / diag for the backward rows.
// Masked (M): IF (DO_SLV(i)) guards the row (gckpp_LinearAlgebra.F90:
// 653-2309), x_i stays as it was; the rows' x_i are elements xbase + i.
// A row's header and first eight pairs arrive with the previous row (fetched
// while that row's data loads are in flight), so a row of up to eight pairs
// costs one memory round trip and no wait for its table words.
template <bool DIV, bool M = false>
__device__ __forceinline__ void solve_rows(const Blk& b, ctab_t* t, int rows, int xbase = 0) {
  MaskWord ms;
  i4 h = ld4(t), q0 = ld4(t + 4), q1 = ld4(t + 8), q2 = ld4(t + 12), q3 = ld4(t + 16);
  for (int r = 0; r < rows; ++r) {
    const int n = h.x, xi = h.y, dg = h.z;
    ctab_t* const tn = t + 4 + 2 * n;  // the next row
    double x = b.ld(xi);
    double dv = 1.0;
    if (DIV) dv = b.ld(dg);
    double l[8], y[8];
    if (n >= 4) {
      l[0] = b.ld(q0.x), y[0] = b.ld(q0.y), l[1] = b.ld(q0.z), y[1] = b.ld(q0.w);
      l[2] = b.ld(q1.x), y[2] = b.ld(q1.y), l[3] = b.ld(q1.z), y[3] = b.ld(q1.w);
    }
    if (n >= 8) {
      l[4] = b.ld(q2.x), y[4] = b.ld(q2.y), l[5] = b.ld(q2.z), y[5] = b.ld(q2.w);
      l[6] = b.ld(q3.x), y[6] = b.ld(q3.y), l[7] = b.ld(q3.z), y[7] = b.ld(q3.w);
    }
    const i4 hn = ld4(tn), n0 = ld4(tn + 4), n1 = ld4(tn + 8), n2 = ld4(tn + 12),
             n3 = ld4(tn + 16);
    __builtin_amdgcn_sched_barrier(0);
    const double x0 = x;
    if (n >= 4) {
#pragma unroll
      for (int k = 0; k < 4; ++k) x = x - l[k] * y[k];
    }
    if (n >= 8) {
#pragma unroll
      for (int k = 4; k < 8; ++k) x = x - l[k] * y[k];
    }
    t += 20;
    int u = 8;
    for (; u + 8 <= n; u += 8, t += 16) {
      i4 q[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) q[r] = ld4(t + 4 * r);
      double l[8], y[8];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        l[2 * r] = b.ld(q[r].x);
        y[2 * r] = b.ld(q[r].y);
        l[2 * r + 1] = b.ld(q[r].z);
        y[2 * r + 1] = b.ld(q[r].w);
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) x = x - l[r] * y[r];
    }
    for (; u < n; u += 4, t += 8) {
      const i4 q01 = ld4(t), q23 = ld4(t + 4);
      const double l0 = b.ld(q01.x), y0 = b.ld(q01.y), l1 = b.ld(q01.z), y1 = b.ld(q01.w);
      const double l2 = b.ld(q23.x), y2 = b.ld(q23.y), l3 = b.ld(q23.z), y3 = b.ld(q23.w);
      x = x - l0 * y0;
      x = x - l1 * y1;
      x = x - l2 * y2;
      x = x - l3 * y3;
    }
    if (DIV) x = x / dv;
    if (M && !ms.on(b, MECH_E_MSLV, (xi >> 9) - xbase)) x = x0;
    b.st(xi, x);
    h = hn;
    q0 = n0;
    q1 = n1;
    q2 = n2;
    q3 = n3;
    t = tn;
  }
}

// Backward rows whose x_j of the last GC_SEG columns sit in this lane's LDS
// segment slots: header {nc | nh << 16, x_i, U(i,i), lds_i} + nc/2
// {LU, x_j, LU, x_j} + nh/2 {LU, slot, LU, slot} (nc, nh multiples of 4):
// x_i = x_i - LU*x_j over the cold then the LDS pairs (KppSolve's ascending
// column order), / U(i,i); a row of that suffix also leaves x_i in its slot
// (lds_i >= 0), before any later row reads it.  LDS padding pairs read the
// spare slot, set to 0.0 here: x - 0.0*0.0 == x.
template <bool M>
__device__ __forceinline__ void solve_bwd_lds(const Blk& b, ctab_t* t, int rows, int xbase) {
  b.lds_st(GC_SEG * 512, 0.0);
  MaskWord ms;
  // a row's header and the words of its first eight pairs arrive with the
  // previous row, as in solve_rows
  i4 h = ld4(t), q0 = ld4(t + 4), q1 = ld4(t + 8), q2 = ld4(t + 12), q3 = ld4(t + 16);
  for (int r = 0; r < rows; ++r) {
    const int nc = h.x & 0xffff, nh = h.x >> 16;
    ctab_t* const tn = t + 4 + 2 * (nc + nh);  // the next row
    double x = b.ld(h.y);
    const double dv = b.ld(h.z);
    // the first eight pairs: cold ones first (two loads), then LDS ones
    // (their x_j from the slot); cc cold and hc LDS pairs among them
    const int cc = nc < 8 ? nc : 8;
    const int hc = nh < 8 - cc ? nh : 8 - cc;
    double l[8], y[8];
    if (cc >= 4) {
      l[0] = b.ld(q0.x), y[0] = b.ld(q0.y), l[1] = b.ld(q0.z), y[1] = b.ld(q0.w);
      l[2] = b.ld(q1.x), y[2] = b.ld(q1.y), l[3] = b.ld(q1.z), y[3] = b.ld(q1.w);
    } else if (hc >= 4) {
      l[0] = b.ld(q0.x), y[0] = b.lds_ld(q0.y), l[1] = b.ld(q0.z), y[1] = b.lds_ld(q0.w);
      l[2] = b.ld(q1.x), y[2] = b.lds_ld(q1.y), l[3] = b.ld(q1.z), y[3] = b.lds_ld(q1.w);
    }
    if (cc == 8) {
      l[4] = b.ld(q2.x), y[4] = b.ld(q2.y), l[5] = b.ld(q2.z), y[5] = b.ld(q2.w);
      l[6] = b.ld(q3.x), y[6] = b.ld(q3.y), l[7] = b.ld(q3.z), y[7] = b.ld(q3.w);
    } else if (cc + hc == 8) {
      l[4] = b.ld(q2.x), y[4] = b.lds_ld(q2.y), l[5] = b.ld(q2.z), y[5] = b.lds_ld(q2.w);
      l[6] = b.ld(q3.x), y[6] = b.lds_ld(q3.y), l[7] = b.ld(q3.z), y[7] = b.lds_ld(q3.w);
    }
    const i4 hn = ld4(tn), n0 = ld4(tn + 4), n1 = ld4(tn + 8), n2 = ld4(tn + 12),
             n3 = ld4(tn + 16);
    __builtin_amdgcn_sched_barrier(0);
    const double x0 = x;
    if (cc + hc >= 4) {
#pragma unroll
      for (int k = 0; k < 4; ++k) x = x - l[k] * y[k];
    }
    if (cc + hc == 8) {
#pragma unroll
      for (int k = 4; k < 8; ++k) x = x - l[k] * y[k];
    }
    t += 20;  // the pairs after the first eight
    int u = cc;
    for (; u + 8 <= nc; u += 8, t += 16) {
      i4 q[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) q[k] = ld4(t + 4 * k);
      double lq[8], yq[8];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        lq[2 * k] = b.ld(q[k].x);
        yq[2 * k] = b.ld(q[k].y);
        lq[2 * k + 1] = b.ld(q[k].z);
        yq[2 * k + 1] = b.ld(q[k].w);
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) x = x - lq[k] * yq[k];
    }
    for (; u < nc; u += 4, t += 8) {
      const i4 q01 = ld4(t), q23 = ld4(t + 4);
      const double l0 = b.ld(q01.x), y0 = b.ld(q01.y), l1 = b.ld(q01.z), y1 = b.ld(q01.w);
      const double l2 = b.ld(q23.x), y2 = b.ld(q23.y), l3 = b.ld(q23.z), y3 = b.ld(q23.w);
      x = x - l0 * y0;
      x = x - l1 * y1;
      x = x - l2 * y2;
      x = x - l3 * y3;
    }
    u = hc;
    for (; u + 8 <= nh; u += 8, t += 16) {
      i4 q[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) q[k] = ld4(t + 4 * k);
      double lq[8], yq[8];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        lq[2 * k] = b.ld(q[k].x);
        lq[2 * k + 1] = b.ld(q[k].z);
        yq[2 * k] = b.lds_ld(q[k].y);
        yq[2 * k + 1] = b.lds_ld(q[k].w);
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) x = x - lq[k] * yq[k];
    }
    for (; u < nh; u += 4, t += 8) {
      const i4 q01 = ld4(t), q23 = ld4(t + 4);
      const double l0 = b.ld(q01.x), l1 = b.ld(q01.z), l2 = b.ld(q23.x), l3 = b.ld(q23.z);
      const double y0 = b.lds_ld(q01.y), y1 = b.lds_ld(q01.w);
      const double y2 = b.lds_ld(q23.y), y3 = b.lds_ld(q23.w);
      x = x - l0 * y0;
      x = x - l1 * y1;
      x = x - l2 * y2;
      x = x - l3 * y3;
    }
    x = x / dv;
    if (M && !ms.on(b, MECH_E_MSLV, (h.y >> 9) - xbase)) x = x0;
    b.st(h.y, x);
    if (h.w >= 0) b.lds_st(h.w, x);
    h = hn;
    q0 = n0;
    q1 = n1;
    q2 = n2;
    q3 = n3;
    t = tn;
  }
}

// KppSolve (gckpp_LinearAlgebra.F90:645-2311) in place on stage vector K_s.
template <bool M>
__device__ void solve(const Blk& b, int s) {
  const int* fw = s == 0 ? gc_fwd0 : s == 1 ? gc_fwd1 : s == 2 ? gc_fwd2 : gc_fwd3;
  const int* bw = s == 0   ? GC_CAT(gc_bwd0_, GC_SEG)
                  : s == 1 ? GC_CAT(gc_bwd1_, GC_SEG)
                  : s == 2 ? GC_CAT(gc_bwd2_, GC_SEG)
                           : GC_CAT(gc_bwd3_, GC_SEG);
  const int xbase = MECH_E_K + s * NVAR;
  solve_rows<false, M>(b, tab(fw), MECH_NFWD, xbase);
  solve_bwd_lds<M>(b, tab(bw), MECH_NBWD, xbase);
}

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

// The cell's auto-reduce flags into its mask elements (a null mask: every
// flag on): DO_FUN, DO_JVS, DO_SLV words of the caller's per-cell arrays.
__device__ void gather_masks(const GcBatchArgs& a, long cell, const Blk& b) {
  for (int w = 0; w < MECH_MW_FUN; ++w)
    b.st32((MECH_E_MFUN + w) * 512, a.doFun ? a.doFun[cell * MECH_MW_FUN + w] : ~0u);
  for (int w = 0; w < MECH_MW_JVS; ++w)
    b.st32((MECH_E_MJVS + w) * 512, a.doJvs ? a.doJvs[cell * MECH_MW_JVS + w] : ~0u);
  for (int w = 0; w < MECH_MW_SLV; ++w)
    b.st32((MECH_E_MSLV + w) * 512, a.doSlv ? a.doSlv[cell * MECH_MW_SLV + w] : ~0u);
}

// Copy cell `cell` (0-based column of the caller arrays) into the lane.
template <bool M>
__device__ void gather(const GcBatchArgs& a, long cell, const Blk& b) {
  if (M) gather_masks(a, cell, b);
  const double* c = a.C + cell * NSPEC;
  const double* r = a.RCONST + cell * NREACT;
  vpass<NVAR>([&](int i) { return c[i]; }, [&](int i, double v) { b.st(O_Y + i * 512, v); });
  for (int j = 0; j < NFIX; ++j) b.st(O_FIX + j * 512, c[NVAR + j]);
  vpass<NREACT>([&](int k) { return r[k]; }, [&](int k, double v) { b.st(O_RCT + k * 512, v); });
  b.st(MECH_E_ZERO * 512, 0.0);  // padding operand of the solve tables
  b.st(MECH_E_ONE * 512, 1.0);   // missing factor of a D_VAR term
}

__device__ void scatter(const GcBatchArgs& a, long cell, const Blk& b, const int* ist,
                        const double* rst, int ierr) {
  double* c = a.C + cell * NSPEC;
  vpass<NVAR>([&](int i) { return b.ld(O_Y + i * 512); }, [&](int i, double v) { c[i] = v; });
  for (int k = 0; k < 8; ++k) a.ISTAT[cell * 8 + k] = ist[k];
  for (int k = 0; k < 3; ++k) a.RSTAT[cell * 3 + k] = rst[k];
  a.IERR[cell] = ierr;
}

// Fcn0 <- Fun(Y); for non-autonomous problems also the dFdT replica
// (gckpp_BatchIntegrator.F90:549-561).
template <bool M>
__device__ void step_functions(const Blk& b, const Opts& o, double T, int* ist) {
  fun<M>(b, false, O_FCN0);
  ist[0] += 1;
  if (!o.autonomous) {
    const double Delta = sqrt(ROUNDOFF) * dmax(1.0e-6, fabs(T));
    fun<M>(b, false, O_DFDT);
    ist[0] += 1;
    struct L2 {
      double d, f;
    };
    vpass<NVAR>([&](int k) { return L2{b.ld(O_DFDT + k * 512), b.ld(O_FCN0 + k * 512)}; },
                [&](int k, const L2& l) { b.st(O_DFDT + k * 512, (l.d - l.f) / Delta); });
  }
}

// Ynew = Y + sum_j A_sj K_j, the point of stage S >= 2 (ros_NewF).
template <int S>
__device__ void stage_point(const Blk& b) {
  struct L {
    double y, k[S];
  };
  vpass<NVAR>(
      [&](int k) {
        L l;
        l.y = b.ld(O_Y + k * 512);
#pragma unroll
        for (int j = 0; j < S; ++j) l.k[j] = b.ld(O_K + j * O_VEC + k * 512);
        return l;
      },
      [&](int k, const L& l) {
        double y = l.y;
#pragma unroll
        for (int j = 0; j < S; ++j) y = y + ros_A[S * (S - 1) / 2 + j] * l.k[j];
        b.st(O_YNEW + k * 512, y);
      });
}

__device__ __forceinline__ void stage_point(const Blk& b, int s) {
  if (s == 2) stage_point<2>(b);
  else stage_point<3>(b);
}

// The four stages after a successful factorisation
// (gckpp_BatchIntegrator.F90:594-643).
// Stage S (0-based) after a successful factorisation
// (gckpp_BatchIntegrator.F90:594-643).
template <bool M, int S>
__device__ __forceinline__ void stage(const Blk& b, bool autonomous, double Dir, double H,
                                      int* ist) {
  const int Ks = O_K + S * O_VEC;
  // K_s = Fcn + sum_j HC_j K_j, one pass (the adds in j order, as the
  // reference's successive WAXPY calls)
  int src = Ks;
  if (S == 0 || S == 1) {
    // stage 1: Fcn = Fcn0; stage 2 has ros_NewF = .FALSE. so Fcn is still Fcn0
    src = O_FCN0;
  } else {
    stage_point<S>(b);
    fun<M>(b, true, Ks);
    ist[0] += 1;
  }
  double HC[S > 0 ? S : 1];
#pragma unroll
  for (int j = 0; j < S; ++j) HC[j] = ros_C[S * (S - 1) / 2 + j] / (Dir * H);
  struct L {
    double f, k[S > 0 ? S : 1];
  };
  vpass<NVAR>(
      [&](int k) {
        L l;
        l.f = b.ld(src + k * 512);
#pragma unroll
        for (int j = 0; j < S; ++j) l.k[j] = b.ld(O_K + j * O_VEC + k * 512);
        return l;
      },
      [&](int k, const L& l) {
        double v = l.f;
#pragma unroll
        for (int j = 0; j < S; ++j) v = v + HC[j] * l.k[j];
        b.st(Ks + k * 512, v);
      });
  if (!autonomous && ros_Gamma[S] != 0.0) {
    const double HG = (Dir * H) * ros_Gamma[S];
    struct L2 {
      double k, d;
    };
    vpass<NVAR>([&](int k) { return L2{b.ld(Ks + k * 512), b.ld(O_DFDT + k * 512)}; },
                [&](int k, const L2& l) { b.st(Ks + k * 512, l.k + HG * l.d); });
  }
  solve<M>(b, S);
  ist[6] += 1;
}

template <bool M>
__device__ void stages(const Blk& b, bool autonomous, double Dir, double H, int* ist) {
  stage<M, 0>(b, autonomous, Dir, H, ist);
  stage<M, 1>(b, autonomous, Dir, H, ist);
  stage<M, 2>(b, autonomous, Dir, H, ist);
  stage<M, 3>(b, autonomous, Dir, H, ist);
}

// The stages, then Ynew, Yerr and the error norm
// (gckpp_BatchIntegrator.F90:594-665).  Returns Err.
template <bool M>
__device__ double stages_and_error(const Blk& b, const Opts& o, double Dir, double H,
                                   const double* atol, const double* rtol, int* ist) {
  stages<M>(b, o.autonomous, Dir, H, ist);
  double Err = 0.0;
  struct YE {
    double y, k4[4], at, rt;
  };
  vpass<NVAR>(
      [&](int k) {
        YE v;
        v.y = b.ld(O_Y + k * 512);
#pragma unroll
        for (int j = 0; j < 4; ++j) v.k4[j] = b.ld(O_K + j * O_VEC + k * 512);
        v.at = o.vector_tol ? atol[k] : atol[0];
        v.rt = o.vector_tol ? rtol[k] : rtol[0];
        return v;
      },
      [&](int k, const YE& v) {  // Err accumulates in element order
        double yn = v.y, ye = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) yn = yn + ros_M[j] * v.k4[j];
#pragma unroll
        for (int j = 0; j < 4; ++j) ye = ye + ros_E[j] * v.k4[j];
        b.st(O_YNEW + k * 512, yn);
        const double Ymax = dmax(fabs(v.y), fabs(yn));
        const double Scal = v.at + v.rt * Ymax;
        const double q = ye / Scal;
        Err = Err + q * q;
      });
  Err = sqrt(Err / (double)NVAR);
  return dmax(Err, 1.0e-10);
}

__device__ void accept_y(const Blk& b, const Opts& o) {
  if (o.clamp) {
    vpass<NVAR>([&](int k) { return b.ld(O_YNEW + k * 512); },
                [&](int k, double v) { b.st(O_Y + k * 512, dmax(v, 0.0)); });
  } else {
    vpass<NVAR>([&](int k) { return b.ld(O_YNEW + k * 512); },
                [&](int k, double v) { b.st(O_Y + k * 512, v); });
  }
}

// Y and the four stage vectors of element k, and Ynew = Y + sum_j M_j K_j
// from them (gckpp_BatchIntegrator.F90:636-640, the stepper's order).
struct YK {
  double y, k[4];
};

__device__ __forceinline__ YK ld_yk(const Blk& b, int k) {
  YK l;
  l.y = b.ld(O_Y + k * 512);
#pragma unroll
  for (int j = 0; j < 4; ++j) l.k[j] = b.ld(O_K + j * O_VEC + k * 512);
  return l;
}

__device__ __forceinline__ double ynew_of(const YK& l) {
  double yn = l.y;
#pragma unroll
  for (int j = 0; j < 4; ++j) yn = yn + ros_M[j] * l.k[j];
  return yn;
}

__device__ __forceinline__ bool time_left(int dir, double T, double Tend) {
  return (dir > 0 && ((T - Tend) + ROUNDOFF <= 0.0)) || (dir < 0 && ((Tend - T) + ROUNDOFF <= 0.0));
}

// plain: Integrate_Cell_S per thread, with its own nested loops.
template <bool M>
__device__ int integrate_cell(const Blk& b, const Opts& o, double Tstart, double Tend,
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
    step_functions<M>(b, o, T, ist);
    ist[1] += 1;  // Jac_SP of this step (evaluated in each attempt below)
    for (;;) {    // UntilAccepted
      int ncons = 0;
      for (;;) {  // singular retries
        const double ghinv = 1.0 / ((Dir * H) * ros_Gamma[0]);
        const int sing = jac_lu<M>(b, ghinv);
        ist[5] += 1;
        if (!sing) break;
        ist[7] += 1;
        ncons += 1;
        if (ncons <= 5) H = H * 0.5;
        else return -8;
      }
      const double Err = stages_and_error<M>(b, o, Dir, H, atol, rtol, ist);
      const double Fac = dmin(o.FacMax, dmax(o.FacMin, o.FacSafe / gc_pow_third(Err)));
      double Hnew = H * Fac;
      Nsteps += 1;
      ist[2] += 1;
      if ((Err <= 1.0) || (H <= o.Hmin)) {
        ist[3] += 1;
        accept_y(b, o);
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

enum : int { PH_FREE = 0, PH_STEP = 1, PH_RETRY = 2, PH_SING = 3 };

template <bool M>
__device__ __forceinline__ void plain_body(const GcBatchArgs& a, double* ws) {
  extern __shared__ double lds[];
  const unsigned ic = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const Blk b = make_blk(ws, ic >> 6, lane, lds + (threadIdx.x >> 6) * ((GC_SEG + 1) * 64));
  if (ic >= (unsigned)a.nCell) return;
  const long cell = (long)a.Perm[ic] - 1;
  int ist[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double rst[3] = {0.0, 0.0, 0.0};
  gather<M>(a, cell, b);
  Opts o;
  int ierr = decode_opts(a.ICNTRL + cell * 20, a.RCNTRL + cell * 20, a.Tstart, a.Tend, o);
  if (ierr == 0) ierr = integrate_cell<M>(b, o, a.Tstart, a.Tend, a.ATOL, a.RTOL, ist, rst);
  scatter(a, cell, b, ist, rst, ierr);
}

}  // namespace

extern "C" __global__ void
__attribute__((amdgpu_flat_work_group_size(64, 256), amdgpu_waves_per_eu(GC_WPE, GC_WPE)))
GC_SFX(gc_rodas3_plain)(GcBatchArgs a, double* ws) {
  plain_body<false>(a, ws);
}

// The plain kernel with the cells' auto-reduce masks (DO_FUN, DO_JVS, DO_SLV).
extern "C" __global__ void
__attribute__((amdgpu_flat_work_group_size(64, 256), amdgpu_waves_per_eu(GC_WPE, GC_WPE)))
GC_SFX(gc_rodas3_plain_masked)(GcBatchArgs a, double* ws) {
  plain_body<true>(a, ws);
}

namespace {

// Wave-aggregated slot from a pool counter (one atomic per wave): the lanes
// with `want` get consecutive slots; the others get ~0u.
__device__ __forceinline__ unsigned take_slot(unsigned* counter, bool want, int lane) {
  const unsigned long long mask = __ballot(want);
  if (mask == 0ull) return ~0u;
  const int leader = __ffsll((long long)mask) - 1;
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned)__popcll(mask));
  base = __shfl(base, leader);
  return want ? base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull)) : ~0u;
}

// Record of an accepted step (recording form): its H and the cell's previous
// record; the start state too when the step is a checkpoint (its index within
// the cell, idx, a multiple of rec.every).  Called by every lane of the wave.
__device__ __forceinline__ void record_step(const Blk& b, const GcRecord& rec, bool acc,
                                            int idx, double H, int lane, int& head) {
  const unsigned slot = take_slot(rec.count, acc, lane);
  const bool ck = acc && (idx % rec.every) == 0;
  const unsigned ys = take_slot(rec.ycount, ck, lane);
  if (!acc || slot >= rec.cap) return;  // pool exhausted: the host sees count > cap
  int yslot = -1;
  if (ck && ys < rec.ycap) {
    double* y = rec.Y + (size_t)ys * NVAR;
    vpass<NVAR>([&](int k) { return b.ld(O_Y + k * 512); }, [&](int k, double v) { y[k] = v; });
    if (rec.K) {  // the accepted step's stage vectors K_1..K_4 (contiguous elements)
      double* kk = rec.K + (size_t)ys * 4 * NVAR;
      vpass<4 * NVAR>([&](int k) { return b.ld(O_K + k * 512); },
                      [&](int k, double v) { kk[k] = v; });
    }
    yslot = (int)ys;
  }
  rec.H[slot] = H;
  rec.prev[slot] = head;
  rec.yslot[slot] = yslot;
  head = (int)slot;
}

// Waits until the input chunks of launch positions [q, q_end) have landed
// (GcGate); seen = chunks 0..seen-1 already observed by this wave.  Returns
// false when a wait ran past the gate's bound.  Wave-uniform.
__device__ __forceinline__ bool gate_wait(const GcGate& g, unsigned q, unsigned q_end,
                                          unsigned& seen) {
  const unsigned k_end = (q_end - 1u) / g.cells + 1u;
  if (k_end <= seen) return true;
  for (unsigned k = seen; k < k_end; ++k) {
    unsigned long long n = 0;
    while (__builtin_amdgcn_readfirstlane(
               __hip_atomic_load(g.flags + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) == 0u) {
      if (++n > g.spin) return false;
      __builtin_amdgcn_s_sleep(127);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // the chunks' data is visible from here on
  seen = k_end;
  return true;
}

// The persistent grid over one batch `a`, or (MB) over nb batches described
// by descs in one queue of qn positions: batch k's cells take positions
// descs[k].q0 .. descs[k].q0 + descs[k].a.nCell - 1 in its own launch order,
// each lane working with its current cell's batch arguments.  A cell's record
// chain head (recording form) is head[q0 + cell].  GATE: positions are handed
// out only once their input chunk has landed (GcGate).
template <bool REC, bool M = false, bool MB = false, bool GATE = false>
__device__ __forceinline__ void persistent_body(const GcBatchArgs& a, double* ws, unsigned S,
                                                unsigned* queue, int cells_per_wave, int grain,
                                                const GcRecord& rec,
                                                const GcBatchDesc* descs = nullptr, int nb = 0,
                                                unsigned qn = 0, const GcGate& gate = GcGate{}) {
  extern __shared__ double lds[];
  const unsigned slot = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const Blk b = make_blk(ws, slot >> 6, lane, lds + (threadIdx.x >> 6) * ((GC_SEG + 1) * 64));
  const bool lane_on = (lane < cells_per_wave) && (slot < S);

  int phase = PH_FREE;
  long cell = -1;
  Opts o;
  int ist[8];
  double rst[3];
  double T = 0.0, H = 0.0, Dir = 1.0;
  int dir = 1, Nsteps = 0, ncons = 0;
  bool rejLast = false, rejMore = false;
  bool exhausted = !lane_on;
  unsigned rsv = 0, rsv_end = 0;  // wave-uniform reservation [rsv, rsv_end)
  int head = -1;                  // last record of the lane's cell (recording form)
  GcBatchArgs la = a;             // the arguments of the lane's cell's batch
  unsigned q0 = 0;                // its first queue position
  unsigned* ldone = nullptr;      // its finished-cell counter (MB, may be null)
  int bi = 0;                     // its index (MB)
  unsigned gate_seen = 0;         // input chunks this wave has seen land (GATE)
  const unsigned qtot = MB ? qn : (unsigned)a.nCell;
  // a cell's results are out: count it for the caller's progress report
  auto finished = [&]() {
    if (MB && ldone)
      __hip_atomic_fetch_add(ldone, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  };

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
        const int leader = __ffsll((long long)mask) - 1;
        if (lane == leader) base = atomicAdd(queue, take);
        base = __shfl(base, leader);
        rsv = base;
        rsv_end = base + take;
        if (GATE && rsv < qtot && !gate_wait(gate, rsv, rsv_end < qtot ? rsv_end : qtot,
                                             gate_seen)) {
          // an input chunk never landed: drop the reservation, take no more cells
          if (lane == leader)
            __hip_atomic_fetch_or(gate.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          rsv = rsv_end;
          exhausted = true;
        }
      }
      const unsigned avail = rsv_end - rsv;
      const unsigned rank = (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
      if (need && rank < avail) {
        const unsigned q = rsv + rank;
        if (q < qtot) {
          if (MB) {  // queue positions only grow, so does the batch index
            while (bi + 1 < nb && q >= descs[bi + 1].q0) ++bi;
            la = descs[bi].a;
            q0 = descs[bi].q0;
            ldone = descs[bi].done;
          }
          cell = (long)la.Perm[q - q0] - 1;
          for (int k = 0; k < 8; ++k) ist[k] = 0;
          for (int k = 0; k < 3; ++k) rst[k] = 0.0;
          gather<M>(la, cell, b);
          const int e = decode_opts(la.ICNTRL + cell * 20, la.RCNTRL + cell * 20, la.Tstart,
                                    la.Tend, o);
          if (e != 0) {
            scatter(la, cell, b, ist, rst, e);
            if (REC) rec.head[q0 + cell] = -1;
            finished();
            cell = -1;  // stays free: takes another cell
          } else {
            head = -1;
            Nsteps = 0;
            T = la.Tstart;
            H = dmin(dmax(fabs(o.Hmin), fabs(o.Hstart)), fabs(o.Hmax));
            if (fabs(H) <= 10.0 * ROUNDOFF) H = DELTAMIN;
            dir = (la.Tend >= la.Tstart) ? +1 : -1;
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
      if (rsv >= qtot && rsv >= rsv_end && phase == PH_FREE) exhausted = true;
    }
    if (!__any(phase != PH_FREE)) break;

    // ---- top of TimeLoop for lanes starting a step -----------------------
    if (phase == PH_STEP) {
      int ierr_now = 0;
      if (!time_left(dir, T, la.Tend)) ierr_now = 1;
      else if (Nsteps > o.max_steps) ierr_now = -6;
      else if (((T + 0.1 * H) == T) || (H <= ROUNDOFF)) ierr_now = -7;
      if (ierr_now != 0) {
        scatter(la, cell, b, ist, rst, ierr_now);
        if (REC) rec.head[q0 + cell] = head;
        finished();
        phase = PH_FREE;
        cell = -1;
      } else {
        H = dmin(H, fabs(la.Tend - T));
      }
    }
    if (phase == PH_STEP) {
      step_functions<M>(b, o, T, ist);
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
      const int sing = jac_lu<M>(b, ghinv);
      ist[5] += 1;
      if (sing) {
        ist[7] += 1;
        ncons += 1;
        if (ncons <= 5) {
          H = H * 0.5;
          phase = PH_SING;
        } else {
          scatter(la, cell, b, ist, rst, -8);
          if (REC) rec.head[q0 + cell] = head;
          finished();
          phase = PH_FREE;
          cell = -1;
        }
      } else {
        factored = true;
      }
    }
    bool accepted = false;
    double Hnew = 0.0;
    if (factored) {
      const double Err = stages_and_error<M>(b, o, Dir, H, la.ATOL, la.RTOL, ist);
      const double Fac = dmin(o.FacMax, dmax(o.FacMin, o.FacSafe / gc_pow_third(Err)));
      Hnew = H * Fac;
      Nsteps += 1;
      ist[2] += 1;
      accepted = (Err <= 1.0) || (H <= o.Hmin);
    }
    if (REC) record_step(b, rec, accepted, ist[3], H, lane, head);
    if (factored) {
      if (accepted) {
        ist[3] += 1;
        accept_y(b, o);
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

}  // namespace

extern "C" __global__ void
__attribute__((amdgpu_flat_work_group_size(64, 256), amdgpu_waves_per_eu(GC_WPE, GC_WPE)))
GC_SFX(gc_rodas3_persistent)(GcBatchArgs a, double* ws, unsigned S, unsigned* queue,
                             int cells_per_wave, int grain) {
  persistent_body<false>(a, ws, S, queue, cells_per_wave, grain, GcRecord{});
}

// The persistent kernel behind the input gate of the overlapped upload.
extern "C" __global__ void
__attribute__((amdgpu_flat_work_group_size(64, 256), amdgpu_waves_per_eu(GC_WPE, GC_WPE)))
GC_SFX(gc_rodas3_persistent_gated)(GcBatchArgs a, double* ws, unsigned S, unsigned* queue,
                                   int cells_per_wave, int grain, GcGate gate) {
  persistent_body<false, false, false, true>(a, ws, S, queue, cells_per_wave, grain, GcRecord{},
                                             nullptr, 0, 0, gate);
}

// One persistent launch over nb batches (their own arrays, sizes, launch
// orders and time spans) in one queue of qn positions.
extern "C" __global__ void
__attribute__((amdgpu_flat_work_group_size(64, 256), amdgpu_waves_per_eu(GC_WPE, GC_WPE)))
GC_SFX(gc_rodas3_persistent_multi)(const GcBatchDesc* descs, int nb, unsigned qn, double* ws,
                                   unsigned S, unsigned* queue, int cells_per_wave, int grain) {
  persistent_body<false, false, true>(GcBatchArgs{}, ws, S, queue, cells_per_wave, grain,
                                      GcRecord{}, descs, nb, qn);
}

// Its recording form (heads of the record chains at head[q0 + cell]).
extern "C" __global__ void
__attribute__((amdgpu_flat_work_group_size(64, 256), amdgpu_waves_per_eu(GC_WPE, GC_WPE)))
GC_SFX(gc_rodas3_record_multi)(const GcBatchDesc* descs, int nb, unsigned qn, double* ws,
                               unsigned S, unsigned* queue, int cells_per_wave, int grain,
                               GcRecord rec) {
  persistent_body<true, false, true>(GcBatchArgs{}, ws, S, queue, cells_per_wave, grain, rec,
                                     descs, nb, qn);
}

// The persistent kernel with the cells' auto-reduce masks.
extern "C" __global__ void
__attribute__((amdgpu_flat_work_group_size(64, 256), amdgpu_waves_per_eu(GC_WPE, GC_WPE)))
GC_SFX(gc_rodas3_persistent_masked)(GcBatchArgs a, double* ws, unsigned S, unsigned* queue,
                                    int cells_per_wave, int grain) {
  persistent_body<false, true>(a, ws, S, queue, cells_per_wave, grain, GcRecord{});
}

// Recording form: also keeps every accepted step's start state and H for the
// discrete adjoint.
extern "C" __global__ void
__attribute__((amdgpu_flat_work_group_size(64, 256), amdgpu_waves_per_eu(GC_WPE, GC_WPE)))
GC_SFX(gc_rodas3_record)(GcBatchArgs a, double* ws, unsigned S, unsigned* queue,
                         int cells_per_wave, int grain, GcRecord rec) {
  persistent_body<true>(a, ws, S, queue, cells_per_wave, grain, rec);
}

// ---------------------------------------------------------------------------
// Discrete adjoint of the accepted Rodas3 steps (step sizes held fixed)
// ---------------------------------------------------------------------------
namespace {

constexpr int O_TMP = MECH_E_TMP * 512;
constexpr int O_LAM = MECH_E_LAM * 512;
constexpr int O_LK = MECH_E_LK * 512;
constexpr int O_GY = MECH_E_GY * 512;
constexpr int O_LP = MECH_E_LP * 512;
constexpr int O_DIR = MECH_E_DIR * 512;
constexpr int O_GK = MECH_E_GK * 512;
constexpr int O_GF = MECH_E_GF * 512;

// KppSolveTR (gckpp_LinearAlgebra.F90:2327-3701) in place on TMP: U^T, then L^T.
__device__ void solve_tr(const Blk& b) {
  solve_rows<true>(b, tab(gc_tr1), MECH_NTR1);
  solve_rows<false>(b, tab(gc_tr2), MECH_NTR2);
}

// A term of a sum stream: the sum restarts at a first term (flags & 1) and is
// stored at dst when its last term (flags & 2) is added; word = dst | flags.
__device__ __forceinline__ void stream_term(const Blk& b, int word, double v, double& acc) {
  acc = (word & 1) ? v : acc + v;
  if (word & 2) b.st(word & ~511, acc);
}

// nu(r) = sum_i S(i,r) mu(i) for mu in TMP, S the net stoichiometry: one
// stream of terms {TMP(i), NU(r) | flags, coef} over all reactions, sixteen
// loads in flight per batch (the stream is a multiple of eight terms).
__device__ void adj_nu(const Blk& b) {
  ctab_t* t = tab(gc_adj_nu);
  double acc = 0.0;
  int q = 0;
  for (; q + 16 <= MECH_NT_NU; q += 16, t += 64) {
    i4 tt[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) tt[r] = ld4(t + 4 * r);
    double x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = b.ld(tt[r].x);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < 16; ++r) stream_term(b, tt[r].y, dbl(tt[r].z, tt[r].w) * x[r], acc);
  }
  for (; q < MECH_NT_NU; q += 8, t += 32) {
    i4 tt[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) tt[r] = ld4(t + 4 * r);
    double x[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) x[r] = b.ld(tt[r].x);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < 8; ++r) stream_term(b, tt[r].y, dbl(tt[r].z, tt[r].w) * x[r], acc);
  }
}

// Byte offset of the element index in bits sh..sh+15 of a table word.
__device__ __forceinline__ int eoff(int w, int sh) { return ((w >> sh) & 0xFFFF) * 512; }

// Derivatives of the rate monomials A_r = c k_r x1 [x2 [x3]] weighted by
// nu_r, W_r = c nu_r k_r (in A's elements, free during the adjoint step):
//  first order  : dL/dk_r += (nu c) prod x;          g_f += W prod_{h != f} x_h
//  second order : the gradient of nu_r (grad A_r . d) for the direction d in
//                 DIR: dL/dk_r += (nu c) sum_g prod_{h != g} x_h d_g,
//                 g_f += W d/dx_f (sum_g prod_{h != g} x_h d_g)
// with g_f the gradient target of factor f (LP for variables, GF for fixed
// species).  The k_r records have distinct targets; the g_f targets are
// gathered, each from its terms in reaction order.

// second order, records of element indices {GK | CN<<16, x1 | x2<<16,
// x3 | d1<<16, d2 | d3<<16} of NF factors, CN = c*nu: GK += CN*(sum_g prod_{h!=g} x_h d_g)
struct Gk2Ops {
  double g, cn, d1, x1, x2, d2, x3, d3;
};

template <int NF>
__device__ __forceinline__ Gk2Ops gk2_load(const Blk& b, i4 f) {
  Gk2Ops o;
  o.g = b.ld(eoff(f.x, 0));
  o.cn = b.ld(eoff(f.x, 16));
  o.d1 = b.ld(eoff(f.z, 16));
  if (NF >= 2) {
    o.x1 = b.ld(eoff(f.y, 0));
    o.x2 = b.ld(eoff(f.y, 16));
    o.d2 = b.ld(eoff(f.w, 0));
  }
  if (NF >= 3) {
    o.x3 = b.ld(eoff(f.z, 0));
    o.d3 = b.ld(eoff(f.w, 16));
  }
  return o;
}

template <int NF>
__device__ __forceinline__ double gk2_term(const Gk2Ops& o) {
  if (NF == 1) return o.d1;
  if (NF == 2) return (o.x2 * o.d1) + (o.x1 * o.d2);
  return (((o.x2 * o.x3) * o.d1) + ((o.x1 * o.x3) * o.d2)) + ((o.x1 * o.x2) * o.d3);
}

template <int NF>
__device__ __forceinline__ ctab_t* gk2_group(const Blk& b, ctab_t* q, int n) {
  for (int r = 0; r < n; r += 8, q += 32) {
    i4 f[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) f[u] = ld4(q + 4 * u);
    Gk2Ops o[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) o[u] = gk2_load<NF>(b, f[u]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 8; ++u) b.st(eoff(f[u].x, 0), o[u].g + o[u].cn * gk2_term<NF>(o[u]));
  }
  return q;
}

__device__ void adj_gk2(const Blk& b) {
  ctab_t* q = tab(gc_adj_gk2);
  q = gk2_group<1>(b, q, gc_adjgrp_gk2[0]);
  q = gk2_group<2>(b, q, gc_adjgrp_gk2[1]);
  gk2_group<3>(b, q, gc_adjgrp_gk2[2]);
}

// Gather sums: one stream of terms {W | a<<16, b | c<<16, d, dst | flags}
// (W..d element indices) over all targets, each term W*((a*b)+(c*d)); the
// first term of a target is its initial value.  Sixteen terms per batch (the
// stream is a multiple of eight terms).
__device__ void adj_lp(const Blk& b, const int* table, int nterm) {
  ctab_t* t = tab(table);
  double acc = 0.0;
  int q = 0;
  for (; q + 16 <= nterm; q += 16, t += 64) {
    i4 tt[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) tt[r] = ld4(t + 4 * r);
    double w[16], xa[16], xb[16], xc[16], xd[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      w[r] = b.ld(eoff(tt[r].x, 0));
      xa[r] = b.ld(eoff(tt[r].x, 16));
      xb[r] = b.ld(eoff(tt[r].y, 0));
      xc[r] = b.ld(eoff(tt[r].y, 16));
      xd[r] = b.ld(eoff(tt[r].z, 0));
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < 16; ++r)
      stream_term(b, tt[r].w, w[r] * ((xa[r] * xb[r]) + (xc[r] * xd[r])), acc);
  }
  for (; q < nterm; q += 8, t += 32) {
    i4 tt[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) tt[r] = ld4(t + 4 * r);
    double w[8], xa[8], xb[8], xc[8], xd[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      w[r] = b.ld(eoff(tt[r].x, 0));
      xa[r] = b.ld(eoff(tt[r].x, 16));
      xb[r] = b.ld(eoff(tt[r].y, 0));
      xc[r] = b.ld(eoff(tt[r].y, 16));
      xd[r] = b.ld(eoff(tt[r].z, 0));
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < 8; ++r)
      stream_term(b, tt[r].w, w[r] * ((xa[r] * xb[r]) + (xc[r] * xd[r])), acc);
  }
}

// W(r) = (c*nu)*k into A(r) with the first-order dL/dk GK(r) += (c*nu)*((x1*x2)*x3),
// leaving c*nu in NU(r): records of element indices {A | GK<<16, NU | K<<16,
// x1 | x2<<16, x3 | lit<<16}, eight per batch (distinct targets but SINK).
__device__ void adj_wk(const Blk& b, const int* table) {
  ctab_t* t = tab(table);
  for (int q = 0; q < MECH_NWK; q += 8, t += 32) {
    i4 tt[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) tt[r] = ld4(t + 4 * r);
    double g[8], nu[8], k[8], x1[8], x2[8], x3[8], lit[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      g[r] = b.ld(eoff(tt[r].x, 16));
      nu[r] = b.ld(eoff(tt[r].y, 0));
      k[r] = b.ld(eoff(tt[r].y, 16));
      x1[r] = b.ld(eoff(tt[r].z, 0));
      x2[r] = b.ld(eoff(tt[r].z, 16));
      x3[r] = b.ld(eoff(tt[r].w, 0));
      lit[r] = ((clit_t*)gc_lits)[(tt[r].w >> 16) & 0xFFFF];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const double cn = lit[r] * nu[r];
      b.st(eoff(tt[r].x, 0), cn * k[r]);
      b.st(eoff(tt[r].x, 16), g[r] + cn * ((x1[r] * x2[r]) * x3[r]));
      b.st(eoff(tt[r].y, 0), cn);  // c*nu for the second-order records
    }
  }
}

// GY += LP and lambda_Kj += C_sj/(Dir*H) mu_s (+ A_sj LP for s >= 2), j < S.
template <int S>
__device__ __forceinline__ void adj_update(const Blk& b, double dh) {
  struct U {
    double lp, mu, gy, lk[S > 0 ? S : 1];
  };
  vpass<NVAR>(
      [&](int k) {
        U u;
        u.lp = b.ld(O_LP + k * 512);
        u.mu = b.ld(O_TMP + k * 512);
        u.gy = b.ld(O_GY + k * 512);
#pragma unroll
        for (int j = 0; j < S; ++j) u.lk[j] = b.ld(O_LK + j * O_VEC + k * 512);
        return u;
      },
      [&](int k, const U& u) {
        b.st(O_GY + k * 512, u.gy + u.lp);
#pragma unroll
        for (int j = 0; j < S; ++j) {
          double v = u.lk[j] + (ros_C[S * (S - 1) / 2 + j] / dh) * u.mu;
          if (S >= 2) v = v + ros_A[S * (S - 1) / 2 + j] * u.lp;
          b.st(O_LK + j * O_VEC + k * 512, v);
        }
      });
}

// One accepted step in reverse.  On entry Y holds the step's start state, the
// LU factors of M = I/(gamma*Dir*H) - J(Y) and the stage vectors K_1..K_4 have
// been recomputed, LAM = dL/dY at the end of the step.  On exit LAM = dL/dY
// at its start; GF and GK have collected the step's terms.
__device__ void adj_step(const Blk& b, double Dir, double H, bool clamp) {
  const double dh = Dir * H;
  if (clamp) {  // Y' = max(Ynew, 0): no sensitivity where Ynew <= 0
    vpass<NVAR>([&](int k) { return ld_yk(b, k); },
                [&](int k, const YK& l) {
                  if (!(ynew_of(l) > 0.0)) b.st(O_LAM + k * 512, 0.0);
                });
  }
  vpass<NVAR>([&](int k) { return b.ld(O_LAM + k * 512); },
              [&](int k, double l) {
                for (int j = 0; j < 4; ++j) b.st(O_LK + j * O_VEC + k * 512, ros_M[j] * l);
                b.st(O_GY + k * 512, l);
              });
  for (int s = 3; s >= 0; --s) {
    // mu_s = M^-T lambda_Ks, nu = S^T mu_s
    vpass<NVAR>([&](int k) { return b.ld(O_LK + s * O_VEC + k * 512); },
                [&](int k, double v) { b.st(O_TMP + k * 512, v); });
    solve_tr(b);
    adj_nu(b);
    // first order at the stage point: J(P_s)^T mu_s, d/dk, d/dF
    if (s >= 2) stage_point(b, s);
    adj_wk(b, s >= 2 ? gc_adj_wk_yn : gc_adj_wk_y);
    if (s >= 2) adj_lp(b, gc_adj_lp1_yn, MECH_NT_LP1_YN);
    else adj_lp(b, gc_adj_lp1_y, MECH_NT_LP1_Y);
    switch (s) {
      case 3: adj_update<3>(b, dh); break;
      case 2: adj_update<2>(b, dh); break;
      case 1: adj_update<1>(b, dh); break;
      default: adj_update<0>(b, dh); break;
    }
    // second order: d/dY of mu_s^T J(Y) K_s
    vpass<NVAR>([&](int k) { return b.ld(O_K + s * O_VEC + k * 512); },
                [&](int k, double v) { b.st(O_DIR + k * 512, v); });
    adj_gk2(b);
    adj_lp(b, gc_adj_lp2, MECH_NT_LP2);
    struct L2 {
      double g, p;
    };
    vpass<NVAR>([&](int k) { return L2{b.ld(O_GY + k * 512), b.ld(O_LP + k * 512)}; },
                [&](int k, const L2& l) { b.st(O_GY + k * 512, l.g + l.p); });
  }
  vpass<NVAR>([&](int k) { return b.ld(O_GY + k * 512); },
              [&](int k, double v) { b.st(O_LAM + k * 512, v); });
}

}  // namespace

// One lane per cell in Perm order: walks the cell's records from the last
// to the first one segment at a time (the steps back to a checkpoint),
// replays the segment forward from its checkpoint to regain each step's start
// state and stage vectors, then, last step first, recomputes each step's
// factors and sweeps the adjoint through it.
namespace {

// The reverse sweep of one cell: launch position ic of batch `a`, whose cells
// start at position q0 (cell Perm[ic - q0] - 1 of the batch, its record chain
// head at head[q0 + cell]); b is the lane's workspace slot.
__device__ __forceinline__ void adjoint_cell(const GcAdjArgs& a, unsigned q0, unsigned ic,
                                             const Blk& b) {
  const long cell = (long)a.Perm[ic - q0] - 1;
  const double* c = a.C + cell * NSPEC;
  const double* rc = a.RCONST + cell * NREACT;
  const double* lam = a.lamC + cell * NSPEC;
  double* gc = a.gC + cell * NSPEC;
  double* gk = a.gRCONST + cell * NREACT;
  const int* icu = a.ICNTRL + cell * 20;
  // option merge as decode_opts: user values where non-zero
  const bool autonomous = icu[0] != 0;
  const bool clamp = icu[15] == 1;
  if (!autonomous) {  // the dFdT stage terms are not differentiated
    for (int i = 0; i < NSPEC; ++i) gc[i] = 0.0;
    for (int r = 0; r < NREACT; ++r) gk[r] = 0.0;
    a.IERR[cell] = -3;
    return;
  }
  for (int j = 0; j < NFIX; ++j) b.st(O_FIX + j * 512, c[NVAR + j]);
  vpass<NREACT>([&](int r) { return rc[r]; }, [&](int r, double v) { b.st(O_RCT + r * 512, v); });
  b.st(MECH_E_ZERO * 512, 0.0);
  b.st(MECH_E_ONE * 512, 1.0);
  b.st(MECH_E_NZERO * 512, -0.0);
  vpass<NVAR>([&](int i) { return lam[i]; }, [&](int i, double v) { b.st(O_LAM + i * 512, v); });
  for (int j = 0; j < NFIX; ++j) b.st(O_GF + j * 512, lam[NVAR + j]);
  for (int r = 0; r < NREACT; ++r) b.st(O_GK + r * 512, 0.0);
  const double Dir = (a.Tend >= a.Tstart) ? 1.0 : -1.0;
  int ist[8];
  int ierr = 1;
  // segment buffers after the adjoint block: start states, step sizes (last
  // step first) and stage vectors of the segment's steps
  const int m = a.every;
  const int O_SEG = MECH_E_ADJ_TOTAL * 512;
  const int O_SEGH = O_SEG + m * NVAR * 512;
  const int O_SEGK = O_SEGH + m * 512;
  unsigned walked = 0;
  int slot = a.head[q0 + cell];
  while (slot >= 0 && ierr == 1) {
    // one segment, walked backwards: records up to and including a checkpoint
    int len = 0, ys = -1;
    for (;;) {
      if ((unsigned)slot >= a.cap || ++walked > a.cap || len >= m) {
        ierr = -5;  // not a chain of these pools
        break;
      }
      b.st(O_SEGH + len * 512, a.recH[slot]);
      ++len;
      ys = a.recYslot[slot];
      slot = a.recPrev[slot];
      if (ys >= 0) break;
      if (slot < 0) {
        ierr = -5;
        break;
      }
    }
    if (ierr != 1) break;
    if ((unsigned)ys >= a.ycap) {
      ierr = -5;
      break;
    }
    // replay forward from the checkpoint (bit-identical to the recorded run),
    // keeping the start state and the stage vectors of each of the segment's
    // steps; the checkpoint step's stage vectors come from the recording when
    // it kept them
    const double* y0 = a.recY + (size_t)ys * NVAR;
    const double* kk = a.recK ? a.recK + (size_t)ys * 4 * NVAR : nullptr;
    vpass<NVAR>([&](int i) { return y0[i]; }, [&](int i, double v) { b.st(O_Y + i * 512, v); });
    for (int q = 0; q < len; ++q) {
      vpass<NVAR>([&](int i) { return b.ld(O_Y + i * 512); },
                  [&](int i, double v) { b.st(O_SEG + (q * NVAR + i) * 512, v); });
      const double H = b.ld(O_SEGH + (len - 1 - q) * 512);
      if (q == 0 && kk) {
        if (len == 1) break;
        vpass<4 * NVAR>([&](int k) { return kk[k]; },
                        [&](int k, double v) { b.st(O_K + k * 512, v); });
      } else {
        fun<false>(b, false, O_FCN0);
        if (jac_lu<false>(b, 1.0 / ((Dir * H) * ros_Gamma[0]))) {
          ierr = -8;
          break;
        }
        stages<false>(b, true, Dir, H, ist);
        if (q == len - 1) break;  // its factors and stages stay in place
        vpass<4 * NVAR>([&](int k) { return b.ld(O_K + k * 512); },
                        [&](int k, double v) { b.st(O_SEGK + (q * 4 * NVAR + k) * 512, v); });
      }
      vpass<NVAR>([&](int k) { return ld_yk(b, k); },  // Ynew and the accepted update
                  [&](int k, const YK& l) {
                    const double yn = ynew_of(l);
                    b.st(O_Y + k * 512, clamp ? dmax(yn, 0.0) : yn);
                  });
    }
    if (ierr != 1) break;
    // the adjoint through the segment's steps, last first: each step's start
    // state and stage vectors back in place and its factors recomputed
    // (the last step's are still in place from the replay)
    for (int q = len - 1; q >= 0; --q) {
      const double H = b.ld(O_SEGH + (len - 1 - q) * 512);
      if (q != len - 1 || (q == 0 && kk)) {
        vpass<NVAR>([&](int i) { return b.ld(O_SEG + (q * NVAR + i) * 512); },
                    [&](int i, double v) { b.st(O_Y + i * 512, v); });
        if (q == 0 && kk)
          vpass<4 * NVAR>([&](int k) { return kk[k]; },
                          [&](int k, double v) { b.st(O_K + k * 512, v); });
        else
          vpass<4 * NVAR>([&](int k) { return b.ld(O_SEGK + (q * 4 * NVAR + k) * 512); },
                          [&](int k, double v) { b.st(O_K + k * 512, v); });
        if (jac_lu<false>(b, 1.0 / ((Dir * H) * ros_Gamma[0]))) {
          ierr = -8;
          break;
        }
      }
      adj_step(b, Dir, H, clamp);
    }
  }
  vpass<NVAR>([&](int i) { return b.ld(O_LAM + i * 512); }, [&](int i, double v) { gc[i] = v; });
  for (int j = 0; j < NFIX; ++j) gc[NVAR + j] = b.ld(O_GF + j * 512);
  vpass<NREACT>([&](int r) { return b.ld(O_GK + r * 512); }, [&](int r, double v) { gk[r] = v; });
  a.IERR[cell] = ierr;
}

}  // namespace

// The sweep over nb batches (GcAdjDesc, qn cells in all) as a persistent
// grid: each wave owns one workspace slot block and takes 64 consecutive
// launch positions at a time from the queue counter (zeroed by the caller)
// until the positions run out, so the workspace is bounded by the resident
// waves, not by the cells.  descs[k].done (host-visible, or null) counts the
// finished cells of batch k.
extern "C" __global__ void
__attribute__((amdgpu_flat_work_group_size(64, 256), amdgpu_waves_per_eu(GC_WPE, GC_WPE)))
GC_SFX(gc_rodas3_adjoint)(const GcAdjDesc* descs, int nb, unsigned qn, double* ws,
                          long slot_doubles, unsigned* queue) {
  extern __shared__ double lds[];
  const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const Blk b = make_blk(ws, wave, lane, lds + (threadIdx.x >> 6) * ((GC_SEG + 1) * 64),
                         (int)slot_doubles);
  const unsigned groups = (qn + 63u) >> 6;
  for (;;) {
    unsigned g = 0;
    if (lane == 0) g = atomicAdd(queue, 1u);
    g = (unsigned)__builtin_amdgcn_readfirstlane((int)g);
    if (g >= groups) break;
    const unsigned ic = g * 64u + (unsigned)lane;
    if (ic < qn) {
      int bi = 0;
      while (bi + 1 < nb && ic >= descs[bi + 1].q0) ++bi;
      adjoint_cell(descs[bi].a, descs[bi].q0, ic, b);
      if (descs[bi].done)
        __hip_atomic_fetch_add(descs[bi].done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// ---------------------------------------------------------------------------
// host launch wrappers of this variant
// ---------------------------------------------------------------------------
extern "C" GC_INTERNAL hipError_t GC_SFX(gc_launch_plain)(const GcBatchArgs* a, double* ws, int block,
                                               hipStream_t st) {
  if (a->nCell <= 0) return hipSuccess;
  const unsigned grid = ((unsigned)a->nCell + (unsigned)block - 1u) / (unsigned)block;
  const size_t lds = (size_t)(block / 64) * (GC_SEG + 1) * 64 * sizeof(double);
  if (a->doFun || a->doJvs || a->doSlv)
    hipLaunchKernelGGL(GC_SFX(gc_rodas3_plain_masked), dim3(grid), dim3(block), lds, st, *a, ws);
  else
    hipLaunchKernelGGL(GC_SFX(gc_rodas3_plain), dim3(grid), dim3(block), lds, st, *a, ws);
  return hipGetLastError();
}

extern "C" GC_INTERNAL hipError_t GC_SFX(gc_launch_persistent)(const GcBatchArgs* a, double* ws, unsigned S,
                                                    unsigned* queue, int block,
                                                    int cells_per_wave, int grain,
                                                    hipStream_t st) {
  if (a->nCell <= 0) return hipSuccess;
  const unsigned grid = S / (unsigned)block;
  const size_t lds = (size_t)(block / 64) * (GC_SEG + 1) * 64 * sizeof(double);
  if (a->doFun || a->doJvs || a->doSlv)
    hipLaunchKernelGGL(GC_SFX(gc_rodas3_persistent_masked), dim3(grid), dim3(block), lds, st, *a,
                       ws, S, queue, cells_per_wave, grain);
  else
    hipLaunchKernelGGL(GC_SFX(gc_rodas3_persistent), dim3(grid), dim3(block), lds, st, *a, ws, S,
                       queue, cells_per_wave, grain);
  return hipGetLastError();
}

extern "C" GC_INTERNAL hipError_t GC_SFX(gc_launch_persistent_gated)(
    const GcBatchArgs* a, double* ws, unsigned S, unsigned* queue, int block, int cells_per_wave,
    int grain, const GcGate* gate, hipStream_t st) {
  if (a->nCell <= 0) return hipSuccess;
  const unsigned grid = S / (unsigned)block;
  const size_t lds = (size_t)(block / 64) * (GC_SEG + 1) * 64 * sizeof(double);
  hipLaunchKernelGGL(GC_SFX(gc_rodas3_persistent_gated), dim3(grid), dim3(block), lds, st, *a, ws,
                     S, queue, cells_per_wave, grain, *gate);
  return hipGetLastError();
}

extern "C" GC_INTERNAL hipError_t GC_SFX(gc_launch_record)(const GcBatchArgs* a, double* ws,
                                                         unsigned S, unsigned* queue, int block,
                                                         int cells_per_wave, int grain,
                                                         const GcRecord* rec, hipStream_t st) {
  if (a->nCell <= 0) return hipSuccess;
  const unsigned grid = S / (unsigned)block;
  const size_t lds = (size_t)(block / 64) * (GC_SEG + 1) * 64 * sizeof(double);
  hipLaunchKernelGGL(GC_SFX(gc_rodas3_record), dim3(grid), dim3(block), lds, st, *a, ws, S,
                     queue, cells_per_wave, grain, *rec);
  return hipGetLastError();
}

extern "C" GC_INTERNAL hipError_t GC_SFX(gc_launch_persistent_multi)(
    const GcBatchDesc* descs, int nb, unsigned qn, double* ws, unsigned S, unsigned* queue,
    int block, int cells_per_wave, int grain, const GcRecord* rec, hipStream_t st) {
  if (qn == 0) return hipSuccess;
  const unsigned grid = S / (unsigned)block;
  const size_t lds = (size_t)(block / 64) * (GC_SEG + 1) * 64 * sizeof(double);
  if (rec)
    hipLaunchKernelGGL(GC_SFX(gc_rodas3_record_multi), dim3(grid), dim3(block), lds, st, descs,
                       nb, qn, ws, S, queue, cells_per_wave, grain, *rec);
  else
    hipLaunchKernelGGL(GC_SFX(gc_rodas3_persistent_multi), dim3(grid), dim3(block), lds, st,
                       descs, nb, qn, ws, S, queue, cells_per_wave, grain);
  return hipGetLastError();
}

extern "C" GC_INTERNAL hipError_t GC_SFX(gc_launch_adjoint)(const GcAdjDesc* descs, int nb,
                                                          unsigned qn, double* ws,
                                                          long slot_doubles, unsigned waves,
                                                          unsigned* queue, int block,
                                                          hipStream_t st) {
  if (qn == 0) return hipSuccess;
  const unsigned wpb = (unsigned)block / 64u;
  const unsigned grid = (waves + wpb - 1u) / wpb;
  const size_t lds = (size_t)wpb * (GC_SEG + 1) * 64 * sizeof(double);
  hipLaunchKernelGGL(GC_SFX(gc_rodas3_adjoint), dim3(grid), dim3(block), lds, st, descs, nb, qn,
                     ws, slot_doubles, queue);
  return hipGetLastError();
}

extern "C" GC_INTERNAL void GC_SFX(gc_variant_info)(GcVariantInfo* d) {
  d->nvar = NVAR;
  d->nfix = NFIX;
  d->nreact = NREACT;
  d->nnz = MECH_NNZ;
  d->ws_doubles_per_slot = E_TOTAL;
  d->adj_doubles_per_slot = MECH_E_ADJ_TOTAL;
  d->waves_per_eu = GC_WPE;
  d->seg = GC_SEG;
}
