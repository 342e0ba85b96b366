// wt_internal.h -- shared definitions of the gfx950 WaveRNN kernels and the
// host runtime of libwavetorch_hip.so.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

#include "../../include/wavetorch_hip.h"

namespace wt {

// ---------------------------------------------------------------- tables
//
// Special pixels: every pixel that carries a source and/or a probe tap,
// deduplicated, sorted by (x, y).  One row of 8 int32:
//   [0] x  [1] y  [2] source multiplicity (number of sources containing it)
//   [3] source slot (index among pixels with multiplicity > 0, or -1)
//   [4] tap_lo [5] tap_hi  (range in the sorted tap table)  [6..7] unused
constexpr int SPX_W = 8;
// Sorted taps: (x, y, kind, column) -- column = output index in y[b, n, :].
constexpr int TAP_W = 4;
// Tile descriptor: one row of 16 int32 (domain coordinates):
//   [0] R0 [1] RH      region rows  [R0, R0+RH)   (output rows + halo, clipped)
//   [2] C0 [3] RW      region cols  [C0, C0+RW)
//   [4] o_r0 [5] o_r1  output rows  [o_r0, o_r1)
//   [6] o_c0 [7] o_c1  output cols  [o_c0, o_c1)
//   [8] sp_lo [9] sp_hi  special pixels with x in region rows
//   [10] tp_lo [11] tp_hi taps with x in output rows
//   [12] RWm [13] CS   column-segment kernels: thread-mapping width and LDS
//                      column stride (cs_layout; 0 = RW and the shortest)
//   [14..15] unused
constexpr int TILE_W = 16;

// LDS bank mapping of the field buffers.  A wave's ds_read_b32 / ds_write_b32
// is served 32 lanes per cycle from 32 four-byte banks (bank = dword index
// mod 32; rocprofv3's SQ_LDS_BANK_CONFLICT agrees on these kernels: a region
// 96 columns wide counts ~0.07 conflict cycles per LDS instruction, one 107
// wide ~0.9).  Thread t owns (segment k, column c) = (t / RWm, t % RWm),
// whose cells start at LDS index (c+1)*CS + k*R + 1.  The 32 lanes of a
// half-wave hold consecutive columns except at one wrap from (k, RWm-1) to
// (k+1, 0); counting the lanes after the wrap as columns RWm, RWm+1, ... of
// segment k moves their index by R - RWm*CS.  With CS odd and
// RWm*CS == R (mod 32) that move is a multiple of 32 words, so the 32 lanes
// fall in 32 distinct banks (the 16-byte material reads of the nonlinear
// adjoint, 8 lanes per cycle, then too); a multiple of 32 for RWm puts every
// wrap between half-waves.  cs_layout takes the smallest RWm >= RW (columns
// RW..RWm-1 idle) and then the smallest CS >= the shortest odd stride that
// satisfy either within `maxthr` threads and cs_lds_slack(R) extra words per
// column; failing that, RWm = RW and the shortest stride.
__host__ __device__ inline int cs_lds_slack(int R) {
  int v = 0;  // 2-adic valuation of R: CS is then fixed modulo 32 >> v
  while (v < 4 && ((R >> v) & 1) == 0) ++v;
  const int s = (32 >> v) - 2;
  return s < 14 ? s : 14;
}
__host__ __device__ inline void cs_layout(int RH, int RW, int R, int maxthr, int &RWm, int &CS) {
  const int nseg = (RH + R - 1) / R;
  const int cs0 = (nseg * R + 2) | 1;
  const int slack = cs_lds_slack(R);
  RWm = RW;
  CS = cs0;
  for (int d = 0; d < 16; ++d) {
    const int w = RW + d;
    if (w * nseg > maxthr) return;
    if ((w & 31) == 0) {  // every wrap falls between half-waves
      RWm = w;
      return;
    }
    for (int e = 0; e <= slack; e += 2)
      if (((w * (cs0 + e) - R) & 31) == 0) {
        RWm = w;
        CS = cs0 + e;
        return;
      }
  }
}

// Arguments of the fused (temporally blocked) forward kernel.
struct FwdArgs {
  int B, Nx, Ny, T;
  int ntiles;           // tiles per sample
  const int32_t *tiles; // (ntiles, TILE_W)
  const int32_t *spx;   // (nspx, SPX_W)
  const int32_t *taps;  // (ntap, TAP_W) sorted
  int ntap;
  // linear coefficients (Nx,Ny) or, nonlinear, the base fields
  const float *A, *Bm, *C2;                // linear
  const float *bgeom, *clin, *rho, *rhoc;  // nonlinear: rhoc = rho*c_nl
  float s2, dtm1, dtm2, t1, t4;
  float sat_b0, sat_uth, cmin, cmax;
  int nlf;  // nonlinear flags (bit0 saturable damping, bit1 Kerr, bit2 clamp)
  // state in/out (B,Nx,Ny): u_n, u_{n-1}
  const float *in1, *in2;
  float *out1, *out2;
  int n0, nsteps;
  const float *x;  // (B,T)
  float *y;        // (B,T,ntap) or null
  float *yraw;     // (B,T,ntap) raw amplitudes for the adjoint, or null
  float *fields;   // (B,T/fstride,Nx,Ny) or null
  int fstride;
  // record mode: lap(u_n) (and u_n for the nonlinear cell) of the steps of
  // this launch into the segment buffer, domain-major with row stride Ny+1:
  // rec[(b*K + (n - rec_n0)) * Nx*(Ny+1) + r*(Ny+1) + c]
  float *rec_lap, *rec_u;
  int rec_n0, rec_K;
  // or, when set, sample b's steps start at step slot rec_soff[b] instead of
  // b*K (a record packed by each sample's own length)
  const int32_t *rec_soff;
  // batched instances: per-sample step count (a sample's work items skip
  // launches past its end; NULL = all samples run T steps) and the stride of
  // the C2 field between samples (0 = one field shared by the batch)
  const int32_t *Tlim;
  int c2_bstride;
};

// Arguments of the fused adjoint kernel (row strips only: C0 = 0, RW = Ny).
struct AdjArgs {
  int B, Nx, Ny, T;
  int ntiles;
  const int32_t *tiles;
  const int32_t *spx;
  const int32_t *taps;
  int ntap;
  int nsrcslot;
  const float *A, *Bm, *C2;
  const float *bgeom, *clin, *rho, *rhoc;
  float s2, dtm1, dtm2, t1, t4;
  float sat_b0, sat_uth, cmin, cmax, c_nl;
  int nlf;
  // adjoint state: lam_{m+1}, lam_{m+2} in; lam_{m_lo}, lam_{m_lo+1} out
  const float *in1, *in2;
  float *out1, *out2;
  int m_hi, nsteps;  // this launch forms lambda_m for m = m_hi .. m_hi-nsteps+1
  const float *ybar; // (B,T,ntap) or null
  const float *yraw; // (B,T,ntap)
  const float *seg_lap, *seg_u;  // segment buffer written by the record pass
  int seg_n0, seg_K;             // seg holds steps n = seg_n0 .. seg_n0+K-1
  const int32_t *seg_soff;       // per-sample first step slot, or NULL (b*K)
  float *gacc;       // (B,Nx,Ny) running sum of lap(u_{m-1})*lambda_m (linear)
  float *gacc_b, *gacc_rho;  // nonlinear accumulators (B,Nx,Ny)
  float *gsrc;       // (B,T,nsrcslot) lambda_m at source pixels
  const int32_t *Tlim;  // per-sample step count, or NULL
  int c2_bstride;       // C2 stride between samples (0 = shared)
};

// One launch over several configurations (batched instances of uneven grid
// size and length): work items [start[g], start[g+1]) run grp[g], each with
// its own tiles, tables, buffers and sizes; the kernel variant (threads, rows
// per thread) and the steps of the launch are common.  Passed by value (the
// whole struct is kernel arguments, well under the 4 KB limit), so a
// captured graph holds every launch's arguments itself.
constexpr int WT_MAX_GROUPS = 8;
struct BatchFwd {
  int ngrp;
  int32_t start[WT_MAX_GROUPS + 1];
  FwdArgs grp[WT_MAX_GROUPS];
};
struct BatchAdj {
  int ngrp;
  int32_t start[WT_MAX_GROUPS + 1];
  AdjArgs grp[WT_MAX_GROUPS];
};
static_assert(sizeof(BatchAdj) < 4000, "batched launch arguments exceed the kernarg limit");

// Carried state of the temporally blocked nonlinear adjoint between launches
// (B,Nx,Ny each): w_m = A c^2, P_m, a2_{m+1} and a2_m at the first step m of
// the next launch, and u_{n0-1} of the current segment.
struct NlCarry {
  const float *w_in, *p_in, *a2a_in, *a2b_in;
  float *w_out, *p_out, *a2a_out, *a2b_out;
  const float *uprev0;
};

// Extra buffers of the nonlinear adjoint (one launch per step): products of
// lambda_m with step m-1 factors handed to the next iterations, and where
// u_{n0-1} of the current segment lives.
struct NlAdjBufs {
  const float *W_in;   // w_m * lambda_{m+1}            (B,Nx,Ny)
  float *W_out;        // w_{m-1} * lambda_m            (B,Nx,Ny)
  float *PL;           // P_m lambda_{m+1} in, P_{m-1} lambda_m out (in place)
  float *AL;           // A Bm lambda of this parity slot (in place)
  const float *uprev0; // u_{seg_n0 - 1}                (B,Nx,Ny)
  int rec_ld;          // row stride of the segment record
  int64_t rec_slice;   // stride of one (sample, step) slice of the record
};

// Read-only view of a plan's configuration and device tables, for code
// outside the runtime (the double-precision loop).
struct PlanView {
  int B, T, Nx, Ny;
  int nlf;  // nonlinear flags as in FwdArgs (bit2 also for a clamp-only cell)
  int nspx, ntap, nslot;
  const int32_t *spx, *taps, *spmap, *mult;
};
int plan_view(const void *plan, PlanView *out);

// Optional phase timestamps (diagnostic builds compiled with
// -DWT_PHASE_TIMING): thread 0 of every workgroup records the 100 MHz wall
// clock at kernel start, after set-up, after the step loop and at the end
// into phase_ts[(blockIdx.x) * 8 + i].
#ifdef WT_PHASE_TIMING
extern __device__ unsigned long long *wt_phase_ts;
#define WT_TS(i)                                                             \
  do {                                                                       \
    if (wt_phase_ts && threadIdx.x == 0)                                     \
      wt_phase_ts[blockIdx.x * 8 + (i)] = __builtin_amdgcn_s_memrealtime();  \
  } while (0)
#else
#define WT_TS(i) \
  do {           \
  } while (0)
#endif

// a / b for the nonlinear adjoints' coefficient re-formation (the blocked
// k_adj_cs_nl and the per-step k_loop_adj_nl_plain use it alike): the
// hardware reciprocal estimate (1 ulp) and one residual correction
// (Markstein, "IA-64 and Elementary Functions", 2000: q = a r, e = a - b q,
// q + e r), within about half an ulp of the IEEE quotient for the
// normal-range operands met there (1 + q^2, its square, dt^-2 + b dt^-1, the
// threshold uth).  Four instructions instead of the ten of the IEEE
// sequence; the forward kernels keep IEEE division (their results are
// compared bit for bit).
__device__ __forceinline__ float adj_div(float a, float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  const float q = a * r;
  const float e = __builtin_fmaf(-b, q, a);
  return __builtin_fmaf(e, r, q);
}

// XCD-aware work-item order.  Workgroups are dispatched round-robin over the
// 8 XCDs of the chip (workgroup i runs on XCD i % 8); remap so that
// consecutive work items (the strips of one sample, which exchange halos
// through L2 between launches) land on the same XCD and share its L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  if ((nb & 7) != 0) return bid;
  int per = nb >> 3;
  return (bid & 7) * per + (bid >> 3);
}

}  // namespace wt

// Host-side helpers implemented in wt_runtime.hip.
namespace wt {
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

// Largest dynamic LDS a workgroup may request on the current device (cached
// per device id below 64).
inline int max_dyn_lds_dev() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  const bool slot = dev >= 0 && dev < 64;
  int v = slot ? cache[dev].load(std::memory_order_relaxed) : 0;
  if (v > 0) return v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
      v <= 0)
    v = 64 * 1024;
  if (slot) cache[dev].store(v, std::memory_order_relaxed);
  return v;
}

// Lets kernel `kfn` request the dynamic LDS the current device leaves next
// to the kernel's static LDS.  The attribute is set once per device:
// `done` (one per kernel) holds a bit per device id below 64; higher ids set
// it on every call.
inline hipError_t allow_dyn_lds(const void *kfn, std::atomic<uint64_t> &done) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const uint64_t bit = (dev >= 0 && dev < 64) ? (uint64_t)1 << dev : 0;
  if (bit && (done.load(std::memory_order_acquire) & bit)) return hipSuccess;
  hipFuncAttributes at{};
  e = hipFuncGetAttributes(&at, kfn);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                          max_dyn_lds_dev() - (int)at.sharedSizeBytes);
  if (e == hipSuccess && bit) done.fetch_or(bit, std::memory_order_release);
  return e;
}
}  // namespace wt

#define WT_HIP(call)                                   \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return wt::hip_fail(_e, #call); \
  } while (0)
