// wt_colseg.hip -- column-segment, temporally blocked WaveRNN kernels (gfx950).
//
// Same contract as the flat kernels of wt_tiled.hip (one workgroup advances
// one region = tile + halo of one sample by `nsteps` steps, no inter-workgroup
// synchronisation), with a thread mapping that keeps vertical neighbours in
// registers:
//
//   thread t -> column c = t % RW, segment k = t / RW; it owns the R cells
//   (rows k*R .. k*R+R-1, column c) of the region.  u_n and u_{n-1} of those
//   cells, and their coefficients, live in VGPRs; up/down neighbours are the
//   thread's own registers except at the segment ends.
//
//   LDS holds the field column-major: cell (row, col) at (col+1)*CS + row+1,
//   CS odd (consecutive lanes = consecutive columns hit distinct banks), with
//   a zero column on each side and a zero row above/below.  Per cell and step
//   one LDS write (the new value) and two reads (left, right neighbours);
//   every LDS address is a per-thread base plus an immediate offset.
//
//   Halo cone: rows whose value can no longer reach the output rows before
//   the end of the launch are not computed (whole segments are skipped).
//
// Forward per cell, in the reference's operation order (cell.py:12-17):
//     s = up + lf;  s = fma(-4, ctr, s)  (== s + (-4*ctr): the product is
//     exact);  s = s + rt;  lap = s + dn
//     v = A*(((s2*ctr) - (Bm*u_prev)) + (C2*lap))   then + x_n at sources.
// Adjoint per cell (linear cell, row strips):
//     lam_m = a1*lam_{m+1} + L(w*lam_{m+1}) - a2*lam_{m+2} (+ probe seed)
//     gacc += lap(u_{m-1}) * lam_m
// Cells with a source pixel or a probe tap take a slower code path that only
// the waves containing them execute (wave-uniform branch).
#include "wt_internal.h"

namespace wt {

// Largest number of steps one launch of these kernels may take (the input
// samples of a launch are staged in LDS).
constexpr int CS_MAX_TBLOCK = 256;
// Largest number of special pixels / probe taps one tile may hold.
constexpr int CS_MAX_TABLE = 256;

struct CsGeom {
  int R0, RH, C0, RW;     // region
  int or0, or1, oc0, oc1; // output rows / cols (domain)
  int k, c;               // segment, column of this thread
  bool active;            // thread owns cells
  int CS;                 // LDS column stride (odd)
  int base;               // LDS index of this thread's first cell
};

// LAYOUT: take the thread-mapping width and LDS column stride the host chose
// (cs_layout, descriptor words 12 and 13; the plain mapping unless they fit
// this launch's rows per thread and threads) -- the forward kernels, whose
// LDS allocation holds the padded stride.  Otherwise (the adjoints, which
// wait little on LDS) the plain mapping t -> (t / RW, t % RW).
template <int R, bool LAYOUT = false>
__device__ __forceinline__ CsGeom cs_geom(const int32_t *td, int tid) {
  CsGeom g;
  g.R0 = td[0]; g.RH = td[1]; g.C0 = td[2]; g.RW = td[3];
  g.or0 = td[4]; g.or1 = td[5]; g.oc0 = td[6]; g.oc1 = td[7];
  const int nseg = (g.RH + R - 1) / R;
  if constexpr (LAYOUT) {
    int RWm = td[12], CS = td[13];
    if (RWm < g.RW || CS < nseg * R + 2 || CS > ((nseg * R + 2) | 1) + cs_lds_slack(R) ||
        RWm * nseg > (int)blockDim.x) {
      RWm = g.RW;
      CS = (nseg * R + 2) | 1;
    }
    g.CS = CS;
    g.k = tid / RWm;
    g.c = tid - g.k * RWm;
    g.active = g.k < nseg && g.c < g.RW;
  } else {
    g.k = tid / g.RW;
    g.c = tid - g.k * g.RW;
    g.active = g.k < nseg;
    g.CS = (nseg * R + 2) | 1;
  }
  g.base = (g.c + 1) * g.CS + g.k * R + 1;
  return g;
}

// Rows [lo, hi) of the region that still influence the output after step s
// of a launch of n steps.
__device__ __forceinline__ void cone(const CsGeom &g, int s, int n, int &lo, int &hi) {
  const int rem = n - 1 - s;
  const int olo = g.or0 - g.R0, ohi = g.or1 - g.R0;
  lo = olo > 0 ? max(0, olo - rem) : 0;
  hi = ohi < g.RH ? min(g.RH, ohi + rem) : g.RH;
}

// Raw buffer access with hardware range checking (gfx9 descriptor word 3 =
// 0x00020000): a load outside [0, num_records) bytes returns 0 and a store
// there is dropped.  The slices below put the buffer base at the first
// output row of the tile, so rows above the output rows give negative byte
// offsets (huge as unsigned) and rows below give offsets >= num_records:
// both fall outside the range without any per-row predicate.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t cs_rsrc(const float *p, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float cs_bload(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0));
}
__device__ __forceinline__ void cs_bstore(__amdgpu_buffer_rsrc_t rs, uint32_t off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs, (int)off, 0, 0);
}
// The segment record (lap u, u) is written once by the recompute and read
// once by the adjoint.  Cache policies of those accesses (the buffer
// instructions' cache-policy operand: 2 = nt, 16 = sc1).  Measured on MI355X
// for the training step: nt writes double the recompute's time (the record
// rows are narrower than a cache line); nt reads with default writes slow
// the adjoint by 8%; the default policy on both is fastest overall.
#ifndef WT_RECORD_LOAD_CPOL
#define WT_RECORD_LOAD_CPOL 0
#endif
#ifndef WT_RECORD_STORE_CPOL
#define WT_RECORD_STORE_CPOL 0
#endif
__device__ __forceinline__ float cs_rec_load(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, WT_RECORD_LOAD_CPOL));
}
__device__ __forceinline__ void cs_rec_store(__amdgpu_buffer_rsrc_t rs, uint32_t off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs, (int)off, 0,
                                        WT_RECORD_STORE_CPOL);
}

// Byte offset of this thread's first cell relative to the first output row
// of its tile (column C0 + c); an out-of-range marker if the thread's column
// is outside the tile's output columns.
__device__ __forceinline__ uint32_t cs_out_voff(const CsGeom &g, int row0, int Ny) {
  const int gc = g.C0 + g.c;
  if (!g.active || gc < g.oc0 || gc >= g.oc1) return 0x80000000u;
  return (uint32_t)(((g.R0 + row0 - g.or0) * Ny + gc) * 4);
}

// Zeroes LDS buffer b1 and, of b0, only the cells that the workgroup's fill
// of b0 (every cell of every segment of the region's columns) leaves alone:
// the ring columns 0 and RW+1 and, in the region's columns, the ring row 0
// and the rows past the last segment (CS - nrows cells per column).
__device__ __forceinline__ void cs_zero_lds(float *b0, float *b1, const CsGeom &g, int nrows,
                                            int NBUF) {
  for (int i = threadIdx.x; i < NBUF; i += blockDim.x) b1[i] = 0.0f;
  const int tail = g.CS - nrows;
  const int nring = 2 * g.CS + g.RW * tail;
  for (int i = threadIdx.x; i < nring; i += blockDim.x) {
    int idx;
    if (i < g.CS) {
      idx = i;
    } else if (i < 2 * g.CS) {
      idx = (g.RW + 1) * g.CS + (i - g.CS);
    } else {
      const int j = i - 2 * g.CS, col = j / tail, r = j - col * tail;
      idx = (col + 1) * g.CS + (r == 0 ? 0 : nrows + r);
    }
    b0[idx] = 0.0f;
  }
}

// ------------------------------------------------------------------ forward

// Writes the output tile of the field in LDS buffer Q to fields[(b,idx)]:
// a flat loop over the tile (used only at snapshot steps).
__device__ __forceinline__ void cs_snapshot(const CsGeom &g, const float *Q, float *fo, int Ny) {
  const int th = g.or1 - g.or0, tw = g.oc1 - g.oc0;
  for (int i = threadIdx.x; i < th * tw; i += blockDim.x) {
    const int rr = i / tw, cc = i - (i / tw) * tw;
    const int row = g.or0 - g.R0 + rr, col = g.oc0 - g.C0 + cc;
    fo[(int64_t)(g.or0 + rr) * Ny + (g.oc0 + cc)] = Q[(col + 1) * g.CS + row + 1];
  }
}

// One forward step of this thread's segment; REC records lap(u_n) (and u_n
// for the nonlinear cell) of the output rows through range-checked buffers.
template <int R, int NL, bool REC>
__device__ __forceinline__ void cs_fwd_rows(const FwdArgs &a, const float *__restrict__ P,
                                            float *__restrict__ Q,
                                            float (&Uc)[R], float (&Up)[R], const float (&KA)[R],
                                            const float (&KB)[R], const float (&KC)[R],
                                            const float (&KD)[R], const float (&KS)[R],
                                            int base, int CS, uint32_t vmask, float xv,
                                            __amdgpu_buffer_rsrc_t rsl, __amdgpu_buffer_rsrc_t rsu,
                                            uint32_t voff, int Ny) {
  const int lfb = base - CS, rtb = base + CS;
  float upv = P[base - 1];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const float ct = Uc[r];
    const float dn = (r < R - 1) ? Uc[r + 1] : P[base + R];
    const float lf = P[lfb + r], rt = P[rtb + r];
    float sacc = upv + lf;
    sacc = __builtin_fmaf(-4.0f, ct, sacc);
    sacc = sacc + rt;
    const float lap = sacc + dn;
    float A, Bm, C2;
    if (NL == 0) {
      A = KA[r]; Bm = KB[r]; C2 = KC[r];
    } else {
      float bb = KA[r], cc = KB[r];
      if (a.nlf & 1) {
        float q = ct / a.sat_uth;
        q = fabsf(q);
        q = q * q;
        q = q + 1.0f;
        q = a.sat_b0 / q;
        bb = bb + KC[r] * q;
      }
      if (a.nlf & 2) cc = cc + KD[r] * (ct * ct);
      if (a.nlf & 4) cc = fminf(fmaxf(cc, a.cmin), a.cmax);
      const float bd = bb * a.dtm1;
      A = 1.0f / (a.dtm2 + bd);
      Bm = a.dtm2 - bd;
      C2 = cc * cc;
      if (!((vmask >> r) & 1u)) A = 0.0f;
    }
    const float p1 = a.s2 * ct;
    const float p2 = Bm * Up[r];
    const float p3 = C2 * lap;
    float v = A * ((p1 - p2) + p3);
    // source: weight 1 at source pixels, 0 elsewhere; fma(1, x, v) rounds
    // once, exactly like the reference's v + x (source.py:22)
    v = __builtin_fmaf(KS[r], xv, v);
    upv = ct;
    Up[r] = v;
    Q[base + r] = v;
    if (REC) {
      const uint32_t o = voff + (uint32_t)(r * Ny * 4);
      cs_rec_store(rsl, o, lap);
      if (NL) cs_rec_store(rsu, o, ct);
    }
  }
}

#ifdef WT_PHASE_TIMING
__device__ unsigned long long *wt_phase_ts = nullptr;
#endif

// Work item w (sample w / ntiles, tile w % ntiles) of the forward launch
// described by `a`; shared by the single-configuration kernel and the
// batched kernel (one launch over several configurations; BATCH compiles in
// the per-sample lengths and record slots).
template <int NT, int R, int NL, bool REC, bool BATCH>
__device__ __forceinline__ void fwd_cs_body(const FwdArgs &a, const int w) {
  WT_TS(0);
  extern __shared__ float smem[];
  const int tid = threadIdx.x;
  const int b = w / a.ntiles, ti = w - (w / a.ntiles) * a.ntiles;
  // a sample whose own length ends before this launch has nothing to do (a
  // recording launch too: the adjoint reads no record of a step at or past
  // the sample's length)
  if (BATCH && a.Tlim && a.n0 >= a.Tlim[b]) return;
  const int32_t *td = a.tiles + ti * TILE_W;
  const CsGeom g = cs_geom<R, true>(td, tid);
  const int NBUF = (g.RW + 2) * g.CS;
  float *buf0 = smem, *buf1 = smem + NBUF;
  float *xs = smem + 2 * NBUF;  // the input samples of this launch
  const int Ny = a.Ny;
  const int64_t F = (int64_t)a.Nx * Ny;
  const int64_t boff = (int64_t)b * F;

  // tables of this tile staged in LDS (one parallel load, then scans hit
  // LDS): special pixels (x, y, multiplicity) and probe taps (x, y, kind, col)
  __shared__ int tb_sp[CS_MAX_TABLE * 3];
  __shared__ int tb_tp[CS_MAX_TABLE * 4];
  const int nsp = min(td[9] - td[8], CS_MAX_TABLE), ntp = min(td[11] - td[10], CS_MAX_TABLE);

  float U0[R], U1[R], KA[R], KB[R], KC[R], KD[R], KS[R];
  const int row0 = g.k * R;  // region row of cell 0
  uint32_t vmask = 0;        // cells that are inside the region
  // 1) all global loads first (one round trip for the whole set-up)
#pragma unroll
  for (int r = 0; r < R; ++r) {
    float u1 = 0.f, u2 = 0.f, v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f;
    if (g.active && row0 + r < g.RH) {
      vmask |= 1u << r;
      const int64_t q = (int64_t)(g.R0 + row0 + r) * Ny + (g.C0 + g.c);
      if (NL == 0) {
        v0 = a.A[q]; v1 = a.Bm[q]; v2 = a.C2[q + (int64_t)b * a.c2_bstride];
      } else {
        v0 = a.bgeom[q]; v1 = a.clin[q]; v2 = a.rho[q]; v3 = a.rhoc[q];
      }
      u1 = a.in1[boff + q];
      u2 = a.in2[boff + q];
    }
    U0[r] = u1; U1[r] = u2;
    KA[r] = v0; KB[r] = v1; KC[r] = v2; KD[r] = v3; KS[r] = 0.0f;
  }
  float xsv = 0.0f;
  if (tid < a.nsteps) xsv = a.x[(int64_t)b * a.T + a.n0 + tid];
  int spv0 = 0, spv1 = 0, spv2 = 0, tpv0 = 0, tpv1 = 0, tpv2 = 0, tpv3 = 0;
  if (tid < nsp) {
    const int32_t *e = a.spx + (td[8] + tid) * SPX_W;
    spv0 = e[0]; spv1 = e[1]; spv2 = e[2];
  }
  if (tid < ntp) {
    const int32_t *t = a.taps + (td[10] + tid) * TAP_W;
    tpv0 = t[0]; tpv1 = t[1]; tpv2 = t[2]; tpv3 = t[3];
  }
  // 2) LDS (buf0's region cells are filled from U0 below)
  cs_zero_lds(buf0, buf1, g, ((g.RH + R - 1) / R) * R, NBUF);
  WT_TS(4);
  if (tid < a.nsteps) xs[tid] = xsv;
  if (tid < nsp) { tb_sp[3 * tid] = spv0; tb_sp[3 * tid + 1] = spv1; tb_sp[3 * tid + 2] = spv2; }
  if (tid < ntp) {
    tb_tp[4 * tid] = tpv0; tb_tp[4 * tid + 1] = tpv1;
    tb_tp[4 * tid + 2] = tpv2; tb_tp[4 * tid + 3] = tpv3;
  }
  __syncthreads();
  WT_TS(5);
  if (g.active) {
#pragma unroll
    for (int r = 0; r < R; ++r) buf0[g.base + r] = U0[r];
  }
  // source pixels of this thread (multiplicity 1; the host routes pixels
  // shared by several sources to the other kernels)
  bool any_src = false;
  for (int kk = 0; kk < nsp; ++kk) {
    if (tb_sp[3 * kk + 2] <= 0) continue;
    const int rr = tb_sp[3 * kk] - g.R0 - row0, cc = tb_sp[3 * kk + 1] - g.C0;
    if (!g.active || cc != g.c || rr < 0 || rr >= R || row0 + rr >= g.RH) continue;
    any_src = true;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (r == rr) KS[r] = 1.0f;
  }
  const bool blk_src = __syncthreads_or(any_src);
  WT_TS(6);
  // probe tap handled by this thread (taps in the output region)
  int pidx = -1, pkind = 0, pcol = 0;
  {
    int kk = 0;
    for (int t = 0; t < ntp; ++t) {
      const int tx = tb_tp[4 * t], ty = tb_tp[4 * t + 1];
      if (ty < g.oc0 || ty >= g.oc1) continue;
      if (kk == tid) {
        pidx = (ty - g.C0 + 1) * g.CS + (tx - g.R0) + 1;
        pkind = tb_tp[4 * t + 2];
        pcol = tb_tp[4 * t + 3];
      }
      ++kk;
    }
  }
  const uint32_t voff = cs_out_voff(g, row0, Ny);
  const int64_t out_bytes = (int64_t)(g.or1 - g.or0) * Ny * 4;
  // record slots of this sample; steps at or past its length are not
  // recorded (empty slice: the stores are dropped)
  const int64_t rslot0 =
      (BATCH && a.rec_soff) ? (int64_t)a.rec_soff[b] : (int64_t)b * a.rec_K;
  const int Lb = (BATCH && a.Tlim) ? a.Tlim[b] : 0x7fffffff;
  __syncthreads();

  auto step = [&](const float *P, float *Q, float(&Uc)[R], float(&Up)[R], int s) {
    const int n = a.n0 + s;
    int lo, hi;
    cone(g, s, a.nsteps, lo, hi);
    const bool need = g.active && row0 < hi && row0 + R > lo;
    // the (sample, step) slice of the record, based at the first output row
    const int64_t ro = (rslot0 + (n - a.rec_n0)) * F + (int64_t)g.or0 * Ny;
    const int64_t rbytes = n < Lb ? out_bytes : 0;
    const __amdgpu_buffer_rsrc_t rsl = cs_rsrc(REC ? a.rec_lap + ro : a.in1, REC ? rbytes : 0);
    const __amdgpu_buffer_rsrc_t rsu =
        cs_rsrc((REC && NL) ? a.rec_u + ro : a.in1, (REC && NL) ? rbytes : 0);
    const float xv = blk_src ? xs[s] : 0.0f;
    if (need)
      cs_fwd_rows<R, NL, REC>(a, P, Q, Uc, Up, KA, KB, KC, KD, KS, g.base, g.CS, vmask, xv, rsl,
                              rsu, voff, Ny);
    __syncthreads();
    if (pidx >= 0) {
      const float v = Q[pidx];
      const int64_t o = ((int64_t)b * a.T + n) * a.ntap + pcol;
      if (a.y) a.y[o] = pkind ? v * v : v;
      if (a.yraw) a.yraw[o] = v;
    }
    if (a.fields && ((n + 1) % a.fstride) == 0) {
      const int64_t nf = a.T / a.fstride;
      cs_snapshot(g, Q, a.fields + ((int64_t)b * nf + (n + 1) / a.fstride - 1) * F, Ny);
    }
  };

  WT_TS(1);
  int s = 0;
  for (; s + 1 < a.nsteps; s += 2) {
    step(buf0, buf1, U0, U1, s);
    step(buf1, buf0, U1, U0, s + 1);
  }
  if (s < a.nsteps) step(buf0, buf1, U0, U1, s);
  WT_TS(2);
  // final state: after an even number of steps u_n is in U0, u_{n-1} in U1
  const bool odd = (a.nsteps & 1) != 0;
  if (g.active) {
    const int gc = g.C0 + g.c;
    if (gc >= g.oc0 && gc < g.oc1) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int gr = g.R0 + row0 + r;
        if (row0 + r < g.RH && gr >= g.or0 && gr < g.or1) {
          const int64_t q = boff + (int64_t)gr * Ny + gc;
          a.out1[q] = odd ? U1[r] : U0[r];
          a.out2[q] = odd ? U0[r] : U1[r];
        }
      }
    }
  }
  WT_TS(3);
}

template <int NT, int R, int NL, bool REC>
__global__ __launch_bounds__(NT) void k_fwd_cs(FwdArgs a) {
  fwd_cs_body<NT, R, NL, REC, false>(a, xcd_remap(blockIdx.x, gridDim.x));
}

// Group of a batched launch that work item w belongs to.
__device__ __forceinline__ int batch_group(const int32_t (&start)[WT_MAX_GROUPS + 1], int ngrp,
                                           int w) {
  int gi = 0;
  while (gi + 1 < ngrp && w >= start[gi + 1]) ++gi;
  return gi;
}

template <int NT, int R, int NL, bool REC>
__global__ __launch_bounds__(NT) void k_fwd_cs_batch(BatchFwd bl) {
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int gi = batch_group(bl.start, bl.ngrp, w);
  fwd_cs_body<NT, R, NL, REC, true>(bl.grp[gi], w - bl.start[gi]);
}

// ------------------------------------------------------------------ adjoint

// Probe-seed table: at most this many probe pixels per region.
constexpr int CS_MAX_SEEDS = 16;

// One reverse step of this thread's segment.  SP: the wave holds
// probe/source pixels (sr0, sr1).  a2 = A*Bm is formed as a1 - 1 (with
// a1 = A*s2 = 2/(1+b*dt), a2 = (1-b*dt)/(1+b*dt)).  The thread's own w*lambda
// values (up/centre/down neighbours) are re-formed as KW*Lc -- the same fp32
// product the previous step stored to LDS -- instead of being kept in R more
// registers.
template <int R, int NSP, bool GS>
__device__ __forceinline__ void cs_adj_rows(
    const float *__restrict__ P, float *__restrict__ Q, float (&Lc)[R], float (&Ln)[R],
    const float (&KA1)[R], const float (&KW)[R], int base, int CS, uint32_t omask, int sr0,
    int sr1, float gs0, float gs1, int sl0, int sl1, float *__restrict__ gsrc_row) {
  const int lfb = base - CS, rtb = base + CS;
  float upv = P[base - 1];
  float ct = KW[0] * Lc[0];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const float dn = (r < R - 1) ? KW[r + 1] * Lc[r + 1] : P[base + R];
    const float lf = P[lfb + r], rt = P[rtb + r];
    float lapw = upv + lf;
    lapw = __builtin_fmaf(-4.0f, ct, lapw);
    lapw = lapw + rt;
    lapw = lapw + dn;
    // a1*L1 + lapw - (a1-1)*L2 = a1*(L1 - L2) + (lapw + L2)
    float lam = __builtin_fmaf(KA1[r], Lc[r] - Ln[r], lapw + Ln[r]);
    if (NSP >= 1 && r == sr0) lam += gs0;
    if (NSP >= 2 && r == sr1) lam += gs1;
    upv = ct;
    ct = dn;
    Ln[r] = lam;
    Q[base + r] = KW[r] * lam;
  }
  if (GS) {
    // lambda_m at source pixels of the output rows -> dJ/dx[b, m-1]: pick
    // the rows' values with selects, then at most NSP stores (one masked
    // store per row would put R branches on every step of the waves that
    // hold special pixels, which the step's barrier waits for)
    float v0 = Ln[0], v1 = Ln[0];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      v0 = (r == sr0) ? Ln[r] : v0;
      if (NSP >= 2) v1 = (r == sr1) ? Ln[r] : v1;
    }
    if (sl0 >= 0 && sr0 >= 0 && ((omask >> sr0) & 1u)) gsrc_row[sl0] = v0;
    if (NSP >= 2 && sl1 >= 0 && sr1 >= 0 && ((omask >> sr1) & 1u)) gsrc_row[sl1] = v1;
  }
}

// Linear cell.  LDS holds w*lambda.  Work item w of the launch `a` (BATCH:
// the batched kernel, with per-sample lengths and record slots).
template <int NT, int R, bool BATCH>
__device__ __forceinline__ void adj_cs_body(const AdjArgs &a, const int w) {
  WT_TS(0);
  extern __shared__ float smem[];
  const int tid = threadIdx.x;
  const int b = w / a.ntiles, ti = w - (w / a.ntiles) * a.ntiles;
  // lambda_m = 0 for every m of this launch when the sample ends before it
  if (BATCH && a.Tlim && a.m_hi - a.nsteps + 1 > a.Tlim[b]) return;
  const int32_t *td = a.tiles + ti * TILE_W;
  const CsGeom g = cs_geom<R>(td, tid);
  const int NBUF = (g.RW + 2) * g.CS;
  float *buf0 = smem, *buf1 = smem + NBUF;
  float *seeds = smem + 2 * NBUF;  // (CS_MAX_SEEDS, nsteps) probe seeds
  __shared__ int seed_px[CS_MAX_SEEDS];
  __shared__ int nseeds;
  const int Ny = a.Ny;
  const int64_t F = (int64_t)a.Nx * Ny;
  const int64_t boff = (int64_t)b * F;
  const float s2 = a.s2;

  __shared__ int tb_sp[CS_MAX_TABLE * 6];  // x, y, mult, slot, tap_lo, tap_hi
  const int nsp = min(td[9] - td[8], CS_MAX_TABLE);

  float L1[R], L2[R], KA1[R], KW[R], G[R], SG0[R], SG1[R];
  const int row0 = g.k * R;
  // 1) all global loads first
  float KC2[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    float A = 0.f, C2 = 0.f, u1 = 0.f, u2 = 0.f;
    if (g.active && row0 + r < g.RH) {
      const int64_t q = (int64_t)(g.R0 + row0 + r) * Ny + (g.C0 + g.c);
      A = a.A[q]; C2 = a.C2[q + (int64_t)b * a.c2_bstride];
      u1 = a.in1[boff + q];
      u2 = a.in2[boff + q];
    }
    KA1[r] = A; KC2[r] = C2;
    L1[r] = u1; L2[r] = u2;
  }
  int e0 = 0, e1 = 0, e2 = 0, e3 = 0, e4 = 0, e5 = 0;
  if (tid < nsp) {
    const int32_t *e = a.spx + (td[8] + tid) * SPX_W;
    e0 = e[0]; e1 = e[1]; e2 = e[2]; e3 = e[3]; e4 = e[4]; e5 = e[5];
  }
  // 2) LDS (buf0's region cells are filled from the lambda pair below)
  cs_zero_lds(buf0, buf1, g, ((g.RH + R - 1) / R) * R, NBUF);
  if (tid < nsp) {
    int *t = tb_sp + 6 * tid;
    t[0] = e0; t[1] = e1; t[2] = e2; t[3] = e3; t[4] = e4; t[5] = e5;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const float A = KA1[r];
    KA1[r] = A * s2; KW[r] = A * KC2[r];
    G[r] = 0.0f; SG1[r] = 0.0f;

  }
  __syncthreads();
  if (tid == 0) {
    // probe pixels of the region, in table order
    int n = 0;
    for (int kk = 0; kk < nsp && n < CS_MAX_SEEDS; ++kk)
      if (tb_sp[6 * kk + 5] > tb_sp[6 * kk + 4]) seed_px[n++] = kk;
    nseeds = n;
  }
  if (g.active) {
#pragma unroll
    for (int r = 0; r < R; ++r) buf0[g.base + r] = KW[r] * L1[r];
  }
  __syncthreads();
  // seeds[k][s] = sum over taps of pixel k of dy/du * ybar at step m = m_hi - s
  for (int i = tid; i < nseeds * a.nsteps; i += blockDim.x) {
    const int k = i / a.nsteps, s = i - (i / a.nsteps) * a.nsteps;
    const int m = a.m_hi - s;
    const int *e = tb_sp + 6 * seed_px[k];
    float gs = 0.0f;
    if (a.ybar && (!a.Tlim || m - 1 < a.Tlim[b])) {
      for (int t = e[4]; t < e[5]; ++t) {
        const int32_t *tp = a.taps + t * TAP_W;
        const int64_t o = ((int64_t)b * a.T + (m - 1)) * a.ntap + tp[3];
        const float gy = a.ybar[o];
        gs += tp[2] ? 2.0f * a.yraw[o] * gy : gy;
      }
    }
    seeds[i] = gs;
  }
  // special rows of this thread: probe taps (seed slot) / source pixels (slot)
  int sr0 = -1, sr1 = -1, sd0 = -1, sd1 = -1, sl0 = -1, sl1 = -1;
  {
    int nseed = 0;
    for (int kk = 0; kk < nsp; ++kk) {
      const int *e = tb_sp + 6 * kk;
      const bool has_tap = e[5] > e[4];
      const int seed_slot = has_tap ? nseed : -1;
      if (has_tap) ++nseed;
      const int rr = e[0] - g.R0 - row0;
      if (!g.active || e[1] - g.C0 != g.c || rr < 0 || rr >= R || row0 + rr >= g.RH) continue;
      const int src_slot = e[2] > 0 ? e[3] : -1;
      if (sr0 < 0) { sr0 = rr; sd0 = seed_slot; sl0 = src_slot; }
      else { sr1 = rr; sd1 = seed_slot; sl1 = src_slot; }
    }
  }
  // wave-uniform special-pixel path: none, one slot, two slots; source
  // gradient stores only in waves that hold a source pixel
  const bool wave_sp = __ballot(sr0 >= 0) != 0;
  const bool wave_sp2 = __ballot(sr1 >= 0) != 0;
  const bool wave_gs = __ballot(sl0 >= 0 || sl1 >= 0) != 0;
  // output cells of this thread (gradient write-back, dJ/dx)
  uint32_t omask = 0;
  if (g.active && g.C0 + g.c >= g.oc0 && g.C0 + g.c < g.oc1) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int gr = g.R0 + row0 + r;
      if (row0 + r < g.RH && gr >= g.or0 && gr < g.or1) omask |= 1u << r;
    }
  }
  const uint32_t voff = cs_out_voff(g, row0, Ny);
  const int64_t out_bytes = (int64_t)(g.or1 - g.or0) * Ny * 4;
  // record of step j (read at reverse step m = j + 1): a sample's steps at or
  // past its length are not recorded, and lambda_m is exactly 0 there, so
  // those reads are sent out of range (they return 0)
  const int Lb = (BATCH && a.Tlim) ? a.Tlim[b] : 0x7fffffff;
  auto rec_off = [&](int j) { return (!BATCH || j < Lb) ? voff : 0x80000000u; };
  __syncthreads();

  const int64_t sslot0 =
      (BATCH && a.seg_soff) ? (int64_t)a.seg_soff[b] : (int64_t)b * a.seg_K;
  auto seg_rsrc = [&](int m) {
    return cs_rsrc(a.seg_lap + (sslot0 + (m - 1 - a.seg_n0)) * F + (int64_t)g.or0 * Ny,
                   out_bytes);
  };
  // lap(u_{m-1}) of the first two steps of this launch (SG0: even steps,
  // SG1: odd steps)
  {
    const __amdgpu_buffer_rsrc_t rs0 = seg_rsrc(a.m_hi);
    const uint32_t voff0 = rec_off(a.m_hi - 1);
#pragma unroll
    for (int r = 0; r < R; ++r) SG0[r] = cs_rec_load(rs0, voff0 + (uint32_t)(r * Ny * 4));
    // issued even for a one-step launch (out of range, reads nothing) so
    // that the loop's wait counts need not cover a path without them
    const bool two = a.nsteps > 1;
    const __amdgpu_buffer_rsrc_t rs1 = seg_rsrc(two ? a.m_hi - 1 : a.m_hi);
    const uint32_t voff1 = two ? rec_off(a.m_hi - 2) : 0x80000000u;
#pragma unroll
    for (int r = 0; r < R; ++r) SG1[r] = cs_rec_load(rs1, voff1 + (uint32_t)(r * Ny * 4));
  }

  auto step = [&](const float *P, float *Q, float(&Lc)[R], float(&Ln)[R], float(&SG)[R],
                  int s) {
    const int m = a.m_hi - s;
    int lo, hi;
    cone(g, s, a.nsteps, lo, hi);
    const bool need = g.active && row0 < hi && row0 + R > lo;
    // record index m-3-seg_n0 >= 0: m-2 >= m_hi-nsteps+1 >= seg_n0+1
    const bool prefetch = s + 2 < a.nsteps;
    const __amdgpu_buffer_rsrc_t rsn = seg_rsrc(prefetch ? m - 2 : m);
    const uint32_t pvoff = prefetch ? rec_off(m - 3) : 0x80000000u;
    float *gsrc_row = a.gsrc + ((int64_t)b * a.T + (m - 1)) * a.nsrcslot;
    if (need) {
      if (wave_sp) {
        const float gs0 = sd0 >= 0 ? seeds[sd0 * a.nsteps + s] : 0.0f;
        if (wave_sp2) {
          const float gs1 = sd1 >= 0 ? seeds[sd1 * a.nsteps + s] : 0.0f;
          cs_adj_rows<R, 2, true>(P, Q, Lc, Ln, KA1, KW, g.base, g.CS, omask, sr0, sr1, gs0,
                                  gs1, sl0, sl1, gsrc_row);
        } else if (wave_gs) {
          cs_adj_rows<R, 1, true>(P, Q, Lc, Ln, KA1, KW, g.base, g.CS, omask, sr0, -1, gs0,
                                  0.f, sl0, -1, gsrc_row);
        } else {
          cs_adj_rows<R, 1, false>(P, Q, Lc, Ln, KA1, KW, g.base, g.CS, omask, sr0, -1, gs0,
                                   0.f, -1, -1, gsrc_row);
        }
      } else {
        cs_adj_rows<R, 0, false>(P, Q, Lc, Ln, KA1, KW, g.base, g.CS, omask, -1, -1, 0.f, 0.f,
                                 -1, -1, gsrc_row);
      }
    }
    // gradient sum, then refill SG with lap(u) of step m-2 (the record is
    // read through a range-checked buffer based at the tile's first output
    // row, zero outside the output rows): one issue point
    // outside the branches, so the loads stay in flight across two steps.
    // Threads whose segment left the cone (or own no cells) accumulate
    // stale values into rows that are not output rows and never written.
    // The loads are always issued (pvoff is out of range when there is no
    // step m-2 in this launch) so that the wait counts stay exact.
    // The empty asm keeps each sum in this step: deferred into the next
    // step, the old and the refilled SG would both be live and the loop
    // would copy the fresh loads (waiting for them) at its back edge.
#pragma unroll
    for (int r = 0; r < R; ++r) {
      G[r] = __builtin_fmaf(SG[r], Ln[r], G[r]);
      asm volatile("" : "+v"(G[r]));
      SG[r] = cs_rec_load(rsn, pvoff + (uint32_t)(r * Ny * 4));
    }
    __syncthreads();
  };

  WT_TS(1);
  int s = 0;
  for (; s + 1 < a.nsteps; s += 2) {
    step(buf0, buf1, L1, L2, SG0, s);   // lambda_m -> L2
    step(buf1, buf0, L2, L1, SG1, s + 1);
  }
  if (s < a.nsteps) step(buf0, buf1, L1, L2, SG0, s);
  WT_TS(2);
  const bool odd = (a.nsteps & 1) != 0;
  if (g.active) {
    // running gradient sums: all loads first, then the stores
    const int64_t q0 = boff + (int64_t)(g.R0 + row0) * Ny + (g.C0 + g.c);
    float gold[R];
#pragma unroll
    for (int r = 0; r < R; ++r) gold[r] = ((omask >> r) & 1u) ? a.gacc[q0 + (int64_t)r * Ny] : 0.0f;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if ((omask >> r) & 1u) {
        const int64_t q = q0 + (int64_t)r * Ny;
        a.out1[q] = odd ? L2[r] : L1[r];
        a.out2[q] = odd ? L1[r] : L2[r];
        a.gacc[q] = gold[r] + G[r];
      }
    }
  }
  WT_TS(3);
}

template <int NT, int R>
__global__ __launch_bounds__(NT) void k_adj_cs(AdjArgs a) {
  adj_cs_body<NT, R, false>(a, xcd_remap(blockIdx.x, gridDim.x));
}

template <int NT, int R>
__global__ __launch_bounds__(NT) void k_adj_cs_batch(BatchAdj bl) {
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int gi = batch_group(bl.start, bl.ngrp, w);
  adj_cs_body<NT, R, true>(bl.grp[gi], w - bl.start[gi]);
}

// ------------------------------------------------------------------ nonlinear adjoint
//
// Temporally blocked adjoint of the nonlinear cell (cell.py:94-102 inside
// cell.py:12-17; the per-step form is k_loop_adj_nl_plain in wt_plain.hip,
// whose formulas this follows term by term).  The step's b and c depend on
// u, so
//   lambda_m = P_m lambda_{m+1} + L(w_m lambda_{m+1}) - a2_{m+1} lambda_{m+2} + g_m
// with w_m = A_m c_m^2, a2_m = A_m Bm_m and P_m the pointwise factor.  Each
// iteration m forms lambda_m, then the coefficients of step m-1 from
// u_{m-1}, u_{m-2} and lap(u_{m-1}) (segment record), accumulates that step's
// parameter gradients and carries w_{m-1}, P_{m-1}, a2_{m-1} in registers to
// the next iterations.  Between launches the carried coefficients travel
// through global memory with the two lambdas (NlCarry), so a launch may start
// at any step.  The halo cells also need the record (their coefficients
// drive lambda inside the cone), so the record is read over the whole region.
template <int NT, int R, bool GB>
__global__ __launch_bounds__(NT) void k_adj_cs_nl(AdjArgs a, NlCarry cy) {
  extern __shared__ float smem[];
  const int tid = threadIdx.x;
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int b = w / a.ntiles, ti = w - (w / a.ntiles) * a.ntiles;
  const int32_t *td = a.tiles + ti * TILE_W;
  const CsGeom g = cs_geom<R>(td, tid);
  const int NBUF = (g.RW + 2) * g.CS;
  float *buf0 = smem, *buf1 = smem + NBUF;
  float *seeds = smem + 2 * NBUF;
  // material (b_geom, c_linear, rho, 0) of every region cell, LDS index as
  // the field buffers, one 16-byte read per cell and step
  float4 *mat = (float4 *)(smem + ((2 * NBUF + CS_MAX_SEEDS * CS_MAX_TBLOCK + 3) & ~3));
  __shared__ int seed_px[CS_MAX_SEEDS];
  __shared__ int nseeds;
  __shared__ int tb_sp[CS_MAX_TABLE * 6];
  const int Ny = a.Ny;
  const int64_t F = (int64_t)a.Nx * Ny;
  const int64_t boff = (int64_t)b * F;
  const int nsp = min(td[9] - td[8], CS_MAX_TABLE);
  const int row0 = g.k * R;
  const int col = g.C0 + g.c;

  float Lc[R], Ln[R], Wm[R], Pm[R], A2a[R], A2b[R], U1[R], Gc[R], Gr[R], Gb[R];
  uint32_t vmask = 0;
  // byte offset of this thread's first cell within one (Nx,Ny) field
  const uint32_t voff = (uint32_t)(((g.R0 + row0) * Ny + col) * 4);
  // 1) global loads: lambdas, carried coefficients, u_{m_hi - 1}, material
  {
    const __amdgpu_buffer_rsrc_t ru = cs_rsrc(
        a.seg_u + ((int64_t)b * a.seg_K + (a.m_hi - 1 - a.seg_n0)) * F, F * 4);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      float l1 = 0.f, l2 = 0.f, wv = 0.f, pv = 0.f, xa = 0.f, xb = 0.f, uu = 0.f;
      float4 mt = make_float4(0.f, 0.f, 0.f, 0.f);
      if (g.active && row0 + r < g.RH) {
        vmask |= 1u << r;
        const int64_t q = (int64_t)(g.R0 + row0 + r) * Ny + col;
        l1 = a.in1[boff + q]; l2 = a.in2[boff + q];
        wv = cy.w_in[boff + q]; pv = cy.p_in[boff + q];
        xa = cy.a2a_in[boff + q]; xb = cy.a2b_in[boff + q];
        uu = cs_bload(ru, voff + (uint32_t)(r * Ny * 4));
        mt = make_float4(a.bgeom[q], a.clin[q], a.rho[q], 0.f);
      }
      Lc[r] = l1; Ln[r] = l2; Wm[r] = wv; Pm[r] = pv; A2a[r] = xa; A2b[r] = xb; U1[r] = uu;
      Gc[r] = 0.f; Gr[r] = 0.f; Gb[r] = 0.f;
      if (g.active) mat[g.base + r] = mt;
    }
  }
  int e0 = 0, e1 = 0, e2 = 0, e3 = 0, e4 = 0, e5 = 0;
  if (tid < nsp) {
    const int32_t *e = a.spx + (td[8] + tid) * SPX_W;
    e0 = e[0]; e1 = e[1]; e2 = e[2]; e3 = e[3]; e4 = e[4]; e5 = e[5];
  }
  // 2) LDS: w_m * lambda_{m+1} of the region, tables, probe seeds
  for (int i = tid; i < 2 * NBUF; i += blockDim.x) smem[i] = 0.0f;
  if (tid < nsp) {
    int *t = tb_sp + 6 * tid;
    t[0] = e0; t[1] = e1; t[2] = e2; t[3] = e3; t[4] = e4; t[5] = e5;
  }
  __syncthreads();
  if (tid == 0) {
    int n = 0;
    for (int kk = 0; kk < nsp && n < CS_MAX_SEEDS; ++kk)
      if (tb_sp[6 * kk + 5] > tb_sp[6 * kk + 4]) seed_px[n++] = kk;
    nseeds = n;
  }
  if (g.active) {
#pragma unroll
    for (int r = 0; r < R; ++r) buf0[g.base + r] = Wm[r] * Lc[r];
  }
  __syncthreads();
  for (int i = tid; i < nseeds * a.nsteps; i += blockDim.x) {
    const int k = i / a.nsteps, s = i - (i / a.nsteps) * a.nsteps;
    const int m = a.m_hi - s;
    const int *e = tb_sp + 6 * seed_px[k];
    float gs = 0.0f;
    if (a.ybar) {
      for (int t = e[4]; t < e[5]; ++t) {
        const int32_t *tp = a.taps + t * TAP_W;
        const int64_t o = ((int64_t)b * a.T + (m - 1)) * a.ntap + tp[3];
        const float gy = a.ybar[o];
        gs += tp[2] ? 2.0f * a.yraw[o] * gy : gy;
      }
    }
    seeds[i] = gs;
  }
  int sr0 = -1, sr1 = -1, sd0 = -1, sd1 = -1, sl0 = -1, sl1 = -1;
  {
    int nseed = 0;
    for (int kk = 0; kk < nsp; ++kk) {
      const int *e = tb_sp + 6 * kk;
      const bool has_tap = e[5] > e[4];
      const int seed_slot = has_tap ? nseed : -1;
      if (has_tap) ++nseed;
      const int rr = e[0] - g.R0 - row0;
      if (!g.active || e[1] - g.C0 != g.c || rr < 0 || rr >= R || row0 + rr >= g.RH) continue;
      const int src_slot = e[2] > 0 ? e[3] : -1;
      if (sr0 < 0) { sr0 = rr; sd0 = seed_slot; sl0 = src_slot; }
      else { sr1 = rr; sd1 = seed_slot; sl1 = src_slot; }
    }
  }
  uint32_t omask = 0;
  if (g.active && col >= g.oc0 && col < g.oc1) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int gr = g.R0 + row0 + r;
      if (row0 + r < g.RH && gr >= g.or0 && gr < g.or1) omask |= 1u << r;
    }
  }
  __syncthreads();

  const float s2 = a.s2, dtm1 = a.dtm1, dtm2 = a.dtm2;
  const float uth = a.sat_uth, b0 = a.sat_b0, cnl = a.c_nl;
  const int nlf = a.nlf;
  // Record values of the step being formed: LL = lap(u_{m-1}) and UL =
  // u_{m-2} (from the checkpoint's u_{n0-1} when m-2 precedes the segment).
  // Each step loads those of the next step at its end, outside every branch
  // and always issued (out of range, reading nothing, past the launch's last
  // step or outside the region: vmask), so that they travel during the
  // barrier and the next step's lambda pass.
  float LL[R], UL[R];
  auto load_record = [&](int m, bool valid) {
    const __amdgpu_buffer_rsrc_t rl =
        cs_rsrc(a.seg_lap + ((int64_t)b * a.seg_K + (valid ? m - 1 - a.seg_n0 : 0)) * F, F * 4);
    const __amdgpu_buffer_rsrc_t ru =
        (!valid || (m - 2) < a.seg_n0)
            ? cs_rsrc(cy.uprev0 + boff, F * 4)
            : cs_rsrc(a.seg_u + ((int64_t)b * a.seg_K + (m - 2 - a.seg_n0)) * F, F * 4);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t o =
          (valid && ((vmask >> r) & 1u)) ? voff + (uint32_t)(r * Ny * 4) : 0x80000000u;
      LL[r] = cs_rec_load(rl, o);
      UL[r] = cs_rec_load(ru, o);
    }
  };
  load_record(a.m_hi, true);
  auto step = [&](const float *P, float *Q, float(&Lx)[R], float(&Ly)[R], int s) {
    // Lx = lambda_{m+1}, Ly = lambda_{m+2} on entry, lambda_m on exit
    const int m = a.m_hi - s;
    int lo, hi;
    cone(g, s, a.nsteps, lo, hi);
    const bool need = g.active && row0 < hi && row0 + R > lo;
    if (need) {
      const float gs0 = sd0 >= 0 ? seeds[sd0 * a.nsteps + s] : 0.0f;
      const float gs1 = sd1 >= 0 ? seeds[sd1 * a.nsteps + s] : 0.0f;
      const int lfb = g.base - g.CS, rtb = g.base + g.CS;
      // 1) lambda_m of every row (needs only this step's coefficients)
      float upv = P[g.base - 1];
      float ct = Wm[0] * Lx[0];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float dn = (r < R - 1) ? Wm[r + 1] * Lx[r + 1] : P[g.base + R];
        const float lf = P[lfb + r], rt = P[rtb + r];
        float lapw = upv + lf;
        lapw = __builtin_fmaf(-4.0f, ct, lapw);
        lapw = lapw + rt;
        lapw = lapw + dn;
        float lam = (Pm[r] * Lx[r] + lapw) - A2a[r] * Ly[r];
        if (r == sr0) lam += gs0;
        if (r == sr1) lam += gs1;
        upv = ct;
        ct = dn;
        Ly[r] = lam;
      }
      // 2) coefficients of step m-1 at each cell (u = u_{m-1}, up = u_{m-2}),
      // that step's parameter gradients, w_{m-1} * lambda_m to LDS
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float lam = Ly[r];
        const float4 mt = mat[g.base + r];
        const float u = U1[r], up = UL[r], lapu = LL[r], rho = mt.z;
        float bb = mt.x, cc = mt.y, sd = 0.0f, c2raw = 0.0f, dbdu = 0.0f, dcdu = 0.0f;
        if (nlf & 1) {
          const float q = adj_div(u, uth), d = 1.0f + q * q;
          sd = adj_div(b0, d);
          bb += rho * sd;
          dbdu = adj_div(rho * b0 * adj_div(-2.0f * q, uth), d * d);
        }
        if (nlf & 2) {
          c2raw = cnl * (u * u);
          cc += rho * c2raw;
          dcdu = 2.0f * rho * cnl * u;
        }
        bool clamped = false;
        if (nlf & 4) {
          if (cc < a.cmin) { cc = a.cmin; clamped = true; }
          if (cc > a.cmax) { cc = a.cmax; clamped = true; }
        }
        const float bd = bb * dtm1;
        const float A = adj_div(1.0f, dtm2 + bd);
        const float Bm = dtm2 - bd;
        const float v = A * ((s2 * u - Bm * up) + (cc * cc) * lapu);
        const float dvdb = A * dtm1 * (up - v);
        const float dvdc = clamped ? 0.0f : 2.0f * A * cc * lapu;
        Gc[r] = __builtin_fmaf(dvdc, lam, Gc[r]);
        if (GB) Gb[r] = __builtin_fmaf(dvdb, lam, Gb[r]);
        Gr[r] = __builtin_fmaf(dvdb * sd + dvdc * c2raw, lam, Gr[r]);
        const float wn = A * (cc * cc);
        Wm[r] = wn;
        Pm[r] = s2 * A + dvdb * dbdu + dvdc * dcdu;
        A2a[r] = A2b[r];
        A2b[r] = A * Bm;
        U1[r] = up;
        Q[g.base + r] = wn * lam;
      }
      if (sr0 >= 0) {
        // lambda_m at source pixels of the output rows -> dJ/dx[b, m-1]:
        // picked with selects, at most two stores
        float *gsrc_row = a.gsrc + ((int64_t)b * a.T + (m - 1)) * a.nsrcslot;
        float v0 = Ly[0], v1 = Ly[0];
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v0 = (r == sr0) ? Ly[r] : v0;
          v1 = (r == sr1) ? Ly[r] : v1;
        }
        if (sl0 >= 0 && ((omask >> sr0) & 1u)) gsrc_row[sl0] = v0;
        if (sl1 >= 0 && sr1 >= 0 && ((omask >> sr1) & 1u)) gsrc_row[sl1] = v1;
      }
    }
    load_record(m - 1, s + 1 < a.nsteps);
    __syncthreads();
  };

  int s = 0;
  for (; s + 1 < a.nsteps; s += 2) {
    step(buf0, buf1, Lc, Ln, s);      // lambda_m -> Ln
    step(buf1, buf0, Ln, Lc, s + 1);  // lambda_{m-1} -> Lc
  }
  if (s < a.nsteps) step(buf0, buf1, Lc, Ln, s);
  const bool odd = (a.nsteps & 1) != 0;
  if (g.active) {
    const int64_t q0 = boff + (int64_t)(g.R0 + row0) * Ny + col;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if ((omask >> r) & 1u) {
        const int64_t q = q0 + (int64_t)r * Ny;
        a.out1[q] = odd ? Ln[r] : Lc[r];
        a.out2[q] = odd ? Lc[r] : Ln[r];
        cy.w_out[q] = Wm[r];
        cy.p_out[q] = Pm[r];
        cy.a2a_out[q] = A2a[r];
        cy.a2b_out[q] = A2b[r];
        a.gacc[q] += Gc[r];
        a.gacc_rho[q] += Gr[r];
        if (GB) a.gacc_b[q] += Gb[r];
      }
    }
  }
}

// ------------------------------------------------------------------ launchers

static int cs_max_dyn_lds() { return max_dyn_lds_dev(); }

#define WT_CS_VARIANTS(X) \
  X(1024, 4)              \
  X(1024, 6)              \
  X(1024, 8)              \
  X(1024, 10)             \
  X(1024, 12)             \
  X(512, 16)              \
  X(512, 20)              \
  X(512, 24)

#ifdef WT_PHASE_TIMING
extern "C" int wt_debug_phase_buffer(unsigned long long *p) {
  return hipMemcpyToSymbol(HIP_SYMBOL(wt_phase_ts), &p, sizeof(p)) == hipSuccess ? 0 : -3;
}
#endif

bool cs_variant_exists(int nt, int r) {
#define X(A, B) if (nt == A && r == B) return true;
  WT_CS_VARIANTS(X)
#undef X
  return false;
}

// Largest LDS column stride of a region of at most RH rows; `padded`:
// with cs_layout's slack (forward tiles, whose descriptors carry a layout),
// so that every tile of the launch fits the allocation.
static int cs_stride_alloc(int R, int RH, bool padded) {
  int nseg = (RH + R - 1) / R;
  return ((nseg * R + 2) | 1) + (padded ? cs_lds_slack(R) : 0);
}

// LDS bytes of a region with RH rows and RW columns for rows-per-thread R.
size_t cs_shmem_bytes(int R, int RH, int RW, bool padded) {
  int CS = cs_stride_alloc(R, RH, padded);
  // field double buffer + staged inputs (forward) or probe seeds (adjoint)
  return ((size_t)2 * (RW + 2) * CS + (size_t)CS_MAX_SEEDS * CS_MAX_TBLOCK) * sizeof(float);
}

// Threads of a launch whose largest region is RH x RW: whole waves for the
// plain mapping.  cs_layout pads columns only within these waves: an extra
// wave costs a VALU-bound block more than the bank conflicts it removes
// (measured: 12 -> 13 waves made the nonlinear forward 13% slower).
int cs_threads(int R, int RH, int RW) {
  int nseg = (RH + R - 1) / R;
  int t = nseg * RW;
  return (t + 63) / 64 * 64;
}

template <int NT, int R>
static hipError_t cs_fwd_dispatch(int nl, bool rec, const FwdArgs &a, int nblocks,
                                  int nthreads, size_t shmem, hipStream_t s) {
#define FT(NL, REC)                                                                   \
  do {                                                                                \
    auto kfn = k_fwd_cs<NT, R, NL, REC>;                                              \
    static std::atomic<uint64_t> done{0};                                             \
    hipError_t e = allow_dyn_lds((const void *)kfn, done);                            \
    if (e != hipSuccess) return e;                                                    \
    hipLaunchKernelGGL(kfn, dim3(nblocks), dim3(nthreads), shmem, s, a);              \
  } while (0)
  if (nl == 0) {
    if (rec) FT(0, true); else FT(0, false);
  } else {
    if (rec) FT(8, true); else FT(8, false);
  }
#undef FT
  return hipGetLastError();
}

template <int NT, int R>
static hipError_t cs_adj_dispatch(const AdjArgs &a, int nblocks, int nthreads, size_t shmem,
                                  hipStream_t s) {
  auto kfn = k_adj_cs<NT, R>;
  static std::atomic<uint64_t> done{0};
  hipError_t e = allow_dyn_lds((const void *)kfn, done);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kfn, dim3(nblocks), dim3(nthreads), shmem, s, a);
  return hipGetLastError();
}

// The region geometry (RH, RW) of the largest tile decides the block size
// and LDS of the launch; smaller tiles leave threads idle.
hipError_t launch_fwd_cs(int nt, int R, int nl, bool rec, const FwdArgs &a, int RHmax,
                         int RWmax, hipStream_t s) {
  int nblocks = a.B * a.ntiles;
  size_t shmem = cs_shmem_bytes(R, RHmax, RWmax, true);
  int nthreads = cs_threads(R, RHmax, RWmax);
  if (shmem + 1024 > (size_t)cs_max_dyn_lds() || nthreads > nt) return hipErrorInvalidConfiguration;
#define X(A, B) \
  if (nt == A && R == B) return cs_fwd_dispatch<A, B>(nl, rec, a, nblocks, nthreads, shmem, s);
  WT_CS_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_adj_cs(int nt, int R, const AdjArgs &a, int RHmax, int RWmax, hipStream_t s) {
  int nblocks = a.B * a.ntiles;
  size_t shmem = cs_shmem_bytes(R, RHmax, RWmax, false);
  int nthreads = cs_threads(R, RHmax, RWmax);
  if (shmem + 1024 > (size_t)cs_max_dyn_lds() || nthreads > nt) return hipErrorInvalidConfiguration;
#define X(A, B) if (nt == A && R == B) return cs_adj_dispatch<A, B>(a, nblocks, nthreads, shmem, s);
  WT_CS_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

// Nonlinear adjoint: 4 rows per thread at 1024 threads (the carried
// coefficients fill the register file; 6 rows spill).
bool cs_nl_variant_exists(int nt, int r) { return nt == 1024 && r == 4; }

size_t cs_nl_shmem_bytes(int R, int RH, int RW) {
  int CS = cs_stride_alloc(R, RH, false);
  return cs_shmem_bytes(R, RH, RW, false) + 16 + (size_t)(RW + 2) * CS * 16;
}

hipError_t launch_adj_cs_nl(int nt, int R, bool gb, const AdjArgs &a, const NlCarry &cy,
                            int RHmax, int RWmax, hipStream_t s) {
  int nblocks = a.B * a.ntiles;
  size_t shmem = cs_nl_shmem_bytes(R, RHmax, RWmax);
  int nthreads = cs_threads(R, RHmax, RWmax);
  if (shmem + 1024 > (size_t)cs_max_dyn_lds() || nthreads > nt || !cs_nl_variant_exists(nt, R))
    return hipErrorInvalidConfiguration;
#define NLK(RR, G)                                                                     \
  do {                                                                                 \
    auto kfn = k_adj_cs_nl<1024, RR, G>;                                               \
    static std::atomic<uint64_t> done{0};                                              \
    hipError_t e = allow_dyn_lds((const void *)kfn, done);                             \
    if (e != hipSuccess) return e;                                                     \
    hipLaunchKernelGGL(kfn, dim3(nblocks), dim3(nthreads), shmem, s, a, cy);           \
  } while (0)
  if (gb) NLK(4, true); else NLK(4, false);
#undef NLK
  return hipGetLastError();
}

// Batched launches (linear cell): one launch over the work items of every
// configuration in `bl`; RHmax/RWmax are the largest region over all of them.
template <int NT, int R>
static hipError_t cs_fwd_batch_dispatch(bool rec, const BatchFwd &bl, int nblocks, int nthreads,
                                        size_t shmem, hipStream_t s) {
#define FB(REC)                                                                       \
  do {                                                                                \
    auto kfn = k_fwd_cs_batch<NT, R, 0, REC>;                                         \
    static std::atomic<uint64_t> done{0};                                             \
    hipError_t e = allow_dyn_lds((const void *)kfn, done);                            \
    if (e != hipSuccess) return e;                                                    \
    hipLaunchKernelGGL(kfn, dim3(nblocks), dim3(nthreads), shmem, s, bl);             \
  } while (0)
  if (rec) FB(true); else FB(false);
#undef FB
  return hipGetLastError();
}

template <int NT, int R>
static hipError_t cs_adj_batch_dispatch(const BatchAdj &bl, int nblocks, int nthreads,
                                        size_t shmem, hipStream_t s) {
  auto kfn = k_adj_cs_batch<NT, R>;
  static std::atomic<uint64_t> done{0};
  hipError_t e = allow_dyn_lds((const void *)kfn, done);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kfn, dim3(nblocks), dim3(nthreads), shmem, s, bl);
  return hipGetLastError();
}

hipError_t launch_fwd_cs_batch(int nt, int R, bool rec, const BatchFwd &bl, int RHmax,
                               int RWmax, hipStream_t s) {
  int nblocks = bl.start[bl.ngrp];
  if (nblocks <= 0) return hipSuccess;
  size_t shmem = cs_shmem_bytes(R, RHmax, RWmax, true);
  int nthreads = cs_threads(R, RHmax, RWmax);
  if (shmem + 1024 > (size_t)cs_max_dyn_lds() || nthreads > nt) return hipErrorInvalidConfiguration;
#define X(A, B) \
  if (nt == A && R == B) return cs_fwd_batch_dispatch<A, B>(rec, bl, nblocks, nthreads, shmem, s);
  WT_CS_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_adj_cs_batch(int nt, int R, const BatchAdj &bl, int RHmax, int RWmax,
                               hipStream_t s) {
  int nblocks = bl.start[bl.ngrp];
  if (nblocks <= 0) return hipSuccess;
  size_t shmem = cs_shmem_bytes(R, RHmax, RWmax, false);
  int nthreads = cs_threads(R, RHmax, RWmax);
  if (shmem + 1024 > (size_t)cs_max_dyn_lds() || nthreads > nt) return hipErrorInvalidConfiguration;
#define X(A, B) \
  if (nt == A && R == B) return cs_adj_batch_dispatch<A, B>(bl, nblocks, nthreads, shmem, s);
  WT_CS_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

}  // namespace wt
