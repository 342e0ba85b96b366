// wt_tiled.hip -- temporally blocked WaveRNN kernels for gfx950.
//
// One workgroup advances one rectangular region (a tile plus a halo) of one
// sample by `nsteps` time steps with the field resident in LDS, then writes
// back the tile.  The halo is as deep as the number of steps, so every
// workgroup is independent within a launch: no inter-workgroup
// synchronisation, no spin-waits.  When a tile covers the whole domain the
// halo is empty and one launch can run any number of steps.
//
// LDS layout of one field buffer (row stride LW = RW + 1, the extra column
// is a shared zero pad: right neighbour of a row's last cell and left
// neighbour of the next row's first cell):
//     [0, LW)                 zero row above the region
//     [LW + r*LW + c]         cell (r, c) of the region, c == RW is the pad
//     [LW + RH*LW, ...)       zero rows / dummy cells below
// Thread t owns the flat cells f = t + j*NT, j < CPT (consecutive lanes ->
// consecutive LDS words: conflict-free ds_read/ds_write).  Per-cell material
// coefficients and the previous field live in VGPRs; the current field is
// double buffered in LDS (read P, write Q, one barrier per step).  Pad,
// dummy and out-of-domain cells have all coefficients zero and stay zero,
// which implements conv2d's zero padding (operators.py:11) without branches.
//
// Forward (k_fwd_tiled) per cell, in the reference's order (cell.py:12-17):
//     lap = (((up + lf) + (-4*ctr)) + rt) + dn
//     v   = A*(((s2*ctr) - (Bm*y2)) + (c2*lap))
// then the owning thread adds the source samples (source.py:19-22), and
// after the barrier probe threads read u_{n+1} (probe.py:15,27).
//
// Adjoint (k_adj_tiled, row strips) per cell (the transpose of the step;
// what autograd computes from cell.py:37-42 chained over rnn.py:50):
//     lam_m = a1*lam_{m+1} + L(w*lam_{m+1}) - a2*lam_{m+2}   (+ probe seed)
//     gacc += lap(u_{m-1}) * lam_m          (dJ/dc, cell.py:36, summed over t)
// with a1 = s2*A, a2 = A*Bm, w = A*C2; w*lam is what lives in LDS.
// Cells carrying a probe tap or a source pixel are "special": their main
// loop coefficients are zero and the owning thread recomputes them after
// the loop with scalar state, adding the probe seed 2*u*ybar and recording
// lam at source pixels (dJ/dx, source.py:22).
#include "wt_internal.h"

namespace wt {

constexpr int SRC_SLOTS = 4;  // source-pixel entries one thread can own
constexpr int SPC_SLOTS = 2;  // special cells one thread can own (adjoint)

template <bool UNIT_H>
__device__ __forceinline__ float lap_lds(const float *P, int i, int LW, float t1,
                                         float t4) {
  float up = P[i - LW], lf = P[i - 1], ct = P[i], rt = P[i + 1], dn = P[i + LW];
  if (UNIT_H) {
    float s = up + lf;
    s = s + (-4.0f) * ct;
    s = s + rt;
    return s + dn;
  } else {
    float s = t1 * up;
    s = s + t1 * lf;
    s = s + t4 * ct;
    s = s + t1 * rt;
    return s + t1 * dn;
  }
}

// Nonlinear material of one cell for the current field value u
// (cell.py:94-102), reference operation order.
// NL: bit0 saturable damping, bit1 Kerr speed, bit2 clamp; the value 8 means
// "flags at run time" (nlf), which is what the tiled kernels instantiate.
template <int NL>
__device__ __forceinline__ void nl_coef(float u, float bg, float cl, float rh,
                                        float rc, float sat_uth, float sat_b0,
                                        float cmin, float cmax, float dtm1,
                                        float dtm2, int nlf, float &A, float &Bm,
                                        float &c2) {
  const int fl = NL == 8 ? nlf : NL;
  float bb = bg, cc = cl;
  if (fl & 1) {
    float r = u / sat_uth;
    r = fabsf(r);
    r = r * r;
    r = r + 1.0f;
    r = sat_b0 / r;
    bb = bb + rh * r;
  }
  if (fl & 2) cc = cc + rc * (u * u);
  if (fl & 4) cc = fminf(fmaxf(cc, cmin), cmax);
  float bd = bb * dtm1;
  A = 1.0f / (dtm2 + bd);
  Bm = dtm2 - bd;
  c2 = cc * cc;
}

// Integer f / LW for f < 2^22 via a float reciprocal with one correction.
__device__ __forceinline__ int div_lw(int f, int LW, float inv) {
  int r = (int)((float)f * inv);
  int rem = f - r * LW;
  if (rem < 0) { --r; rem += LW; }
  if (rem >= LW) { ++r; }
  return r;
}

// ------------------------------------------------------------------ forward
template <int NT, int CPT, bool UNIT_H, int NL, bool REC>
__global__ __launch_bounds__(NT) void k_fwd_tiled(FwdArgs a) {
  extern __shared__ float smem[];
  const int tid = threadIdx.x;
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int b = w / a.ntiles, ti = w - (w / a.ntiles) * a.ntiles;
  const int32_t *td = a.tiles + ti * TILE_W;
  const int R0 = td[0], RH = td[1], C0 = td[2], RW = td[3];
  const int or0 = td[4], or1 = td[5], oc0 = td[6], oc1 = td[7];
  const int LW = RW + 1;
  const int NB = 2 * LW + NT * CPT;
  const float invLW = 1.0f / (float)LW;
  float *buf0 = smem, *buf1 = smem + NB;
  const int Nx = a.Nx, Ny = a.Ny;
  const int64_t F = (int64_t)Nx * Ny;
  const int64_t boff = (int64_t)b * F;

  for (int i = tid; i < 2 * NB; i += NT) smem[i] = 0.0f;
  __syncthreads();

  // per-cell state: coefficient registers and the previous field
  float c0r[CPT], c1r[CPT], c2r[CPT], c3r[CPT], y2[CPT];
  uint64_t vmask = 0;  // nonlinear cell: which of this thread's cells are real
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    int f = tid + j * NT;
    int r = div_lw(f, LW, invLW);
    int cc = f - r * LW;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, u1 = 0.f, u2 = 0.f;
    if (r < RH && cc < RW) {
      vmask |= 1ull << j;
      int64_t q = (int64_t)(R0 + r) * Ny + (C0 + cc);
      if (NL == 0) {
        v0 = a.A[q]; v1 = a.Bm[q]; v2 = a.C2[q];
      } else {
        v0 = a.bgeom[q]; v1 = a.clin[q]; v2 = a.rho[q]; v3 = a.rhoc[q];
      }
      u1 = a.in1[boff + q];
      u2 = a.in2[boff + q];
    }
    c0r[j] = v0; c1r[j] = v1; c2r[j] = v2; c3r[j] = v3;
    y2[j] = u2;
    buf0[LW + f] = u1;
  }

  // sources owned by this thread (LDS index; repeated for multiplicity)
  int sidx[SRC_SLOTS];
#pragma unroll
  for (int s = 0; s < SRC_SLOTS; ++s) sidx[s] = -1;
  {
    int nmine = 0;
    for (int k = td[8]; k < td[9]; ++k) {
      const int32_t *e = a.spx + k * SPX_W;
      int mult = e[2];
      if (mult <= 0) continue;
      int rr = e[0] - R0, yy = e[1] - C0;
      if (rr < 0 || rr >= RH || yy < 0 || yy >= RW) continue;
      int f = rr * LW + yy;
      if (f % NT != tid) continue;
      for (int m = 0; m < mult; ++m) {
#pragma unroll
        for (int s = 0; s < SRC_SLOTS; ++s)
          if (s == nmine) sidx[s] = LW + f;
        ++nmine;
      }
    }
  }
  // probe tap handled by this thread (taps in the output region)
  int pidx = -1, pkind = 0, pcol = 0;
  {
    int kk = 0;
    for (int k = td[10]; k < td[11]; ++k) {
      const int32_t *tp = a.taps + k * TAP_W;
      if (tp[1] < oc0 || tp[1] >= oc1) continue;
      if (kk == tid) {
        pidx = LW + (tp[0] - R0) * LW + (tp[1] - C0);
        pkind = tp[2];
        pcol = tp[3];
      }
      ++kk;
    }
  }
  // record mode (row strips): lap(u_n) of output rows, domain-major
  const int flo = (or0 - R0) * LW, fspan = (or1 - or0) * LW;
  __syncthreads();

  for (int s = 0; s < a.nsteps; ++s) {
    const int n = a.n0 + s;
    float *P = (s & 1) ? buf1 : buf0;
    float *Q = (s & 1) ? buf0 : buf1;
    float *recl = nullptr, *recu = nullptr;
    if (REC) {
      int64_t o = ((int64_t)b * a.rec_K + (n - a.rec_n0)) * Nx * (int64_t)LW +
                  (int64_t)R0 * LW;
      recl = a.rec_lap + o;
      if (NL) recu = a.rec_u + o;
    }
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int f = tid + j * NT;
      const int i = LW + f;
      const float ct = P[i];
      const float lap = lap_lds<UNIT_H>(P, i, LW, a.t1, a.t4);
      float A, Bm, c2;
      if (NL == 0) {
        A = c0r[j]; Bm = c1r[j]; c2 = c2r[j];
      } else {
        nl_coef<NL>(ct, c0r[j], c1r[j], c2r[j], c3r[j], a.sat_uth, a.sat_b0,
                    a.cmin, a.cmax, a.dtm1, a.dtm2, a.nlf, A, Bm, c2);
        // pad / dummy / outside cells: keep them at zero
        if (!((vmask >> j) & 1ull)) A = 0.0f;
      }
      const float p1 = a.s2 * ct;
      const float p2 = Bm * y2[j];
      const float p3 = c2 * lap;
      Q[i] = A * ((p1 - p2) + p3);
      y2[j] = ct;
      if (REC) {
        if ((unsigned)(f - flo) < (unsigned)fspan) {
          recl[f] = lap;
          if (NL) recu[f] = ct;
        }
      }
    }
    // source injection by the owning thread (rnn.py:56-57, source.py:22)
    if (sidx[0] >= 0) {
      const float xv = a.x[(int64_t)b * a.T + n];
#pragma unroll
      for (int s2i = 0; s2i < SRC_SLOTS; ++s2i)
        if (sidx[s2i] >= 0) Q[sidx[s2i]] = Q[sidx[s2i]] + xv;
    }
    __syncthreads();
    if (pidx >= 0) {
      const float v = Q[pidx];
      const int64_t o = ((int64_t)b * a.T + n) * a.ntap + pcol;
      if (a.y) a.y[o] = pkind ? v * v : v;
      if (a.yraw) a.yraw[o] = v;
    }
    if (a.fields && ((n + 1) % a.fstride) == 0) {
      const int64_t nf = a.T / a.fstride;
      float *fo = a.fields + ((int64_t)b * nf + (n + 1) / a.fstride - 1) * F;
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        const int f = tid + j * NT;
        const int r = div_lw(f, LW, invLW);
        const int cc = f - r * LW;
        const int gr = R0 + r, gc = C0 + cc;
        if (r < RH && cc < RW && gr >= or0 && gr < or1 && gc >= oc0 && gc < oc1)
          fo[(int64_t)gr * Ny + gc] = Q[LW + f];
      }
    }
  }
  // write back the tile: u_{n0+nsteps}, u_{n0+nsteps-1}
  const float *Pf = (a.nsteps & 1) ? buf1 : buf0;
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int f = tid + j * NT;
    const int r = div_lw(f, LW, invLW);
    const int cc = f - r * LW;
    const int gr = R0 + r, gc = C0 + cc;
    if (r < RH && cc < RW && gr >= or0 && gr < or1 && gc >= oc0 && gc < oc1) {
      const int64_t q = boff + (int64_t)gr * Ny + gc;
      a.out1[q] = Pf[LW + f];
      a.out2[q] = y2[j];
    }
  }
}

// ------------------------------------------------------------------ adjoint
// Linear cell, row strips (C0 == 0, RW == Ny).
template <int NT, int CPT, bool UNIT_H>
__global__ __launch_bounds__(NT) void k_adj_tiled(AdjArgs a) {
  extern __shared__ float smem[];
  const int tid = threadIdx.x;
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int b = w / a.ntiles, ti = w - (w / a.ntiles) * a.ntiles;
  const int32_t *td = a.tiles + ti * TILE_W;
  const int R0 = td[0], RH = td[1], RW = td[3];
  const int or0 = td[4], or1 = td[5];
  const int LW = RW + 1;
  const int NB = 2 * LW + NT * CPT;
  const float invLW = 1.0f / (float)LW;
  float *buf0 = smem, *buf1 = smem + NB;
  const int Nx = a.Nx, Ny = a.Ny;
  const int64_t F = (int64_t)Nx * Ny;
  const int64_t boff = (int64_t)b * F;
  const float s2 = a.s2;

  for (int i = tid; i < 2 * NB; i += NT) smem[i] = 0.0f;
  __syncthreads();

  float ka1[CPT], ka2[CPT], kw[CPT], l1[CPT], l2[CPT], ga[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int f = tid + j * NT;
    const int r = div_lw(f, LW, invLW);
    const int cc = f - r * LW;
    float A = 0.f, Bm = 0.f, C2 = 0.f, u1 = 0.f, u2 = 0.f;
    if (r < RH && cc < RW) {
      const int64_t q = (int64_t)(R0 + r) * Ny + cc;
      A = a.A[q]; Bm = a.Bm[q]; C2 = a.C2[q];
      u1 = a.in1[boff + q];
      u2 = a.in2[boff + q];
    }
    ka1[j] = A * s2; ka2[j] = A * Bm; kw[j] = A * C2;
    l1[j] = u1; l2[j] = u2; ga[j] = 0.0f;
    buf0[LW + f] = kw[j] * u1;
  }

  // special cells owned by this thread
  int sp_i[SPC_SLOTS], sp_j[SPC_SLOTS], sp_e[SPC_SLOTS];
  float sp_a1[SPC_SLOTS], sp_a2[SPC_SLOTS], sp_w[SPC_SLOTS], sp_l1[SPC_SLOTS],
      sp_l2[SPC_SLOTS], sp_g[SPC_SLOTS];
#pragma unroll
  for (int s = 0; s < SPC_SLOTS; ++s) {
    sp_i[s] = -1; sp_j[s] = -1; sp_e[s] = 0;
    sp_a1[s] = sp_a2[s] = sp_w[s] = sp_l1[s] = sp_l2[s] = sp_g[s] = 0.f;
  }
  {
    int nmine = 0;
    for (int k = td[8]; k < td[9]; ++k) {
      const int32_t *e = a.spx + k * SPX_W;
      const int rr = e[0] - R0;
      if (rr < 0 || rr >= RH) continue;
      const int f = rr * LW + e[1];
      if (f % NT != tid) continue;
      const int jj = f / NT;
      const int64_t q = (int64_t)e[0] * Ny + e[1];
      const float A = a.A[q];
      const float ca1 = A * s2, ca2 = A * a.Bm[q], cw = A * a.C2[q];
      const float u1 = a.in1[boff + q], u2 = a.in2[boff + q];
#pragma unroll
      for (int s = 0; s < SPC_SLOTS; ++s)
        if (s == nmine) {
          sp_i[s] = LW + f; sp_j[s] = jj; sp_e[s] = k;
          sp_a1[s] = ca1; sp_a2[s] = ca2; sp_w[s] = cw;
          sp_l1[s] = u1; sp_l2[s] = u2;
        }
      ++nmine;
      // zero the main-loop coefficients of that cell
#pragma unroll
      for (int j = 0; j < CPT; ++j)
        if (j == jj) { ka1[j] = 0.f; ka2[j] = 0.f; kw[j] = 0.f; }
    }
  }
  const int flo = (or0 - R0) * LW, fspan = (or1 - or0) * LW;
  __syncthreads();

  for (int s = 0; s < a.nsteps; ++s) {
    const int m = a.m_hi - s;  // forming lambda_m; gradient of step m-1
    float *P = (s & 1) ? buf1 : buf0;
    float *Q = (s & 1) ? buf0 : buf1;
    const float *segp = a.seg_lap +
                        ((int64_t)b * a.seg_K + (m - 1 - a.seg_n0)) * Nx * (int64_t)LW +
                        (int64_t)R0 * LW;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int f = tid + j * NT;
      const int i = LW + f;
      const float lapw = lap_lds<UNIT_H>(P, i, LW, a.t1, a.t4);
      const float lam = fmaf(-ka2[j], l2[j], fmaf(ka1[j], l1[j], lapw));
      l2[j] = l1[j];
      l1[j] = lam;
      Q[i] = kw[j] * lam;
      if ((unsigned)(f - flo) < (unsigned)fspan) ga[j] = fmaf(segp[f], lam, ga[j]);
    }
#pragma unroll
    for (int sl = 0; sl < SPC_SLOTS; ++sl) {
      if (sp_i[sl] < 0) continue;
      const int i = sp_i[sl];
      const int f = i - LW;
      const int32_t *e = a.spx + sp_e[sl] * SPX_W;
      float gseed = 0.0f;
      if (a.ybar) {
        for (int t = e[4]; t < e[5]; ++t) {
          const int32_t *tp = a.taps + t * TAP_W;
          const int64_t o = ((int64_t)b * a.T + (m - 1)) * a.ntap + tp[3];
          const float gy = a.ybar[o];
          gseed += tp[2] ? 2.0f * a.yraw[o] * gy : gy;
        }
      }
      const float lapw = lap_lds<UNIT_H>(P, i, LW, a.t1, a.t4);
      const float lam =
          fmaf(-sp_a2[sl], sp_l2[sl], fmaf(sp_a1[sl], sp_l1[sl], lapw)) + gseed;
      sp_l2[sl] = sp_l1[sl];
      sp_l1[sl] = lam;
      Q[i] = sp_w[sl] * lam;
      if ((unsigned)(f - flo) < (unsigned)fspan) {
        sp_g[sl] = fmaf(segp[f], lam, sp_g[sl]);
        if (e[2] > 0) a.gsrc[((int64_t)b * a.T + (m - 1)) * a.nsrcslot + e[3]] = lam;
      }
    }
    __syncthreads();
  }
  // write back: lambda_{m_lo} -> out1, lambda_{m_lo+1} -> out2, gradient sums
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int f = tid + j * NT;
    const int r = div_lw(f, LW, invLW);
    const int cc = f - r * LW;
    const int gr = R0 + r;
    bool own = r < RH && cc < RW && gr >= or0 && gr < or1;
#pragma unroll
    for (int sl = 0; sl < SPC_SLOTS; ++sl) own = own && (sp_j[sl] != j);
    if (own) {
      const int64_t q = boff + (int64_t)gr * Ny + cc;
      a.out1[q] = l1[j];
      a.out2[q] = l2[j];
      a.gacc[q] += ga[j];
    }
  }
#pragma unroll
  for (int sl = 0; sl < SPC_SLOTS; ++sl) {
    if (sp_i[sl] < 0) continue;
    const int32_t *e = a.spx + sp_e[sl] * SPX_W;
    if (e[0] < or0 || e[0] >= or1) continue;
    const int64_t q = boff + (int64_t)e[0] * Ny + e[1];
    a.out1[q] = sp_l1[sl];
    a.out2[q] = sp_l2[sl];
    a.gacc[q] += sp_g[sl];
  }
}

// ------------------------------------------------------------------ launchers

// Largest dynamic LDS a workgroup may request on the current device.
static int max_dyn_lds() { return max_dyn_lds_dev(); }

template <int NT, int CPT>
static hipError_t fwd_dispatch(int nl, bool rec, const FwdArgs &a, int nblocks, size_t shmem,
                               hipStream_t s) {
#define FT(NL, REC)                                                                  \
  do {                                                                               \
    auto kfn = k_fwd_tiled<NT, CPT, true, NL, REC>;                                  \
    static std::atomic<uint64_t> done{0};                                            \
    hipError_t e = allow_dyn_lds((const void *)kfn, done);                           \
    if (e != hipSuccess) return e;                                                   \
    hipLaunchKernelGGL(kfn, dim3(nblocks), dim3(NT), shmem, s, a);                   \
  } while (0)
  if (nl == 0) {
    if (rec) FT(0, true); else FT(0, false);
  } else {
    if (rec) FT(8, true); else FT(8, false);
  }
#undef FT
  return hipGetLastError();
}

template <int NT, int CPT>
static hipError_t adj_dispatch(const AdjArgs &a, int nblocks, size_t shmem, hipStream_t s) {
  auto kfn = k_adj_tiled<NT, CPT, true>;
  static std::atomic<uint64_t> done{0};
  hipError_t e = allow_dyn_lds((const void *)kfn, done);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kfn, dim3(nblocks), dim3(NT), shmem, s, a);
  return hipGetLastError();
}

// (NT, CPT) variants compiled into the library.
#define WT_TILED_VARIANTS(X) \
  X(1024, 8)                 \
  X(1024, 12)                \
  X(1024, 16)                \
  X(512, 16)                 \
  X(512, 24)                 \
  X(512, 32)                 \
  X(256, 32)                 \
  X(256, 48)

bool tiled_variant_exists(int nt, int cpt) {
#define X(A, B) if (nt == A && cpt == B) return true;
  WT_TILED_VARIANTS(X)
#undef X
  return false;
}

size_t tiled_shmem_bytes(int nt, int cpt, int LW) {
  return (size_t)2 * (2 * LW + nt * cpt) * sizeof(float);
}

// The tiled kernels are instantiated for h == 1 only (every configuration of
// the reference's studies); other h run the plain kernels.
hipError_t launch_fwd_tiled(int nt, int cpt, bool unit_h, int nl, bool rec, const FwdArgs &a,
                            int LWmax, hipStream_t s) {
  if (!unit_h) return hipErrorInvalidValue;
  int nblocks = a.B * a.ntiles;
  size_t shmem = tiled_shmem_bytes(nt, cpt, LWmax);
  if (shmem > (size_t)max_dyn_lds()) return hipErrorInvalidConfiguration;
#define X(A, B) \
  if (nt == A && cpt == B) return fwd_dispatch<A, B>(nl, rec, a, nblocks, shmem, s);
  WT_TILED_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_adj_tiled(int nt, int cpt, bool unit_h, const AdjArgs &a, int LWmax,
                            hipStream_t s) {
  if (!unit_h) return hipErrorInvalidValue;
  int nblocks = a.B * a.ntiles;
  size_t shmem = tiled_shmem_bytes(nt, cpt, LWmax);
  if (shmem > (size_t)max_dyn_lds()) return hipErrorInvalidConfiguration;
#define X(A, B) \
  if (nt == A && cpt == B) return adj_dispatch<A, B>(a, nblocks, shmem, s);
  WT_TILED_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

}  // namespace wt
