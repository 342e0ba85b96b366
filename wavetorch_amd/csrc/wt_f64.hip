// wt_f64.hip -- the WaveRNN time loop in double precision (gfx950).
//
// The reference computes in float64 when a user calls set_dtype('float64')
// (utils.py:14-20; its own example configurations train that way).  This path
// keeps every field u_n of the loop in a caller-provided tape -- what the
// reference's autograd tape holds, (T + 2) x B x Nx x Ny doubles -- and runs
// one launch per step:
//
//   k64_fwd      one step of rnn.py:50-64: the leapfrog update of
//                cell.py:12-17 in the reference's operation order (with the
//                nonlinear b and c of cell.py:94-102), sources (source.py:15-22)
//                and probes (probe.py:14-27), compiled without contraction
//   k64_adj      adjoint iteration m: step m's coefficients at u_m give the
//                pointwise term (s2 A + dv/db db/du + dv/dc dc/du) *
//                lambda_{m+1} and, staged with a halo in LDS, w_m * lambda_{m+1};
//                lambda_m = that term + L(w lambda_{m+1})
//                - (A Bm)_{m+1} lambda_{m+2} + probe seeds, then step m-1's
//                parameter gradients (dJ/dc_linear, dJ/db, dJ/drho) and
//                lambda_m at source pixels (dJ/dx)
//
// i.e. the transpose of cell.py:12-17 that autograd replays from
// cell.py:26-44 (with the derivatives of saturable_damping, cell.py:8-9, and
// of the Kerr-like speed), the same recurrence as the fp32 adjoints.  Plan
// tables (sources, probes) come from a plan made by wt_plan_create; the
// double-precision scalars (dt, h, nonlinear coefficients) are passed
// explicitly, since wt_config holds them as float.
#include <algorithm>

#include "wt_internal.h"

namespace wt {

struct F64Args {
  int B, Nx, Ny, T, ntap, nslot, nlf;
  const int32_t *spx, *taps, *spmap;
  const double *b, *c, *rho, *x;
  double dtm1, dtm2, s2, t1, t4;
  double sat_b0, sat_uth, c_nl, cmin, cmax;
};

// 5-point Laplacian with zero padding, in conv2d's accumulation order
// (operators.py:5-11): (((t1*up + t1*lf) + t4*ct) + t1*rt) + t1*dn.
__device__ __forceinline__ double lap64(const double *__restrict__ u, int64_t q, int i, int j,
                                        int Nx, int Ny, double t1, double t4) {
  const double up = i > 0 ? u[q - Ny] : 0.0;
  const double dn = i < Nx - 1 ? u[q + Ny] : 0.0;
  const double lf = j > 0 ? u[q - 1] : 0.0;
  const double rt = j < Ny - 1 ? u[q + 1] : 0.0;
  double s = t1 * up;
  s = s + t1 * lf;
  s = s + t4 * u[q];
  s = s + t1 * rt;
  return s + t1 * dn;
}

// Material of a cell at field value u (cell.py:94-102; the clamp is an
// extension).  Returns b and c; sd = saturable term, c2raw = c_nl u^2.
__device__ __forceinline__ void mat64(const F64Args &a, int64_t q, double u, double &bb,
                                      double &cc, double &sd, double &c2raw, bool &clamped) {
  bb = a.b[q];
  cc = a.c[q];
  sd = 0.0;
  c2raw = 0.0;
  clamped = false;
  if (a.nlf & 1) {
    // b0 / (1 + |u/uth|^2), as saturable_damping (cell.py:8-9)
    double r = u / a.sat_uth;
    r = fabs(r);
    r = r * r;
    r = r + 1.0;
    sd = a.sat_b0 / r;
    bb = bb + a.rho[q] * sd;
  }
  if (a.nlf & 2) {
    // c_linear + rho * c_nl * u^2 (cell.py:99-100)
    const double rc = a.rho[q] * a.c_nl;
    c2raw = a.c_nl * (u * u);
    cc = cc + rc * (u * u);
  }
  if (a.nlf & 4) {
    if (cc < a.cmin) { cc = a.cmin; clamped = true; }
    if (cc > a.cmax) { cc = a.cmax; clamped = true; }
  }
}

// One forward step: u_{n+1} = S(u_n, u_{n-1}) + sources; probe outputs.
__global__ void k64_fwd(F64Args a, const double *__restrict__ u1, const double *__restrict__ u2,
                        double *__restrict__ un, double *__restrict__ y,
                        double *__restrict__ fields, int fstride, int n) {
  const int Nx = a.Nx, Ny = a.Ny;
  const int64_t F = (int64_t)Nx * Ny;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= F * a.B) return;
  const int b = (int)(gid / F);
  const int64_t q = gid - (int64_t)b * F;
  const int i = (int)(q / Ny), j = (int)(q - (int64_t)i * Ny);
  const double *u = u1 + (int64_t)b * F;
  const double ct = u[q];
  const double lap = lap64(u, q, i, j, Nx, Ny, a.t1, a.t4);
  double bb, cc, sd, c2raw;
  bool clamped;
  mat64(a, q, ct, bb, cc, sd, c2raw, clamped);
  const double bd = bb * a.dtm1;
  const double A = 1.0 / (a.dtm2 + bd);
  const double Bm = a.dtm2 - bd;
  const double p1 = a.s2 * ct;
  const double p2 = Bm * u2[gid];
  const double p3 = (cc * cc) * lap;
  double v = A * ((p1 - p2) + p3);
  const int sp = a.spmap[q];
  if (sp >= 0) {
    const int32_t *e = a.spx + sp * SPX_W;
    const double xv = a.x[(int64_t)b * a.T + n];
    for (int k = 0; k < e[2]; ++k) v = v + xv;  // one add per source (rnn.py:56-57)
    if (y)
      for (int t = e[4]; t < e[5]; ++t) {
        const int32_t *tp = a.taps + t * TAP_W;
        y[((int64_t)b * a.T + n) * a.ntap + tp[3]] = tp[2] ? v * v : v;
      }
  }
  un[gid] = v;
  if (fields && ((n + 1) % fstride) == 0) {
    const int64_t nf = a.T / fstride;
    fields[((int64_t)b * nf + (n + 1) / fstride - 1) * F + q] = v;
  }
}

// Adjoint iteration m in one launch: lambda_m, then step m-1's gradients.
// A workgroup owns a tile of K64_TX rows x K64_TY columns of one sample and
// stages w_m lambda_{m+1} (w = A c^2 at u_m, pointwise) on the tile and a
// one-cell halo in LDS, so L(w lambda_{m+1}) needs no second launch.
//   lambda_m = (s2 A + dv/db db/du + dv/dc dc/du) lambda_{m+1}
//              + L(w lambda_{m+1}) - (A Bm)_{m+1} lambda_{m+2}
//              + probe seeds + field seeds
// (A Bm)_{m+1}, the factor of u_m in step m+1, is re-formed from u_{m+1}
// (saturable damping) or from b alone; `up1` is NULL when step m+1 does not
// exist (m >= T-1, where lambda_{m+2} = 0).
constexpr int K64_TX = 16, K64_TY = 32, K64_NT = K64_TX * K64_TY;
constexpr int K64_HW = K64_TY + 2, K64_HN = (K64_TX + 2) * K64_HW;

__global__ __launch_bounds__(K64_NT) void k64_adj(
    F64Args a, int tiles_x, int tiles_y, const double *__restrict__ up1,
    const double *__restrict__ um, const double *__restrict__ umm1,
    const double *__restrict__ umm2, const double *__restrict__ L1,
    const double *__restrict__ L2, const double *__restrict__ ybar,
    const double *__restrict__ fbar, double *__restrict__ L0, double *__restrict__ gacc_c,
    double *__restrict__ gacc_b, double *__restrict__ gacc_r, double *__restrict__ gsrc,
    int m) {
  __shared__ double wls[K64_HN];
  const int Nx = a.Nx, Ny = a.Ny;
  const int64_t F = (int64_t)Nx * Ny;
  // neighbouring tiles of a sample on one XCD: their halo lines meet in its L2
  int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int tcol = blk % tiles_y;
  blk /= tiles_y;
  const int trow = blk % tiles_x;
  const int b = blk / tiles_x;
  const int i0 = trow * K64_TX, j0 = tcol * K64_TY;
  const int64_t boff = (int64_t)b * F;
  // w_m lambda_{m+1} on the tile and its halo (zero outside the domain)
  for (int p = threadIdx.x; p < K64_HN; p += K64_NT) {
    const int pi = p / K64_HW, pj = p - pi * K64_HW;
    const int i = i0 - 1 + pi, j = j0 - 1 + pj;
    double w = 0.0;
    if (i >= 0 && i < Nx && j >= 0 && j < Ny) {
      const int64_t q = (int64_t)i * Ny + j;
      double bb, cc, sd, c2raw;
      bool clamped;
      mat64(a, q, um[boff + q], bb, cc, sd, c2raw, clamped);
      const double bd = bb * a.dtm1;
      const double A = 1.0 / (a.dtm2 + bd);
      w = A * (cc * cc) * L1[boff + q];
    }
    wls[p] = w;
  }
  __syncthreads();
  const int li = threadIdx.x / K64_TY, lj = threadIdx.x - li * K64_TY;
  const int i = i0 + li, j = j0 + lj;
  if (i >= Nx || j >= Ny) return;
  const int64_t q = (int64_t)i * Ny + j, gid = boff + q;
  // step m at u_m, u_{m-1}: the pointwise term
  double lam;
  {
    const double *u = um + boff;
    const double uu = u[q];
    const double lap = lap64(u, q, i, j, Nx, Ny, a.t1, a.t4);
    double bb, cc, sd, c2raw;
    bool clamped;
    mat64(a, q, uu, bb, cc, sd, c2raw, clamped);
    double dbdu = 0.0, dcdu = 0.0;
    if (a.nlf & 1) {
      const double r = uu / a.sat_uth, d = 1.0 + r * r;
      dbdu = a.rho[q] * a.sat_b0 * (-2.0 * r / a.sat_uth) / (d * d);
    }
    if ((a.nlf & 2) && !clamped) dcdu = 2.0 * a.rho[q] * a.c_nl * uu;
    const double bd = bb * a.dtm1;
    const double A = 1.0 / (a.dtm2 + bd), Bm = a.dtm2 - bd;
    const double um1 = umm1[gid];
    const double v = A * ((a.s2 * uu - Bm * um1) + (cc * cc) * lap);
    const double dvdb = A * a.dtm1 * (um1 - v);
    const double dvdc = A * 2.0 * cc * lap;
    lam = (a.s2 * A + dvdb * dbdu + dvdc * dcdu) * L1[gid];
  }
  // + L(w lambda_{m+1}), the Laplacian's order of lap64 on the staged values
  {
    const int c0 = (li + 1) * K64_HW + (lj + 1);
    double s = a.t1 * wls[c0 - K64_HW];
    s = s + a.t1 * wls[c0 - 1];
    s = s + a.t4 * wls[c0];
    s = s + a.t1 * wls[c0 + 1];
    lam = lam + (s + a.t1 * wls[c0 + K64_HW]);
  }
  if (up1) {  // - (A Bm)_{m+1} lambda_{m+2}
    double bb = a.b[q];
    if (a.nlf & 1) {
      double r = up1[gid] / a.sat_uth;
      r = fabs(r);
      r = r * r;
      r = r + 1.0;
      bb = bb + a.rho[q] * (a.sat_b0 / r);
    }
    const double bd = bb * a.dtm1;
    const double A = 1.0 / (a.dtm2 + bd), Bm = a.dtm2 - bd;
    lam = lam - (A * Bm) * L2[gid];
  }
  const int sp = a.spmap[q];
  if (sp >= 0 && ybar) {
    // direct seeds: y[m-1] reads u_m (probe.py:14-27)
    const int32_t *e = a.spx + sp * SPX_W;
    for (int t = e[4]; t < e[5]; ++t) {
      const int32_t *tp = a.taps + t * TAP_W;
      const double g = ybar[((int64_t)b * a.T + (m - 1)) * a.ntap + tp[3]];
      lam += tp[2] ? 2.0 * um[gid] * g : g;
    }
  }
  if (fbar) lam += fbar[((int64_t)b * a.T + (m - 1)) * F + q];
  L0[gid] = lam;
  // step m-1 produced u_m = v_m + sources from (u_{m-1}, u_{m-2})
  const double *u = umm1 + boff;
  const double uu = u[q];
  const double lap = lap64(u, q, i, j, Nx, Ny, a.t1, a.t4);
  double bb, cc, sd, c2raw;
  bool clamped;
  mat64(a, q, uu, bb, cc, sd, c2raw, clamped);
  const double bd = bb * a.dtm1;
  const double A = 1.0 / (a.dtm2 + bd), Bm = a.dtm2 - bd;
  const double v = A * ((a.s2 * uu - Bm * umm2[gid]) + (cc * cc) * lap);
  const double dvdb = A * a.dtm1 * (umm2[gid] - v);
  const double dvdc = clamped ? 0.0 : A * 2.0 * cc * lap;
  gacc_c[gid] += dvdc * lam;
  if (gacc_b) gacc_b[gid] += dvdb * lam;
  if (gacc_r) gacc_r[gid] += (dvdb * sd + dvdc * c2raw) * lam;
  if (sp >= 0) {
    const int32_t *e = a.spx + sp * SPX_W;
    if (e[2] > 0) gsrc[((int64_t)b * a.T + (m - 1)) * a.nslot + e[3]] = lam;
  }
}

__global__ void k64_fill(double *__restrict__ p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// out[q] = sum_b acc[b, q] in fixed order (autograd's reduction of the
// per-sample gradient to the (Nx, Ny) parameter)
__global__ void k64_batch_sum(const double *__restrict__ acc, int B, int64_t F,
                              double *__restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= F) return;
  double s = 0.0;
  for (int b = 0; b < B; ++b) s += acc[(int64_t)b * F + q];
  out[q] = s;
}

// dJ/dx[b, n] = sum over source pixels of multiplicity * lambda_{n+1}[pixel]
__global__ void k64_finalize_gx(const double *__restrict__ gsrc, int B, int T, int nslot,
                                const int32_t *__restrict__ mult, double *__restrict__ gx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * T) return;
  double s = 0.0;
  for (int k = 0; k < nslot; ++k) s += (double)mult[k] * gsrc[i * nslot + k];
  gx[i] = s;
}

static inline unsigned nblk64(int64_t n) { return (unsigned)((n + 255) / 256); }

static hipError_t fill64(double *p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k64_fill, dim3(g), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

// Workspace layout (doubles).  saved: the tape U[n + 1] = u_n, n = -1..T.
// temp (forward without gradients): three rotating fields; (adjoint): three
// lambdas, three gradient accumulators, and lambda at source slots
// (B, T, nslot).
struct Ws64 {
  int64_t BF, tape, fwd_temp, adj_temp;
};
static Ws64 ws64(const PlanView &v) {
  Ws64 w;
  w.BF = (int64_t)v.B * v.Nx * v.Ny;
  w.tape = (int64_t)(v.T + 2) * w.BF;
  w.fwd_temp = 3 * w.BF;
  w.adj_temp = 6 * w.BF + (int64_t)v.B * v.T * std::max(v.nslot, 1);
  return w;
}

static F64Args args64(const PlanView &v, const wt_scalars_f64 &sc, const double *b,
                      const double *c, const double *rho, const double *x) {
  F64Args a{};
  a.B = v.B; a.Nx = v.Nx; a.Ny = v.Ny; a.T = v.T; a.ntap = v.ntap;
  a.nslot = std::max(v.nslot, 1); a.nlf = v.nlf;
  a.spx = v.spx; a.taps = v.taps; a.spmap = v.spmap;
  a.b = b; a.c = c; a.rho = rho; a.x = x;
  // dt**-2 == 1/(dt*dt), dt**-1 == 1/dt, 2/dt**2 (cell.py:13-14); h**(-2)
  // == 1/(h*h) (operators.py:7), for the float64 0-dim tensors
  a.dtm2 = 1.0 / (sc.dt * sc.dt);
  a.dtm1 = 1.0 / sc.dt;
  a.s2 = 2.0 * a.dtm2;
  a.t1 = 1.0 / (sc.h * sc.h);
  a.t4 = -4.0 * a.t1;
  a.sat_b0 = sc.sat_b0; a.sat_uth = sc.sat_uth; a.c_nl = sc.c_nl;
  a.cmin = sc.c_min; a.cmax = sc.c_max;
  return a;
}

}  // namespace wt

using namespace wt;

extern "C" int wt_workspace_bytes_f64(const void *plan, int32_t with_grad, int64_t *saved,
                                      int64_t *temp) {
  PlanView v;
  if (!plan || !saved || !temp || plan_view(plan, &v) != WT_OK) {
    set_error("wt_workspace_bytes_f64: invalid argument");
    return WT_EINVAL;
  }
  const Ws64 w = ws64(v);
  *saved = with_grad ? w.tape * (int64_t)sizeof(double) : 0;
  *temp = (with_grad ? std::max(w.fwd_temp, w.adj_temp) : w.fwd_temp) * (int64_t)sizeof(double);
  return WT_OK;
}

extern "C" int wt_loop_fwd_f64(void *plan, const wt_scalars_f64 *sc, const double *b,
                               const double *c, const double *rho, const double *x, double *y,
                               double *fields, int32_t field_stride, int32_t with_grad,
                               void *saved, void *temp, void *stream) {
  PlanView v;
  if (!plan || !sc || !b || !c || !x || (with_grad ? !saved : !temp) ||
      plan_view(plan, &v) != WT_OK || ((v.nlf & 3) && !rho)) {
    set_error("wt_loop_fwd_f64: invalid argument");
    return WT_EINVAL;
  }
  if (field_stride < 1 || (fields && v.T % field_stride != 0)) {
    set_error("wt_loop_fwd_f64: field_stride %d must divide T = %d", field_stride, v.T);
    return WT_EINVAL;
  }
  if (!(sc->dt > 0) || !(sc->h > 0) || ((v.nlf & 1) && !(sc->sat_uth != 0))) {
    set_error("wt_loop_fwd_f64: invalid scalars (dt %g, h %g, sat_uth %g)", sc->dt, sc->h,
              sc->sat_uth);
    return WT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const F64Args a = args64(v, *sc, b, c, rho, x);
  const Ws64 w = ws64(v);
  const unsigned g = nblk64(w.BF);
  if (with_grad) {
    double *U = (double *)saved;
    WT_HIP(fill64(U, 2 * w.BF, 0.0, s));  // u_{-1} = u_0 = 0 (rnn.py:39-41)
    for (int n = 0; n < v.T; ++n) {
      hipLaunchKernelGGL(k64_fwd, dim3(g), dim3(256), 0, s, a, U + (int64_t)(n + 1) * w.BF,
                         U + (int64_t)n * w.BF, U + (int64_t)(n + 2) * w.BF, y, fields,
                         (int)field_stride, n);
      WT_HIP(hipGetLastError());
    }
  } else {
    double *st = (double *)temp;
    WT_HIP(fill64(st, 2 * w.BF, 0.0, s));
    double *p1 = st, *p2 = st + w.BF, *p3 = st + 2 * w.BF;  // u_n, u_{n-1}, u_{n+1}
    for (int n = 0; n < v.T; ++n) {
      hipLaunchKernelGGL(k64_fwd, dim3(g), dim3(256), 0, s, a, p1, p2, p3, y, fields,
                         (int)field_stride, n);
      WT_HIP(hipGetLastError());
      double *t = p2; p2 = p1; p1 = p3; p3 = t;
    }
  }
  return WT_OK;
}

extern "C" int wt_loop_adj_f64(void *plan, const wt_scalars_f64 *sc, const double *b,
                               const double *c, const double *rho, const double *ybar,
                               const double *fbar, void *saved, void *temp, double *grad_c,
                               double *grad_b, double *grad_rho, double *grad_x, void *stream) {
  PlanView v;
  if (!plan || !sc || !b || !c || !saved || !temp || plan_view(plan, &v) != WT_OK ||
      ((v.nlf & 3) && !rho)) {
    set_error("wt_loop_adj_f64: invalid argument");
    return WT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const F64Args a = args64(v, *sc, b, c, rho, nullptr);
  const Ws64 w = ws64(v);
  const int64_t BF = w.BF, F = (int64_t)v.Nx * v.Ny;
  const double *U = (const double *)saved;
  auto Un = [&](int n) { return U + (int64_t)(n + 1) * BF; };  // u_n, n >= -1
  double *t = (double *)temp;
  double *lam[3] = {t, t + BF, t + 2 * BF};
  double *gc = t + 3 * BF, *gb = t + 4 * BF, *gr = t + 5 * BF;
  double *gsrc = t + 6 * BF;
  WT_HIP(fill64(t, w.adj_temp, 0.0, s));  // lambda_{T+1} = lambda_{T+2} = 0
  const int tiles_x = (v.Nx + K64_TX - 1) / K64_TX, tiles_y = (v.Ny + K64_TY - 1) / K64_TY;
  const unsigned g = (unsigned)((int64_t)v.B * tiles_x * tiles_y);
  double *l1 = lam[0], *l2 = lam[1], *l0 = lam[2];
  for (int m = v.T; m >= 1; --m) {
    // step m+1 (factor of u_m in u_{m+2}) exists for m + 1 <= T - 1
    const double *up1 = m + 1 <= v.T - 1 ? Un(m + 1) : nullptr;
    hipLaunchKernelGGL(k64_adj, dim3(g), dim3(K64_NT), 0, s, a, tiles_x, tiles_y, up1, Un(m),
                       Un(m - 1), Un(m - 2), l1, l2, ybar, fbar, l0, gc,
                       grad_b ? gb : nullptr, (v.nlf & 3) ? gr : nullptr, gsrc, m);
    WT_HIP(hipGetLastError());
    double *tmp = l2; l2 = l1; l1 = l0; l0 = tmp;  // (l2, l1) <- (l1, l0)
  }
  const unsigned gf = nblk64(F);
  if (grad_c) {
    hipLaunchKernelGGL(k64_batch_sum, dim3(gf), dim3(256), 0, s, gc, v.B, F, grad_c);
    WT_HIP(hipGetLastError());
  }
  if (grad_b) {  // dJ/db of the linear cell as well (b enters A and Bm)
    hipLaunchKernelGGL(k64_batch_sum, dim3(gf), dim3(256), 0, s, gb, v.B, F, grad_b);
    WT_HIP(hipGetLastError());
  }
  if (grad_rho) {
    if (v.nlf & 3) {
      hipLaunchKernelGGL(k64_batch_sum, dim3(gf), dim3(256), 0, s, gr, v.B, F, grad_rho);
      WT_HIP(hipGetLastError());
    } else {
      WT_HIP(fill64(grad_rho, F, 0.0, s));
    }
  }
  if (grad_x) {
    if (v.nslot > 0) {
      hipLaunchKernelGGL(k64_finalize_gx, dim3(nblk64((int64_t)v.B * v.T)), dim3(256), 0, s,
                         gsrc, v.B, v.T, v.nslot, v.mult, grad_x);
      WT_HIP(hipGetLastError());
    } else {
      WT_HIP(fill64(grad_x, (int64_t)v.B * v.T, 0.0, s));
    }
  }
  return WT_OK;
}
