/*
 * wave_oracle.c -- CPU restatement of wavetorch's WaveRNN time loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg may load it.  The product
 * path (wavetorch_amd) never calls into it.
 *
 * What is restated (reference = fancompute/wavetorch v0.2.1):
 *   wo_laplacian_*   operators.py:5-11   5-point stencil, conv2d padding=1
 *                                         (zero Dirichlet), taps h^-2*[0,1,0;1,-4,1;0,1,0]
 *   wo_forward_*     rnn.py:21-72        time loop: cell -> sources -> probes
 *                    cell.py:12-17       _time_step leapfrog
 *                    cell.py:8-9,94-102  saturable damping / Kerr-like speed
 *                    source.py:15-22     additive injection, scale dt**2 = 1.0
 *                    probe.py:14-15,26-27 amplitude / intensity readout
 *   wo_adjoint_f64   cell.py:26-44       TimeStep.backward, chained over the
 *                                         loop the way autograd replays it
 *                                         (discrete adjoint), plus the
 *                                         nonlinear chain rule through
 *                                         cell.py:94-102.
 *
 * fp32 variant: every arithmetic operation is performed in the order and
 * with the rounding the reference's torch ops use (one IEEE rounding per op,
 * no contraction -- build with -ffp-contract=off), so that a kernel that
 * reproduces the same order is bitwise identical:
 *     dtm2 = 1/(dt*dt)   dtm1 = 1/dt   s2 = 2*dtm2          (cell.py:13-14)
 *     A    = 1/(dtm2 + b*dtm1)   Bm = dtm2 - b*dtm1   c2 = c*c
 *     lap  = (((t1*up + t1*lf) + t4*ctr) + t1*rt) + t1*dn   (conv2d order)
 *     y    = A*(((s2*y1) - (Bm*y2)) + (c2*lap))
 * fp64 variant: the same algebra in double precision (error reference).
 *
 * Source / probe descriptions:
 *   src_px : nsrcpx rows of (x, y, source_index); each source's pixels are
 *            deduplicated by the caller (index_put assigns, source.py:20);
 *            sources are applied in increasing source_index order.
 *   taps   : ntap rows of (x, y, kind); kind 0 = WaveProbe (u),
 *            kind 1 = WaveIntensityProbe (u*u).  Output y[b, n, tap].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define IDX3(b, i, j) (((size_t)(b) * Nx + (i)) * Ny + (j))

/* ---------------------------------------------------------------- */
/* Laplacian (operators.py:5-11)                                    */
/* ---------------------------------------------------------------- */
#define DEF_LAP(T, SUF)                                                        \
  static void lap_##SUF(const T *u, T *out, int Nx, int Ny, T t1, T t4) {      \
    for (int i = 0; i < Nx; ++i)                                               \
      for (int j = 0; j < Ny; ++j) {                                           \
        T up = i > 0 ? u[(size_t)(i - 1) * Ny + j] : (T)0;                     \
        T dn = i < Nx - 1 ? u[(size_t)(i + 1) * Ny + j] : (T)0;                \
        T lf = j > 0 ? u[(size_t)i * Ny + j - 1] : (T)0;                       \
        T rt = j < Ny - 1 ? u[(size_t)i * Ny + j + 1] : (T)0;                  \
        T ct = u[(size_t)i * Ny + j];                                          \
        T a = t1 * up;                                                         \
        T s = a + t1 * lf;                                                     \
        s = s + t4 * ct;                                                       \
        s = s + t1 * rt;                                                       \
        s = s + t1 * dn;                                                       \
        out[(size_t)i * Ny + j] = s;                                           \
      }                                                                        \
  }
DEF_LAP(float, f32)
DEF_LAP(double, f64)

/* h**(-2) as torch computes it for a 0-dim tensor: 1/(h*h). */
static float taps_f32(float h) { return 1.0f / (h * h); }
static double taps_f64(double h) { return 1.0 / (h * h); }

void wo_laplacian_f32(const float *u, float *out, int B, int Nx, int Ny,
                      float h) {
  float t1 = taps_f32(h), t4 = -4.0f * t1;
  for (int b = 0; b < B; ++b)
    lap_f32(u + (size_t)b * Nx * Ny, out + (size_t)b * Nx * Ny, Nx, Ny, t1, t4);
}
void wo_laplacian_f64(const double *u, double *out, int B, int Nx, int Ny,
                      double h) {
  double t1 = taps_f64(h), t4 = -4.0 * t1;
  for (int b = 0; b < B; ++b)
    lap_f64(u + (size_t)b * Nx * Ny, out + (size_t)b * Nx * Ny, Nx, Ny, t1, t4);
}

/* ---------------------------------------------------------------- */
/* Forward loop (rnn.py:21-72)                                       */
/* ---------------------------------------------------------------- */
/*
 * nl flags: bit0 = saturable damping active (satdamp_b0 > 0, cell.py:94)
 *           bit1 = Kerr-like speed active   (c_nl != 0,        cell.py:99)
 *           bit2 = clamp c to [cmin, cmax]  (extension; not in the reference)
 * fields_out: if non-NULL, receives u_{n+1} for the steps n with
 *             (n+1) % fstride == 0 as (B, T/fstride, Nx, Ny) (rnn.py:66
 *             output_fields path; fstride 1 = every step).
 */
#define DEF_FWD(T, SUF)                                                        \
  int wo_forward_##SUF(int B, int Nx, int Ny, int Tn, const T *bfield,         \
                       const T *cfield, T dt, T h, const T *x, int nsrcpx,     \
                       const int32_t *src_px, int ntap, const int32_t *taps,   \
                       T *y_out, T *fields_out, int fstride, int nl,           \
                       const T *rho, T sat_b0, T sat_uth, T c_nl, T cmin,      \
                       T cmax) {                                               \
    size_t F = (size_t)Nx * Ny;                                                \
    if (fstride < 1) return -2;                                                \
    const int nsnap = Tn / fstride;                                            \
    T *u1 = (T *)calloc(B * F, sizeof(T));                                     \
    T *u2 = (T *)calloc(B * F, sizeof(T));                                     \
    T *un = (T *)calloc(B * F, sizeof(T));                                     \
    T *lp = (T *)calloc(F, sizeof(T));                                         \
    if (!u1 || !u2 || !un || !lp) {                                            \
      free(u1); free(u2); free(un); free(lp);                                  \
      return -1;                                                               \
    }                                                                          \
    T t1 = taps_##SUF(h), t4 = (T)-4 * t1;                                     \
    T dtm2 = (T)1 / (dt * dt), dtm1 = (T)1 / dt, s2 = (T)2 * dtm2;             \
    int maxsrc = -1;                                                           \
    for (int k = 0; k < nsrcpx; ++k)                                           \
      if (src_px[3 * k + 2] > maxsrc) maxsrc = src_px[3 * k + 2];              \
    for (int n = 0; n < Tn; ++n) {                                             \
      for (int b = 0; b < B; ++b) {                                            \
        const T *y1 = u1 + b * F, *y2 = u2 + b * F;                            \
        lap_##SUF(y1, lp, Nx, Ny, t1, t4);                                     \
        for (size_t q = 0; q < F; ++q) {                                       \
          T bb = bfield[q], cc = cfield[q];                                    \
          if (nl & 1) {                                                        \
            T r = y1[q] / sat_uth;                                             \
            r = fabs(r);                                                       \
            r = r * r;                                                         \
            r = r + (T)1;                                                      \
            r = sat_b0 / r;                                                    \
            bb = bb + rho[q] * r;                                              \
          }                                                                    \
          if (nl & 2) {                                                        \
            T rc = rho[q] * c_nl;                                              \
            cc = cc + rc * (y1[q] * y1[q]);                                    \
          }                                                                    \
          if (nl & 4) {                                                        \
            if (cc < cmin) cc = cmin;                                          \
            if (cc > cmax) cc = cmax;                                          \
          }                                                                    \
          T bd = bb * dtm1;                                                    \
          T A = (T)1 / (dtm2 + bd);                                            \
          T Bm = dtm2 - bd;                                                    \
          T c2 = cc * cc;                                                      \
          T p1 = s2 * y1[q];                                                   \
          T p2 = Bm * y2[q];                                                   \
          T p3 = c2 * lp[q];                                                   \
          un[b * F + q] = A * ((p1 - p2) + p3);                                \
        }                                                                      \
      }                                                                        \
      /* sources, in module order (rnn.py:56-57) */                            \
      for (int s = 0; s <= maxsrc; ++s)                                        \
        for (int k = 0; k < nsrcpx; ++k)                                       \
          if (src_px[3 * k + 2] == s) {                                        \
            size_t q = (size_t)src_px[3 * k] * Ny + src_px[3 * k + 1];         \
            for (int b = 0; b < B; ++b)                                        \
              un[b * F + q] = un[b * F + q] + x[(size_t)b * Tn + n];           \
          }                                                                    \
      /* rotate hidden state (cell.py:107): (h1, h2) <- (y, h1) */             \
      T *tmp = u2; u2 = u1; u1 = un; un = tmp;                                 \
      for (int b = 0; b < B; ++b) {                                            \
        for (int k = 0; k < ntap; ++k) {                                       \
          T v = u1[b * F + (size_t)taps[3 * k] * Ny + taps[3 * k + 1]];        \
          if (taps[3 * k + 2] == 1) v = v * v;                                 \
          if (y_out) y_out[((size_t)b * Tn + n) * ntap + k] = v;               \
        }                                                                      \
        if (fields_out && (n + 1) % fstride == 0)                              \
          memcpy(fields_out + ((size_t)b * nsnap + (n + 1) / fstride - 1) * F,   \
                 u1 + b * F, F * sizeof(T));                                   \
      }                                                                        \
    }                                                                          \
    free(u1); free(u2); free(un); free(lp);                                    \
    return 0;                                                                  \
  }
DEF_FWD(float, f32)
DEF_FWD(double, f64)

/* ---------------------------------------------------------------- */
/* Discrete adjoint in fp64 (cell.py:26-44 chained over rnn.py:50-70) */
/* ---------------------------------------------------------------- */
/*
 * Inputs as wo_forward_f64 plus
 *   ybar   : dJ/dy, (B,T,ntap)  (may be NULL if fbar is given)
 *   fbar   : dJ/d(fields), (B,T,Nx,Ny) for the output_fields path (may be NULL)
 * Outputs (any may be NULL):
 *   gc     : dJ/dc_linear (Nx,Ny), summed over batch
 *   gb     : dJ/db        (Nx,Ny), summed over batch
 *   grho   : dJ/drho through the nonlinear terms only (Nx,Ny)
 *   gx     : dJ/dx (B,T)
 *   u32    : NULL, or the fields u_1..u_T (B,T,Nx,Ny) of a forward computed
 *            elsewhere (wo_forward_f32 with fstride 1): the reverse sweep then
 *            runs in fp64 at that trajectory instead of its own fp64 forward,
 *            i.e. the exact adjoint of the computed forward (the step's
 *            Jacobian depends on u through the probe seeds 2*u*ybar, the c
 *            gradient's lap(u) and the nonlinear coefficients)
 * Memory: keeps every forward field (small problems only).
 */
int wo_adjoint_f64(int B, int Nx, int Ny, int Tn, const double *bfield,
                   const double *cfield, double dt, double h, const double *x,
                   int nsrcpx, const int32_t *src_px, int ntap,
                   const int32_t *taps, int nl, const double *rho,
                   double sat_b0, double sat_uth, double c_nl, double cmin,
                   double cmax, const double *ybar, const double *fbar,
                   const float *u32, double *gc, double *gb, double *grho,
                   double *gx) {
  size_t F = (size_t)Nx * Ny;
  /* U[n] = u_n for n = -1..Tn  -> index n+1 */
  double *U = (double *)calloc((size_t)(Tn + 2) * B * F, sizeof(double));
  double *lp = (double *)calloc(F, sizeof(double));
  double *l1 = (double *)calloc(B * F, sizeof(double)); /* lambda_{m+1} */
  double *l2 = (double *)calloc(B * F, sizeof(double)); /* lambda_{m+2} */
  double *l0 = (double *)calloc(B * F, sizeof(double));
  double *wl = (double *)calloc(F, sizeof(double));
  double *lw = (double *)calloc(F, sizeof(double));
  /* -A*Bm factor of each step, in a ring indexed by step parity: iteration m
   * reads step m+1's factor (slot (m+1)&1) before its gradient block writes
   * step m-1's factor into the same slot. */
  double *a2n = (double *)calloc(2 * B * F, sizeof(double));
  if (!U || !lp || !l1 || !l2 || !l0 || !wl || !lw || !a2n) {
    free(U); free(lp); free(l1); free(l2); free(l0); free(wl); free(lw);
    free(a2n);
    return -1;
  }
#define UU(n, b) (U + ((size_t)((n) + 1) * B + (b)) * F)
  double t1 = taps_f64(h), t4 = -4.0 * t1;
  double dtm2 = 1.0 / (dt * dt), dtm1 = 1.0 / dt, s2 = 2.0 * dtm2;
  int maxsrc = -1;
  for (int k = 0; k < nsrcpx; ++k)
    if (src_px[3 * k + 2] > maxsrc) maxsrc = src_px[3 * k + 2];
  /* forward, storing every field (or the given trajectory, upcast) */
  if (u32)
    for (int n = 0; n < Tn; ++n)
      for (int b = 0; b < B; ++b) {
        const float *src = u32 + ((size_t)b * Tn + n) * F;
        double *dst = UU(n + 1, b);
        for (size_t q = 0; q < F; ++q) dst[q] = (double)src[q];
      }
  for (int n = 0; n < Tn && !u32; ++n) {
    for (int b = 0; b < B; ++b) {
      const double *y1 = UU(n, b), *y2 = UU(n - 1, b);
      double *yn = UU(n + 1, b);
      lap_f64(y1, lp, Nx, Ny, t1, t4);
      for (size_t q = 0; q < F; ++q) {
        double bb = bfield[q], cc = cfield[q];
        if (nl & 1) bb += rho[q] * sat_b0 / (1.0 + (y1[q] / sat_uth) * (y1[q] / sat_uth));
        if (nl & 2) cc += rho[q] * c_nl * y1[q] * y1[q];
        if (nl & 4) { if (cc < cmin) cc = cmin; if (cc > cmax) cc = cmax; }
        double A = 1.0 / (dtm2 + bb * dtm1), Bm = dtm2 - bb * dtm1;
        yn[q] = A * (s2 * y1[q] - Bm * y2[q] + cc * cc * lp[q]);
      }
    }
    for (int s = 0; s <= maxsrc; ++s)
      for (int k = 0; k < nsrcpx; ++k)
        if (src_px[3 * k + 2] == s) {
          size_t q = (size_t)src_px[3 * k] * Ny + src_px[3 * k + 1];
          for (int b = 0; b < B; ++b) UU(n + 1, b)[q] += x[(size_t)b * Tn + n];
        }
  }
  if (gc) memset(gc, 0, F * sizeof(double));
  if (gb) memset(gb, 0, F * sizeof(double));
  if (grho) memset(grho, 0, F * sizeof(double));
  if (gx) memset(gx, 0, (size_t)B * Tn * sizeof(double));
  /* reverse sweep.  Iteration m (m = Tn..1) forms lambda_m = dJ/du_m from
   *   - the direct probe / field seed at output index m-1,
   *   - step m   (u_{m+1} = S(u_m, u_{m-1})) through lambda_{m+1},
   *   - step m+1 (u_{m+2} = S(u_{m+1}, u_m)) through lambda_{m+2}  (the -A*Bm term),
   * and, with lambda_m in hand, the parameter gradients of step m-1, which
   * produced u_m = v_m + sources.                                            */
  for (int m = Tn; m >= 1; --m) {
    for (int b = 0; b < B; ++b) {
      const double *um = UU(m, b), *um1 = UU(m - 1, b);
      double *L1 = l1 + b * F, *L2 = l2 + b * F, *L0 = l0 + b * F;
      /* contributions of step m (needs step-m coefficients, depends on u_m) */
      lap_f64(um, lp, Nx, Ny, t1, t4);
      for (size_t q = 0; q < F; ++q) {
        double u = um[q];
        double bb = bfield[q], cc = cfield[q], dbdu = 0.0, dcdu = 0.0;
        if (nl & 1) {
          double r = u / sat_uth, d = 1.0 + r * r;
          bb += rho[q] * sat_b0 / d;
          dbdu = rho[q] * sat_b0 * (-2.0 * r / sat_uth) / (d * d);
        }
        if (nl & 2) { cc += rho[q] * c_nl * u * u; dcdu = 2.0 * rho[q] * c_nl * u; }
        if (nl & 4) {
          if (cc < cmin) { cc = cmin; dcdu = 0.0; }
          if (cc > cmax) { cc = cmax; dcdu = 0.0; }
        }
        double A = 1.0 / (dtm2 + bb * dtm1), Bm = dtm2 - bb * dtm1;
        double S = s2 * u - Bm * um1[q] + cc * cc * lp[q];
        double v = A * S;
        double dvdb = A * dtm1 * (um1[q] - v);
        double dvdc = A * 2.0 * cc * lp[q];
        wl[q] = A * cc * cc * L1[q];
        /* pointwise part (s2*A + nonlinear coefficient derivatives) */
        L0[q] = (s2 * A + dvdb * dbdu + dvdc * dcdu) * L1[q];
      }
      lap_f64(wl, lw, Nx, Ny, t1, t4);
      const double *a2r = a2n + ((size_t)((m + 1) & 1) * B + b) * F;
      for (size_t q = 0; q < F; ++q) {
        L0[q] += lw[q] - a2r[q] * L2[q];
      }
      /* direct seed: y[m-1] reads u_m */
      if (ybar)
        for (int k = 0; k < ntap; ++k) {
          size_t q = (size_t)taps[3 * k] * Ny + taps[3 * k + 1];
          double g = ybar[((size_t)b * Tn + (m - 1)) * ntap + k];
          L0[q] += (taps[3 * k + 2] == 1) ? 2.0 * um[q] * g : g;
        }
      if (fbar) {
        const double *fb = fbar + ((size_t)b * Tn + (m - 1)) * F;
        for (size_t q = 0; q < F; ++q) L0[q] += fb[q];
      }
      /* gradients of step m-1 (u_m = v_m + src; uses lambda_m = L0) */
      const double *umm1 = UU(m - 1, b), *umm2 = UU(m - 2, b);
      lap_f64(umm1, lp, Nx, Ny, t1, t4);
      for (size_t q = 0; q < F; ++q) {
        double u = umm1[q];
        double bb = bfield[q], cc = cfield[q], sd = 0.0, c2raw = 0.0;
        int clamped = 0;
        if (nl & 1) { double r = u / sat_uth; sd = sat_b0 / (1.0 + r * r); bb += rho[q] * sd; }
        if (nl & 2) { c2raw = c_nl * u * u; cc += rho[q] * c2raw; }
        if (nl & 4) {
          if (cc < cmin) { cc = cmin; clamped = 1; }
          if (cc > cmax) { cc = cmax; clamped = 1; }
        }
        double A = 1.0 / (dtm2 + bb * dtm1), Bm = dtm2 - bb * dtm1;
        double v = A * (s2 * u - Bm * umm2[q] + cc * cc * lp[q]);
        double dvdb = A * dtm1 * (umm2[q] - v);
        double dvdc = clamped ? 0.0 : A * 2.0 * cc * lp[q];
        double lam = L0[q];
        if (gc) gc[q] += dvdc * lam;
        if (gb) gb[q] += dvdb * lam;
        if (grho) grho[q] += (dvdb * sd + dvdc * c2raw) * lam;
        /* A*Bm of step m-1 multiplies lambda_m in iteration m-2 */
        a2n[((size_t)((m - 1) & 1) * B + b) * F + q] = A * Bm;
      }
      if (gx)
        for (int k = 0; k < nsrcpx; ++k) {
          size_t q = (size_t)src_px[3 * k] * Ny + src_px[3 * k + 1];
          gx[(size_t)b * Tn + (m - 1)] += L0[q];
        }
    }
    /* shift lambdas: (l2, l1) <- (l1, l0) */
    double *tmp = l2; l2 = l1; l1 = l0; l0 = tmp;
  }
  free(U); free(lp); free(l1); free(l2); free(l0); free(wl); free(lw); free(a2n);
#undef UU
  return 0;
}
