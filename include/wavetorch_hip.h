/*
 * wavetorch_hip.h -- C ABI of the MI355X (gfx950) WaveRNN time-loop library
 * (libwavetorch_hip.so).
 *
 * The reference (fancompute/wavetorch v0.2.1) is pure Python/PyTorch; its hot
 * path is the per-step torch-op sequence driven from rnn.py.  Each entry point
 * below replaces a named piece of that path; the citation on each declaration
 * gives the reference file:line it stands in for.  The Python package
 * wavetorch_amd binds these with ctypes (see wavetorch_amd/_lib.py) behind the
 * reference's own nn.Module API.
 *
 * Conventions
 *   - All tensors are caller-owned device memory, fp32, C-contiguous:
 *     fields (B, Nx, Ny) with Ny fastest; signals (B, T); probe series
 *     (B, T, ntap).  The library never frees caller memory.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Every call is asynchronous on that stream unless stated otherwise.
 *   - Every function returns an int status: 0 on success, a negative
 *     wt_status code otherwise; it never throws or aborts.  The message of
 *     the last failure on the calling thread is available from
 *     wt_last_error().
 *   - Reentrant: no global mutable state except the per-thread error string,
 *     the per-device kernel-attribute cache and wt_autotune's result cache
 *     (both mutex protected), and the static device workspace of
 *     wt_loss_power_ce_static (shared by that entry point's calls on a device).
 */
#ifndef WAVETORCH_HIP_H
#define WAVETORCH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WT_ABI_VERSION 2

enum wt_status {
  WT_OK = 0,
  WT_EINVAL = -1,       /* bad argument: null pointer, bad shape, bad index */
  WT_ECFL = -2,         /* dt violates the CFL bound (cell.py:66-73) */
  WT_EHIP = -3,         /* a HIP runtime call failed */
  WT_EUNSUPPORTED = -4, /* shape/plan not supported by any kernel variant */
  WT_ENOMEM = -5,       /* workspace too small */
};

/* Version of this ABI (WT_ABI_VERSION). */
int wt_abi_version(void);

/* Copies the last error message of the calling thread into buf (always
 * NUL-terminated when len > 0) and returns its full length. */
int wt_last_error(char *buf, size_t len);

/* Sets flag[0] = 1 (device int32) when any of the n fp32 values at p is NaN
 * or infinite, and leaves it unchanged otherwise: one kernel launch on
 * `stream`, no host synchronisation, capturable into a graph.  The caller
 * zeroes the flag.  No counterpart in the reference, where an overflowing
 * field or gradient shows up only in the loss it prints (train.py:110-113);
 * wavetorch_amd runs it on the loop's outputs and gradients when a model's
 * check_finite is set. */
int wt_check_finite(const float *p, int64_t n, int32_t *flag, void *stream);
/* wt_check_finite over n float64 values (the float64 loop's outputs and
 * gradients). */
int wt_check_finite_f64(const double *p, int64_t n, int32_t *flag, void *stream);

/* ------------------------------------------------------------------------ */
/* Material coefficients                                                     */
/* ------------------------------------------------------------------------ */

/* Leapfrog coefficients of cell.py:12-17, with the reference's fp32 rounding:
 *   dtm2 = 1/(dt*dt), dtm1 = 1/dt
 *   A  = 1/(dtm2 + b*dtm1)      ((dt**-2 + b*dt**-1).pow(-1), cell.py:13)
 *   Bm = dtm2 - b*dtm1          (cell.py:14)
 *   C2 = c*c                    (c.pow(2), cell.py:15)
 * over n cells.  Replaces the six material-field ops the reference
 * recomputes on every step (cell.py:13-15). */
int wt_coef_linear(const float *b, const float *c, int64_t n, float dt,
                   float *A, float *Bm, float *C2, void *stream);

/* ------------------------------------------------------------------------ */
/* Laplacian: backs operators._laplacian (operators.py:5-11)                  */
/* ------------------------------------------------------------------------ */

/* out = h^-2 * (5-point stencil of u) with zero padding, u and out
 * (B,Nx,Ny), out not aliasing u; in conv2d's accumulation order, so bitwise
 * equal to the reference's fp32 CPU result for any h.  The operator is its
 * own adjoint (symmetric stencil, zero padding): the backward is the same
 * call on the output gradient. */
int wt_laplacian(const float *u, float *out, int32_t B, int32_t Nx, int32_t Ny, float h,
                 void *stream);
/* wt_laplacian in float64 (set_dtype('float64'), utils.py:14-20). */
int wt_laplacian_f64(const double *u, double *out, int32_t B, int32_t Nx, int32_t Ny,
                     double h, void *stream);

/* ------------------------------------------------------------------------ */
/* Single step: backs cell.TimeStep (cell.py:20-44)                          */
/* ------------------------------------------------------------------------ */

/* y = _time_step(b, c, y1, y2, dt, h)  (cell.py:12-17, operators.py:5-11).
 * b and c are (Nx,Ny) when bc_per_sample == 0, else (B,Nx,Ny).
 * Bitwise equal to the reference's fp32 CPU result for h == 1. */
int wt_step_fwd(const float *b, const float *c, int32_t bc_per_sample,
                const float *y1, const float *y2, int32_t B, int32_t Nx,
                int32_t Ny, float dt, float h, float *y, void *stream);

/* TimeStep.backward (cell.py:26-44) for grad_output g.  Any output pointer
 * may be NULL (the needs_input_grad gating of cell.py:32,35,37,41).
 * grad_b / grad_c are (Nx,Ny) summed over the batch when bc_per_sample == 0
 * (what autograd's reduction to the input shape yields), else (B,Nx,Ny). */
int wt_step_bwd(const float *b, const float *c, int32_t bc_per_sample,
                const float *y1, const float *y2, const float *g, int32_t B,
                int32_t Nx, int32_t Ny, float dt, float h, float *grad_b,
                float *grad_c, float *grad_y1, float *grad_y2, void *stream);

/* wt_step_fwd / wt_step_bwd in float64 (TimeStep under set_dtype('float64'),
 * utils.py:14-20): the same arguments, arrays and scalars double. */
int wt_step_fwd_f64(const double *b, const double *c, int32_t bc_per_sample, const double *y1,
                    const double *y2, int32_t B, int32_t Nx, int32_t Ny, double dt, double h,
                    double *y, void *stream);
int wt_step_bwd_f64(const double *b, const double *c, int32_t bc_per_sample, const double *y1,
                    const double *y2, const double *g, int32_t B, int32_t Nx, int32_t Ny,
                    double dt, double h, double *grad_b, double *grad_c, double *grad_y1,
                    double *grad_y2, void *stream);

/* ------------------------------------------------------------------------ */
/* Whole time loop: backs WaveRNN.forward (rnn.py:21-72) and the autograd     */
/* replay of its tape (cell.py:26-44 per step)                               */
/* ------------------------------------------------------------------------ */

/* Static description of one WaveRNN configuration.  Host memory. */
typedef struct wt_config {
  int32_t Nx, Ny;      /* geom.domain_shape (geom.py:19) */
  float dt, h;         /* cell.dt (cell.py:60), geom.h (geom.py:20) */
  /* sources (source.py:7-22): nsrc_px rows of (x, y, source_index); each
   * source's pixels deduplicated (index_put assigns, source.py:20).  Every
   * source receives the same input sample x[b, n] (rnn.py:56-57), with
   * scale dt**2 = 1.0 (source.py:15). */
  int32_t nsrc_px;
  const int32_t *src_px;
  /* probes (probe.py:6-27): ntap rows of (x, y, kind); kind 0 = WaveProbe
   * (u), kind 1 = WaveIntensityProbe (u*u).  Output column = row index. */
  int32_t ntap;
  const int32_t *taps;
  /* nonlinear cell (cell.py:8-9, 94-102): active when sat_b0 > 0 or
   * c_nl != 0.  c_clamp != 0 clamps c to [c_min, c_max] (an extension). */
  float sat_b0, sat_uth, c_nl;
  int32_t c_clamp;
  float c_min, c_max;
} wt_config;

/* Kernel-variant / tiling choice.  All fields 0 = let the library choose
 * (heuristic); wavetorch_amd.autotune fills these from measurements. */
typedef struct wt_tuning {
  int32_t plain;     /* 1 = one launch per step (the plain GPU form) */
  int32_t strips;    /* row strips per sample (tiled form) */
  int32_t tblock;    /* time steps per launch (temporal-block depth) */
  int32_t nt;        /* threads per workgroup: 256, 512 or 1024 */
  int32_t cpt;       /* cells per thread (compile-time variants) */
  int32_t ck_every;  /* checkpoint interval K in steps (0 = 64) */
  int32_t kind;      /* tiled kernel family: 0 = column segments (cpt = rows
                        per thread), 1 = flat cell list */
  int32_t tcols;     /* column tiles per sample (column-segment family; 0 = 1) */
  /* adjoint kernel choice; 0 = same as the forward fields above */
  int32_t a_strips, a_tcols, a_tblock, a_nt, a_cpt;
  /* batched instances: the batch runs as this many concurrent parts on as
     many streams (0 = library choice, at most 8); single plans ignore it */
  int32_t streams;
} wt_tuning;

/* Creates a plan for `cfg` at batch B and length T: validates the
 * configuration (indices, CFL as cell.py:66-73 with cmax from `cmax`; grids
 * up to Nx * (Ny + 1) < 2^22 cells, about 2047 x 2047, which the kernels'
 * 32-bit in-field cell indices assume -- WT_EUNSUPPORTED beyond),
 * picks the kernel variant and uploads the source/probe/tile tables to the
 * current device (synchronously).  The plan is library-owned; destroy with
 * wt_plan_destroy.  `tuning` may be NULL. */
int wt_plan_create(const wt_config *cfg, int32_t B, int32_t T, float cmax,
                   const wt_tuning *tuning, void **plan_out);

/* Host-only half of wt_plan_create: the same validation (including the CFL
 * check of cell.py:66-73) and tiling/variant choice, with no device work.
 * Writes the resolved tuning to *out and the with-grad workspace sizes of
 * wt_workspace_bytes to *saved / *temp (any may be NULL). */
int wt_plan_resolve(const wt_config *cfg, int32_t B, int32_t T, float cmax,
                    const wt_tuning *tuning, wt_tuning *out, int64_t *saved,
                    int64_t *temp);
int wt_plan_destroy(void *plan);
/* The tuning actually chosen (after heuristics). */
int wt_plan_tuning(const void *plan, wt_tuning *out);

/* Bytes of the two caller-provided workspaces:
 *   saved: lives from wt_loop_fwd to wt_loop_adj (coefficients, checkpoints,
 *          probe amplitudes, and the record of the last checkpoint segment,
 *          which the forward writes so that the reverse sweep need not
 *          recompute that segment); needed only when with_grad.
 *   temp:  scratch of one call (state ping-pong, adjoint state, gradient
 *          partials, and the segment record of the reverse sweep's
 *          recomputes unless the plan keeps the last segment's record in
 *          `saved`, which the sweep then reuses for them).
 * wt_loop_fwd needs only the with_grad = 0 size of `temp`, with or without
 * gradients. */
int wt_workspace_bytes(const void *plan, int32_t with_grad, int64_t *saved,
                       int64_t *temp);

/* Forward time loop (rnn.py:21-72).
 *   b, c        (Nx,Ny) geom.b and c_linear (rnn.py:46)
 *   rho         (Nx,Ny) raw density (rnn.py:47); used by the nonlinear cell
 *   x           (B,T) input signal
 *   y           (B,T,ntap) probe output (rnn.py:59-64,70), may be NULL
 *   fields      (B,T/field_stride,Nx,Ny) fields u_{n+1} for n+1 a multiple of
 *               field_stride (rnn.py:66 with output_fields; stride 1 =
 *               reference behaviour), may be NULL
 *   with_grad   store checkpoints, amplitudes and the last segment's record
 *               into `saved` for wt_loop_adj */
int wt_loop_fwd(void *plan, const float *b, const float *c, const float *rho,
                const float *x, float *y, float *fields, int32_t field_stride,
                int32_t with_grad, void *saved, void *temp, void *stream);

/* Reverse sweep (the autograd replay of cell.py:26-44 over rnn.py:50-70)
 * with checkpoint recompute every K steps (every segment but the last, whose
 * record the forward left in `saved`; with K >= T nothing is recomputed).
 * Consumes `saved`: the recomputes overwrite the last segment's record, so
 * each wt_loop_fwd(with_grad) is followed by at most one wt_loop_adj.
 *   ybar        (B,T,ntap) dJ/dy, or NULL
 *   fbar        (B,T,Nx,Ny) dJ/d(fields) (stride-1 fields only), or NULL
 *   grad_c      (Nx,Ny) dJ/dc_linear, summed over batch, or NULL
 *   grad_b      (Nx,Ny) dJ/db (nonlinear cell only), or NULL
 *   grad_rho    (Nx,Ny) dJ/drho through the nonlinear terms, or NULL
 *   grad_x      (B,T)   dJ/dx, or NULL
 * Output buffers are overwritten (not accumulated). */
int wt_loop_adj(void *plan, const float *b, const float *c, const float *rho,
                const float *x, const float *ybar, const float *fbar,
                void *saved, void *temp, float *grad_c, float *grad_b,
                float *grad_rho, float *grad_x, void *stream);

/* Measured choice of the tuning for one problem shape (no counterpart in the
 * reference, whose time loop is a Python loop, rnn.py:50-67): times
 * wt_loop_fwd (and with with_grad the checkpointed wt_loop_adj) on the
 * caller's b, c, rho, x (device buffers as for wt_loop_fwd) for candidate
 * tilings, starting from the library's heuristic plan and moving one field
 * at a time -- tiles per sample (strips x tcols) near the heuristic's,
 * time-block depth, rows per thread, forward and adjoint separately -- then
 * the coupled (depth, rows) pairs; keeps a change when it is > 0.5% faster;
 * two passes or until budget_ms of wall time.  Each candidate's time is the
 * best of `reps` timed runs (HIP events on `stream`) after one warm-up.
 * `start` (may be NULL) supplies the fixed checkpoint interval (ck_every;
 * 0 = 64).  Writes the winner (as wt_plan_resolve resolves it) to *out and
 * its time to *best_ms (may be NULL).  The search allocates and frees its
 * own device buffers (not for use under stream capture) and synchronises
 * `stream`; results are cached per shape for the life of the process
 * (thread-safe).  Python callers: wavetorch_amd.autotune.tune_native. */
int wt_autotune(const wt_config *cfg, int32_t B, int32_t T, float cmax,
                const wt_tuning *start, int32_t with_grad, const float *b,
                const float *c, const float *rho, const float *x, int32_t reps,
                double budget_ms, wt_tuning *out, double *best_ms, void *stream);

/* ------------------------------------------------------------------------ */
/* The time loop in double precision (set_dtype('float64'), utils.py:14-20)  */
/* ------------------------------------------------------------------------ */

/* Scalars of the float64 loop, which wt_config holds only as float: dt
 * (cell.py:60), h (geom.py:20) and the nonlinear coefficients (cell.py:60-64;
 * c_min / c_max of the clamp extension). */
typedef struct wt_scalars_f64 {
  double dt, h, sat_b0, sat_uth, c_nl, c_min, c_max;
} wt_scalars_f64;

/* Bytes of the workspaces of wt_loop_fwd_f64 / wt_loop_adj_f64 for a plan
 * made by wt_plan_create for the same configuration, batch and length:
 *   saved (with_grad): every field u_n, n = -1..T, ((T+2)*B*Nx*Ny doubles,
 *          what the reference's autograd tape keeps; no checkpoint recompute)
 *   temp:  three fields (forward without gradients), or the adjoint's
 *          three lambdas, gradient accumulators and lambda at the sources. */
int wt_workspace_bytes_f64(const void *plan, int32_t with_grad, int64_t *saved,
                           int64_t *temp);

/* wt_loop_fwd in float64 (rnn.py:21-72 under set_dtype('float64')): one
 * launch per step, the reference's operation order in double precision.
 * Arguments as wt_loop_fwd, all arrays double; `sc` gives dt, h and the
 * nonlinear coefficients in double.  With with_grad the fields go to `saved`
 * and `temp` may be NULL. */
int wt_loop_fwd_f64(void *plan, const wt_scalars_f64 *sc, const double *b, const double *c,
                    const double *rho, const double *x, double *y, double *fields,
                    int32_t field_stride, int32_t with_grad, void *saved, void *temp,
                    void *stream);

/* wt_loop_adj in float64: the reverse sweep over the stored fields (the
 * autograd replay of cell.py:26-44 over rnn.py:50-70 in double precision).
 * Arguments as wt_loop_adj (no x: the fields are stored), all arrays double;
 * output buffers are overwritten; grad_b is dJ/db for the linear cell as well;
 * fbar, when given, holds every step (field_stride 1). */
int wt_loop_adj_f64(void *plan, const wt_scalars_f64 *sc, const double *b, const double *c,
                    const double *rho, const double *ybar, const double *fbar, void *saved,
                    void *temp, double *grad_c, double *grad_b, double *grad_rho,
                    double *grad_x, void *stream);

/* ------------------------------------------------------------------------ */
/* Batched instances                                                         */
/* ------------------------------------------------------------------------ */
/* Many independent linear WaveRNN problems of uneven grid size and length
 * (e.g. a set of inverse-design problems, each a WaveRNN of rnn.py:4-72 with
 * its own c and input length) run as one sequence of launches in which every
 * launch covers every instance that still has steps to run.  Instances are
 * passed in groups that share a configuration (grid, PML field b, sources,
 * probes; at most 8 groups); within a group every instance has its own c and
 * its own length.  Each instance's result is what wt_loop_fwd/wt_loop_adj
 * give for it alone (replaces looping rnn.py:21-72 + its autograd backward
 * over the instances).  The reference runs one configuration at a time. */
typedef struct wt_batch_group_io {
  const float *b;     /* (Nx,Ny) PML damping shared by the group */
  const float *c;     /* (B,Nx,Ny) wave speed of each instance */
  const float *x;     /* (B,T) input, T = the longest length of the batch */
  float *y;           /* (B,T,ntap) probe series (forward); entries at or past
                         an instance's length are unspecified */
  const float *ybar;  /* (B,T,ntap) dJ/dy (adjoint); entries at or past an
                         instance's length are ignored */
  float *grad_c;      /* (B,Nx,Ny) dJ/dc of each instance, or NULL */
  float *grad_x;      /* (B,T) dJ/dx, or NULL */
} wt_batch_group_io;

/* ngroups configurations; B[g] instances in group g; lengths = the steps of
 * every instance, group after group (sum of B entries); cmax[g] for the CFL
 * check of group g.  tunings: NULL (library choice) or ngroups entries whose
 * tblock/nt/cpt/a_tblock/a_nt/a_cpt/ck_every/streams are taken from entry 0
 * (common to every launch) and strips/tcols (and a_strips/a_tcols) per
 * group.  Only the linear cell with h = 1 is batched.
 * A group whose lengths are in non-increasing order (longest first) has its
 * launches cover only the instances that still have steps to run; in any
 * other order they cover every instance while the group's longest one runs
 * (same results, more idle workgroups).
 * The forward records lap(u) of the last checkpoint segment's steps into the
 * saved workspace, where the reverse sweep starts without a recompute; with
 * ck_every >= the longest length (one segment) that is every step of every
 * instance up to its own length, packed by length (about
 * 4 * sum(L_i) * Nx * (Ny + 1) bytes per group), and nothing is recomputed.
 * The plan may split the batch into `streams` parts (contiguous instance
 * ranges of every group, balanced by steps; by default one part per ~512
 * work items of a launch, at most 4): each part runs on a stream of the
 * plan's own, forked from the caller's stream and joined back to it with
 * events, so a captured graph holds all of them and results are those of
 * the single sequence (the instances are independent).  Launches of the
 * parts overlap, filling the CUs that one part's launches leave idle; with
 * groups longest first the parts hold instances of different lengths, so
 * they reach their short-lived tails at different times.
 * Calls with one plan must not overlap (the part streams are the plan's). */
int wt_batch_plan_create(int32_t ngroups, const wt_config *cfgs, const int32_t *B,
                         const int32_t *lengths, const float *cmax,
                         const wt_tuning *tunings, void **plan_out);
int wt_batch_plan_tuning(const void *plan, int32_t group, wt_tuning *out);
/* One saved and one temporary buffer for the whole batch. */
int wt_batch_workspace_bytes(const void *plan, int64_t *saved, int64_t *temp);
/* Forward of every instance (probes into io[g].y; checkpoints into saved
 * when with_grad). */
int wt_batch_fwd(void *plan, const wt_batch_group_io *io, int32_t with_grad,
                 void *saved, void *temp, void *stream);
/* Reverse sweep after wt_batch_fwd(with_grad = 1): io[g].ybar in,
 * io[g].grad_c / grad_x out. */
int wt_batch_adj(void *plan, const wt_batch_group_io *io, void *saved,
                 void *temp, void *stream);
/* wt_batch_fwd + wt_batch_adj for a loss whose dJ/dy is known before the
 * forward (a fixed weighting of the probe series), as one call that a graph
 * can capture whole. */
int wt_batch_fwd_adj(void *plan, const wt_batch_group_io *io, void *saved,
                     void *temp, void *stream);
int wt_batch_plan_destroy(void *plan);

/* ------------------------------------------------------------------------ */
/* Geometry parameterisation (geom.py:207-233)                               */
/* ------------------------------------------------------------------------ */

/* c = c0 + (c1-c0) * P(Blur^N(rho))   (WaveGeometryFreeForm.c, geom.py:231-233)
 *   kernel  (2r+1)^2 normalised blur kernel (geom.py:149-154)
 *   P(x)  = (tanh(beta*eta) + tanh(beta*(x-eta))) / (tanh(beta*eta)+tanh(beta*(1-eta)))
 *   work    2*Nx*Ny floats of device scratch (blur ping-pong)
 *   rho_proj (Nx,Ny) receives P(Blur^N(rho)) (may be NULL)
 *   blur_out (Nx,Ny) receives Blur^N(rho) (needed by the backward; may be NULL) */
int wt_geom_freeform_c(const float *rho, const float *kernel, int32_t radius,
                       int32_t blur_N, int32_t Nx, int32_t Ny, float eta,
                       float beta, float c0, float c1, float *c,
                       float *rho_proj, float *blur_out, float *work,
                       void *stream);

/* Backward of wt_geom_freeform_c: grad_rho = (Blur^T)^N((c1-c0) P'(blur) grad_c).
 *   blur_in (Nx,Ny) Blur^N(rho) from the forward; work 2*Nx*Ny floats. */
int wt_geom_freeform_c_bwd(const float *grad_c, const float *blur_in,
                           const float *kernel, int32_t radius, int32_t blur_N,
                           int32_t Nx, int32_t Ny, float eta, float beta,
                           float c0, float c1, float *grad_rho, float *work,
                           void *stream);

/* c = c0 + (c1-c0) * P(raw) for WaveGeometryHoley (geom.py:109-135):
 *   raw[i,j] = sum_h exp(-sqrt((i-x_h)^2 + (j-y_h)^2) / r_h), in hole order
 *   x, y, r  (nholes) hole centres and radii (device; may be NULL if nholes = 0)
 *   c        (Nx,Ny) output; raw_out (Nx,Ny) receives raw (for the backward)
 * With c0 = 0, c1 = 1 the output is the projected density (the `rho`
 * property, geom.py:130-131). */
int wt_geom_holey_c(const float *x, const float *y, const float *r, int32_t nholes,
                    int32_t Nx, int32_t Ny, float eta, float beta, float c0,
                    float c1, float *c, float *raw_out, void *stream);

/* Backward of wt_geom_holey_c: dJ/dx, dJ/dy, dJ/dr (nholes each; any may be
 * NULL) from grad_c (Nx,Ny) and raw from the forward.  As autograd through
 * the reference's sqrt, a centre exactly on a grid point gives NaN for its
 * x and y gradients.  work: wt_geom_holey_workspace_bytes bytes, 8-aligned. */
int wt_geom_holey_c_bwd(const float *grad_c, const float *raw, const float *x,
                        const float *y, const float *r, int32_t nholes, int32_t Nx,
                        int32_t Ny, float eta, float beta, float c0, float c1,
                        float *grad_x, float *grad_y, float *grad_r, void *work,
                        void *stream);

/* Bytes of `work` for wt_geom_holey_c_bwd (no counterpart in the reference). */
int wt_geom_holey_workspace_bytes(int32_t nholes, int32_t Nx, int32_t Ny, int64_t *bytes);

/* The four geometry entry points above in float64 (the geometry of a model
 * built after set_dtype('float64'), utils.py:14-20): the same arguments,
 * every array and scalar double; the FreeForm work arrays hold 2*Nx*Ny
 * doubles, the Holey backward's work wt_geom_holey_workspace_bytes_f64
 * bytes. */
int wt_geom_freeform_c_f64(const double *rho, const double *kernel, int32_t radius,
                           int32_t blur_N, int32_t Nx, int32_t Ny, double eta, double beta,
                           double c0, double c1, double *c, double *rho_proj, double *blur_out,
                           double *work, void *stream);
int wt_geom_freeform_c_bwd_f64(const double *grad_c, const double *blur_in,
                               const double *kernel, int32_t radius, int32_t blur_N, int32_t Nx,
                               int32_t Ny, double eta, double beta, double c0, double c1,
                               double *grad_rho, double *work, void *stream);
int wt_geom_holey_c_f64(const double *x, const double *y, const double *r, int32_t nholes,
                        int32_t Nx, int32_t Ny, double eta, double beta, double c0, double c1,
                        double *c, double *raw_out, void *stream);
int wt_geom_holey_c_bwd_f64(const double *grad_c, const double *raw, const double *x,
                            const double *y, const double *r, int32_t nholes, int32_t Nx,
                            int32_t Ny, double eta, double beta, double c0, double c1,
                            double *grad_x, double *grad_y, double *grad_r, void *work,
                            void *stream);
int wt_geom_holey_workspace_bytes_f64(int32_t nholes, int32_t Nx, int32_t Ny, int64_t *bytes);

/* ------------------------------------------------------------------------ */
/* Loss head (train.py:61-62 with utils.py:35-36)                            */
/* ------------------------------------------------------------------------ */

/* J = CrossEntropyLoss()(normalize_power(y.sum(dim=1)), labels), the
 * reference's training criterion on its probe series (train.py:61-62; the
 * default mean reduction), and dJ/dI (B,P) with I = y.sum(dim=1): the
 * gradient of J with respect to y[b,t,k] is grad_I[b,k] for every t.
 *   y       (B,T,P) fp32 probe series, contiguous; P <= 256
 *   labels  (B) int64 class indices in [0, P); a sample with any other label
 *           gets a NaN loss term and NaN gradients (the reference raises)
 *   loss    one fp32 (device); grad_I (B,P) fp32 (device)
 *   work    device scratch of wt_loss_power_ce_workspace_bytes(B, T, P) bytes,
 *           8-byte aligned (partial sums, per-sample losses)
 * Three kernel launches on `stream`, no allocation, no memset; 1 <= B <= 65535. */
int wt_loss_power_ce(const float *y, const int64_t *labels, int32_t B, int32_t T,
                     int32_t P, float *loss, float *grad_I, void *work,
                     void *stream);

/* Bytes of `work` that wt_loss_power_ce needs for (B, T, P) (no counterpart
 * in the reference). */
int wt_loss_power_ce_workspace_bytes(int32_t B, int32_t T, int32_t P, int64_t *bytes);

/* wt_loss_power_ce without a workspace argument: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * stream capture, fixed-shape serving loops, harnesses that forbid
 * hipMalloc).  It runs the same three kernels, with the partial sums in a
 * static device array of the library (1 MiB = 131072 doubles; each device
 * the library runs on holds its own copy until the process ends), and
 * allocates nothing.  Shapes needing more than B*C*P + B <= 131072 doubles
 * (C = the chunk count, at most 16; e.g. B <= 2048 at P = 3) give
 * WT_ENOMEM.  Calls that overlap on one device share the array and race;
 * calls on one stream are ordered. */
int wt_loss_power_ce_static(const float *y, const int64_t *labels, int32_t B, int32_t T,
                            int32_t P, float *loss, float *grad_I, void *stream);

/* ------------------------------------------------------------------------ */
/* Kernel timing (no counterpart in the reference, which has no profiler)    */
/* ------------------------------------------------------------------------ */

/* Enables/disables HIP-event timing of every time-loop kernel launch (events
 * recorded on the launch stream).  Not for use under graph capture. */
int wt_profile(int32_t enable);
/* Synchronises the recorded events and returns, per kernel class
 * (0 forward, 1 forward recompute with record, 2 adjoint), the summed
 * kernel milliseconds, the number of launches and the cell-steps they
 * processed; then clears the record.  Each array has 3 entries. */
int wt_profile_read(double *ms, int64_t *launches, int64_t *cell_steps);

/* ------------------------------------------------------------------------ */
/* HIP graph capture of any sequence of the calls above                      */
/* ------------------------------------------------------------------------ */

/* Begin stream capture on `stream` (hipStreamCaptureModeThreadLocal).  Every
 * wt_* call made on that stream until wt_graph_end is recorded (the library
 * records no memset nodes). */
int wt_graph_begin(void *stream);
/* End capture and instantiate; *graph_out is a library-owned handle. */
int wt_graph_end(void *stream, void **graph_out);
int wt_graph_launch(void *graph, void *stream);
int wt_graph_destroy(void *graph);

#ifdef __cplusplus
}
#endif
#endif /* WAVETORCH_HIP_H */
