// wt_runtime.hip -- host runtime and C ABI of libwavetorch_hip.so.
//
// Plans (validated configuration + device tables + tiling), workspace
// layout, the forward time loop with checkpoints every K steps, the reverse
// sweep that recomputes each K-step segment from its checkpoint and runs the
// adjoint through it, the single-step TimeStep entry points and HIP graph
// capture.  See include/wavetorch_hip.h for the contract of every export.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "wt_internal.h"

namespace wt {

// ---------------------------------------------------------------- errors

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int hip_fail(hipError_t e, const char *what) {
  set_error("HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  return WT_EHIP;
}

// launchers implemented in wt_plain.hip / wt_tiled.hip / wt_geom.hip
hipError_t launch_coef_linear(const float *b, const float *c, int64_t n, float dtm1, float dtm2,
                              float *A, float *Bm, float *C2, hipStream_t s);
hipError_t launch_laplacian(const float *u, int B, int Nx, int Ny, float t1, float t4,
                            float *out, hipStream_t s);
hipError_t launch_laplacian(const double *u, int B, int Nx, int Ny, double t1, double t4,
                            double *out, hipStream_t s);
hipError_t launch_step_fwd(bool unit_h, const double *b, const double *c, int bc_ps,
                           const double *y1, const double *y2, int B, int Nx, int Ny,
                           double dtm1, double dtm2, double s2, double t1, double t4, double *y,
                           hipStream_t s);
hipError_t launch_step_bwd(bool unit_h, const double *b, const double *c, int bc_ps,
                           const double *y1, const double *y2, const double *g, int B, int Nx,
                           int Ny, double dt, double t1, double t4, double *gb, double *gc,
                           double *gy1, double *gy2, hipStream_t s);
hipError_t launch_step_fwd(bool unit_h, const float *b, const float *c, int bc_ps,
                           const float *y1, const float *y2, int B, int Nx, int Ny, float dtm1,
                           float dtm2, float s2, float t1, float t4, float *y, hipStream_t s);
hipError_t launch_step_bwd(bool unit_h, const float *b, const float *c, int bc_ps,
                           const float *y1, const float *y2, const float *g, int B, int Nx,
                           int Ny, float dt, float t1, float t4, float *gb, float *gc,
                           float *gy1, float *gy2, hipStream_t s);
hipError_t launch_fwd_plain(bool unit_h, int nl, bool rec, const FwdArgs &a,
                            const int32_t *spmap, int n, hipStream_t s);
hipError_t launch_adj_plain(bool unit_h, const AdjArgs &a, const int32_t *spmap,
                            const float *wl1, float *wl_out, const float *fbar, int m,
                            hipStream_t s);
hipError_t launch_adj_nl_plain(bool unit_h, int nl, const AdjArgs &a, const int32_t *spmap,
                               const NlAdjBufs &nb, const float *fbar, int m, hipStream_t s);
hipError_t launch_batch_sum(const float *acc, int B, int64_t F, float *out, hipStream_t s);
hipError_t launch_square(const float *in, int64_t n, float *out, hipStream_t s);
hipError_t launch_finalize_gc_ps(const float *gacc, int B, int64_t F, const float *c,
                                 const float *A, float *gc, hipStream_t s);
hipError_t launch_fwd_cs_batch(int nt, int R, bool rec, const BatchFwd &bl, int RHmax,
                               int RWmax, hipStream_t s);
hipError_t launch_adj_cs_batch(int nt, int R, const BatchAdj &bl, int RHmax, int RWmax,
                               hipStream_t s);
hipError_t launch_finalize_gc(const float *gacc, int B, int64_t F, const float *c,
                              const float *A, float *gc, hipStream_t s);
hipError_t launch_finalize_gx(const float *gsrc, int B, int T, int nslot, const int32_t *mult,
                              float *gx, hipStream_t s);
hipError_t launch_fill(float *p, int64_t n, float v, hipStream_t s);
hipError_t launch_nonfinite(const float *p, int64_t n, int32_t *flag, hipStream_t s);
hipError_t launch_nonfinite(const double *p, int64_t n, int32_t *flag, hipStream_t s);
hipError_t launch_fwd_tiled(int nt, int cpt, bool unit_h, int nl, bool rec, const FwdArgs &a,
                            int LWmax, hipStream_t s);
hipError_t launch_adj_tiled(int nt, int cpt, bool unit_h, const AdjArgs &a, int LWmax,
                            hipStream_t s);
bool tiled_variant_exists(int nt, int cpt);
size_t tiled_shmem_bytes(int nt, int cpt, int LW);
bool cs_variant_exists(int nt, int r);
size_t cs_shmem_bytes(int R, int RH, int RW, bool padded);
int cs_threads(int R, int RH, int RW);
hipError_t launch_fwd_cs(int nt, int R, int nl, bool rec, const FwdArgs &a, int RHmax,
                         int RWmax, hipStream_t s);
hipError_t launch_adj_cs(int nt, int R, const AdjArgs &a, int RHmax, int RWmax, hipStream_t s);
bool cs_nl_variant_exists(int nt, int r);
size_t cs_nl_shmem_bytes(int R, int RH, int RW);
hipError_t launch_adj_cs_nl(int nt, int R, bool gb, const AdjArgs &a, const NlCarry &cy,
                            int RHmax, int RWmax, hipStream_t s);
hipError_t launch_mul_scalar(const float *in, float v, int64_t n, float *out, hipStream_t s);
hipError_t launch_clamp(const float *in, float lo, float hi, int64_t n, float *out,
                        hipStream_t s);

// ---------------------------------------------------------------- profiling
//
// Optional per-kernel-class timing with HIP events recorded on the launch
// stream around every launch of the loop kernels (enabled by wt_profile).
// Classes: 0 forward, 1 forward-with-record (segment recompute), 2 adjoint.
struct Prof {
  std::mutex mu;
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[3];
  std::vector<int64_t> work[3];  // cell-steps of each launch
  std::vector<hipEvent_t> pool;
};
static Prof g_prof;

static hipEvent_t prof_event() {
  hipEvent_t e = nullptr;
  if (!g_prof.pool.empty()) {
    e = g_prof.pool.back();
    g_prof.pool.pop_back();
  } else {
    (void)hipEventCreate(&e);
  }
  return e;
}

struct ProfScope {
  int cls;
  hipStream_t s;
  hipEvent_t a = nullptr;
  int64_t work;
  ProfScope(int c, hipStream_t st, int64_t w) : cls(c), s(st), work(w) {
    if (!g_prof.on) return;
    std::lock_guard<std::mutex> g(g_prof.mu);
    a = prof_event();
    (void)hipEventRecord(a, s);
  }
  ~ProfScope() {
    if (!a) return;
    std::lock_guard<std::mutex> g(g_prof.mu);
    hipEvent_t b = prof_event();
    (void)hipEventRecord(b, s);
    g_prof.ev[cls].push_back({a, b});
    g_prof.work[cls].push_back(work);
  }
};

constexpr int SRC_SLOTS_HOST = 4;  // must match wt_tiled.hip
constexpr int SPC_SLOTS_HOST = 2;
constexpr size_t LDS_LIMIT = 160 * 1024;

// ---------------------------------------------------------------- plan

struct Plan {
  int dev = 0;
  int B = 0, T = 0, Nx = 0, Ny = 0;
  float dt = 0, h = 1;
  bool unit_h = true;
  int nl = 0;  // 0 linear; bit0 saturable damping, bit1 Kerr speed, bit2 clamp
  bool clamp_only = false;
  float sat_b0 = 0, sat_uth = 1, c_nl = 0, cmin = 0, cmax = 0;
  float dtm1 = 0, dtm2 = 0, s2 = 0, t1 = 1, t4 = -4;
  wt_tuning tun{};
  int K = 64, nseg = 1;
  int nspx = 0, ntap = 0, nslot = 0;
  std::vector<int32_t> h_spx, h_taps, h_tiles, h_mult;
  int32_t *d_spx = nullptr, *d_taps = nullptr, *d_spmap = nullptr, *d_mult = nullptr;
  int32_t *d_tiles = nullptr;
  int ntiles = 1, LWmax = 0, RHmax = 0, RWmax = 0;
  // adjoint tiling (column-segment family), may differ from the forward's
  std::vector<int32_t> h_tiles_a;
  int32_t *d_tiles_a = nullptr;
  int ntiles_a = 1, RHmax_a = 0, RWmax_a = 0, tb_a = 1, nt_a = 0, cpt_a = 0;
  bool tiled = false;
  // batched instances: one c (hence C2) per sample instead of one per plan
  bool c2ps = false;
  // step slots of the saved record (0 = B * K; batched instances in one
  // segment: the sum of their own lengths)
  int64_t rec_steps = 0;
  // workspace layout (bytes)
  size_t F = 0;
  size_t saved_coef = 0, saved_ck = 0, saved_yraw = 0, saved_rec = 0, saved_bytes = 0;
  size_t t_state = 0, t_lam = 0, t_wl = 0, t_gacc = 0, t_seg = 0, t_gsrc = 0, t_extra = 0,
         temp_bytes = 0;
  size_t temp_bytes_nograd = 0;
};

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

static void free_plan(Plan *p) {
  if (!p) return;
  if (p->d_spx) (void)hipFree(p->d_spx);
  if (p->d_taps) (void)hipFree(p->d_taps);
  if (p->d_spmap) (void)hipFree(p->d_spmap);
  if (p->d_mult) (void)hipFree(p->d_mult);
  if (p->d_tiles) (void)hipFree(p->d_tiles);
  if (p->d_tiles_a) (void)hipFree(p->d_tiles_a);
  delete p;
}

// Row strips (optionally 2-D tiles when `cols` > 1) with halo `halo`.
static void build_tiles(Plan &p, int strips, int cols, int halo) {
  p.h_tiles.clear();
  p.LWmax = 0;
  p.RHmax = 0;
  p.RWmax = 0;
  for (int si = 0; si < strips; ++si) {
    int o_r0 = (int)((int64_t)p.Nx * si / strips), o_r1 = (int)((int64_t)p.Nx * (si + 1) / strips);
    for (int ci = 0; ci < cols; ++ci) {
      int o_c0 = (int)((int64_t)p.Ny * ci / cols), o_c1 = (int)((int64_t)p.Ny * (ci + 1) / cols);
      int R0 = std::max(0, o_r0 - (strips > 1 ? halo : 0));
      int R1 = std::min(p.Nx, o_r1 + (strips > 1 ? halo : 0));
      int C0 = std::max(0, o_c0 - (cols > 1 ? halo : 0));
      int C1 = std::min(p.Ny, o_c1 + (cols > 1 ? halo : 0));
      int32_t t[TILE_W] = {0};
      t[0] = R0; t[1] = R1 - R0; t[2] = C0; t[3] = C1 - C0;
      t[4] = o_r0; t[5] = o_r1; t[6] = o_c0; t[7] = o_c1;
      // special pixels sorted by x: range with x in [R0, R1)
      int lo = 0, hi = 0;
      while (lo < p.nspx && p.h_spx[lo * SPX_W] < R0) ++lo;
      hi = lo;
      while (hi < p.nspx && p.h_spx[hi * SPX_W] < R1) ++hi;
      t[8] = lo; t[9] = hi;
      int tl = 0, th = 0;
      while (tl < p.ntap && p.h_taps[tl * TAP_W] < o_r0) ++tl;
      th = tl;
      while (th < p.ntap && p.h_taps[th * TAP_W] < o_r1) ++th;
      t[10] = tl; t[11] = th;
      p.h_tiles.insert(p.h_tiles.end(), t, t + TILE_W);
      p.LWmax = std::max(p.LWmax, (C1 - C0) + 1);
      p.RHmax = std::max(p.RHmax, R1 - R0);
      p.RWmax = std::max(p.RWmax, C1 - C0);
    }
  }
  p.ntiles = strips * cols;
}

// Heuristic tilings of an Nx x Ny grid (s strips x c columns of tiles per
// sample): by tile count, then by how far the tiles are from square
// (|log2 of (Nx/s)/(Ny/c)|), then more strips first.  Tilings whose tiles
// are more than 2:1 are left out, except the single tile.
static std::vector<std::pair<int, int>> tile_order(int Nx, int Ny) {
  struct Cand { int s, c; double bad; };
  std::vector<Cand> v;
  for (int s = 1; s <= std::min(Nx, 128); ++s)
    for (int c = 1; c <= std::min(Ny, 128); ++c) {
      const double bad = std::fabs(std::log2(((double)Nx / s) / ((double)Ny / c)));
      if (s * c == 1 || bad <= 1.0 + 1e-9) v.push_back({s, c, bad});
    }
  std::stable_sort(v.begin(), v.end(), [](const Cand &a, const Cand &b) {
    if (a.s * a.c != b.s * b.c) return a.s * a.c < b.s * b.c;
    if (std::fabs(a.bad - b.bad) > 1e-9) return a.bad < b.bad;
    return a.s > b.s;
  });
  std::vector<std::pair<int, int>> out;
  for (auto &e : v) out.emplace_back(e.s, e.c);
  return out;
}

// Checks that the tiled kernels can take this tiling: region fits the
// (nt, cpt) variant and LDS, per-thread source/special/tap slots suffice.
static bool tiles_ok(const Plan &p, int nt, int cpt) {
  if (!tiled_variant_exists(nt, cpt)) return false;
  if (tiled_shmem_bytes(nt, cpt, p.LWmax) > LDS_LIMIT) return false;
  for (int ti = 0; ti < p.ntiles; ++ti) {
    const int32_t *t = &p.h_tiles[ti * TILE_W];
    int RH = t[1], RW = t[3], LW = RW + 1;
    if ((int64_t)RH * LW > (int64_t)nt * cpt) return false;
    std::vector<int> srcn(nt, 0), spcn(nt, 0);
    for (int k = t[8]; k < t[9]; ++k) {
      const int32_t *e = &p.h_spx[k * SPX_W];
      int rr = e[0] - t[0], yy = e[1] - t[2];
      if (rr < 0 || rr >= RH || yy < 0 || yy >= RW) continue;
      int f = rr * LW + yy;
      srcn[f % nt] += e[2];
      spcn[f % nt] += 1;
    }
    for (int i = 0; i < nt; ++i)
      if (srcn[i] > SRC_SLOTS_HOST || spcn[i] > SPC_SLOTS_HOST) return false;
    int ntp = 0;
    for (int k = t[10]; k < t[11]; ++k) {
      const int32_t *tp = &p.h_taps[k * TAP_W];
      if (tp[1] >= t[6] && tp[1] < t[7]) ++ntp;
    }
    if (ntp > nt) return false;
  }
  return true;
}

// Column-segment family: block of cs_threads() threads, R rows per thread;
// per thread at most 2 special rows (adjoint) and source multiplicity <= 2.
// `fwd`: forward tiles (their LDS allocation holds the padded layout).
static bool tiles_ok_cs(const Plan &p, int nt, int R, bool fwd = true) {
  if (!cs_variant_exists(nt, R)) return false;
  for (int ti = 0; ti < p.ntiles; ++ti) {  // tables staged in LDS per tile
    const int32_t *t = &p.h_tiles[ti * TILE_W];
    if (t[9] - t[8] > 256 || t[11] - t[10] > 256) return false;
  }
  if (cs_shmem_bytes(R, p.RHmax, p.RWmax, fwd) + 1024 > LDS_LIMIT) return false;
  int nthr = cs_threads(R, p.RHmax, p.RWmax);
  if (nthr > nt) return false;
  for (int ti = 0; ti < p.ntiles; ++ti) {
    const int32_t *t = &p.h_tiles[ti * TILE_W];
    int RH = t[1], RW = t[3];
    std::vector<int> spc((size_t)RW * ((RH + R - 1) / R), 0);
    int nseed = 0;
    for (int k = t[8]; k < t[9]; ++k) {
      const int32_t *e = &p.h_spx[k * SPX_W];
      if (e[5] > e[4] && ++nseed > 16) return false;
      int rr = e[0] - t[0], cc = e[1] - t[2];
      if (rr < 0 || rr >= RH || cc < 0 || cc >= RW) continue;
      if (e[2] > 1) return false;
      if (++spc[(size_t)(rr / R) * RW + cc] > 2) return false;
    }
    int ntp = 0;
    for (int k = t[10]; k < t[11]; ++k) {
      const int32_t *tp = &p.h_taps[k * TAP_W];
      if (tp[1] >= t[6] && tp[1] < t[7]) ++ntp;
    }
    if (ntp > nthr) return false;
  }
  return true;
}

static const int kCsVariants[][2] = {{1024, 4},  {1024, 6},  {1024, 8},  {1024, 10},
                                     {1024, 12}, {512, 16},  {512, 20},  {512, 24}};

// A 1024-thread variant with at most maxR rows per thread that holds the
// current tiling (the smallest such R).
static bool fit_cs(const Plan &p, int maxR, int &R, bool fwd = true) {
  for (auto &v : kCsVariants) {
    if (v[0] != 1024 || v[1] > maxR) continue;
    if (tiles_ok_cs(p, v[0], v[1], fwd)) {
      R = v[1];
      return true;
    }
  }
  return false;
}

// Smallest R whose block fits (fewest rows per thread -> most threads).
static bool pick_variant_cs(Plan &p, int &nt, int &R, bool fwd = true) {
  for (auto &v : kCsVariants) {
    if (tiles_ok_cs(p, v[0], v[1], fwd)) {
      nt = v[0];
      R = v[1];
      return true;
    }
  }
  return false;
}

static const int kVariants[][2] = {{1024, 8}, {1024, 12}, {1024, 16}, {512, 16},
                                   {512, 24}, {512, 32},  {256, 32},  {256, 48}};

// Choose (nt, cpt) for the current tiling: the smallest region capacity
// that fits, preferring more threads.
static bool pick_variant(Plan &p, int &nt, int &cpt) {
  int best = -1;
  int64_t best_cap = 0;
  for (int v = 0; v < (int)(sizeof(kVariants) / sizeof(kVariants[0])); ++v) {
    int a = kVariants[v][0], c = kVariants[v][1];
    if (!tiles_ok(p, a, c)) continue;
    int64_t cap = (int64_t)a * c;
    if (best < 0 || cap < best_cap || (cap == best_cap && a > kVariants[best][0])) {
      best = v;
      best_cap = cap;
    }
  }
  if (best < 0) return false;
  nt = kVariants[best][0];
  cpt = kVariants[best][1];
  return true;
}

static int plan_layout(Plan &p) {
  size_t F = (size_t)p.Nx * p.Ny;
  p.F = F;
  size_t BF = (size_t)p.B * F * sizeof(float);
  bool nlc = p.nl != 0;
  // saved: coefficients (3F) or rho*c_nl (F), checkpoints, probe amplitudes
  const size_t ncoef = nlc ? 1 : (p.c2ps ? 2 + (size_t)p.B : 3);
  p.saved_coef = align_up(ncoef * F * sizeof(float));
  p.saved_ck = align_up((size_t)p.nseg * 2 * BF);
  p.saved_yraw = align_up((size_t)p.B * p.T * std::max(p.ntap, 1) * sizeof(float));
  // the record of the last segment, written by the forward itself: the
  // reverse sweep starts there and needs no recompute for it (one segment in
  // all when K >= T)
  const size_t segcells =
      (size_t)(p.rec_steps > 0 ? p.rec_steps : (int64_t)p.B * p.K) * p.Nx * (p.Ny + 1);
  p.saved_rec = align_up((nlc ? 2 : 1) * segcells * sizeof(float));
  p.saved_bytes = p.saved_coef + p.saved_ck + p.saved_yraw + p.saved_rec;
  // temp
  p.t_state = align_up(8 * BF);  // 4 state pairs (ping-pong + segment in/out)
  // adjoint state ping-pong (plain uses 3); the nonlinear blocked adjoint
  // adds two sets of four carried coefficient fields
  p.t_lam = align_up((nlc ? 12 : 4) * BF);
  p.t_wl = align_up(2 * BF);     // plain adjoint w*lambda
  p.t_gacc = align_up((nlc ? 3 : 1) * BF);
  // segment record of the reverse sweep's recomputes: when the forward keeps
  // the last segment's record in `saved`, the sweep consumes that record
  // first and then recomputes every earlier segment into the same buffer
  p.t_seg = p.saved_rec ? 0 : align_up((nlc ? 2 : 1) * segcells * sizeof(float));
  p.t_gsrc = align_up((size_t)p.B * p.T * std::max(p.nslot, 1) * sizeof(float));
  p.t_extra = align_up(ncoef * F * sizeof(float));  // coefficients when not saving
  p.temp_bytes = p.t_state + p.t_lam + p.t_wl + p.t_gacc + p.t_seg + p.t_gsrc + p.t_extra;
  p.temp_bytes_nograd = p.t_state + p.t_extra;
  return WT_OK;
}

}  // namespace wt

using namespace wt;

// ====================================================================== ABI

extern "C" {

int wt_abi_version(void) { return WT_ABI_VERSION; }

int wt_last_error(char *buf, size_t len) {
  const std::string &s = g_last_error;
  if (buf && len > 0) {
    size_t n = std::min(len - 1, s.size());
    memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return (int)s.size();
}

extern "C++" {
template <typename T>
static void scalars(T dt, T h, T &dtm1, T &dtm2, T &s2, T &t1, T &t4) {
  // dt**-2 == 1/(dt*dt), dt**-1 == 1/dt, 2/dt**2 == 2*dt**-2 for the 0-dim
  // tensors of cell.py:13-14; h**(-2) == 1/(h*h) (operators.py:7); in the
  // compute type.
  dtm2 = T(1) / (dt * dt);
  dtm1 = T(1) / dt;
  s2 = T(2) * dtm2;
  t1 = T(1) / (h * h);
  t4 = T(-4) * t1;
}
}  // extern "C++"

int wt_check_finite(const float *p, int64_t n, int32_t *flag, void *stream) {
  if (!p || !flag || n < 0) {
    set_error("wt_check_finite: invalid argument");
    return WT_EINVAL;
  }
  WT_HIP(launch_nonfinite(p, n, flag, (hipStream_t)stream));
  return WT_OK;
}

int wt_check_finite_f64(const double *p, int64_t n, int32_t *flag, void *stream) {
  if (!p || !flag || n < 0) {
    set_error("wt_check_finite_f64: invalid argument");
    return WT_EINVAL;
  }
  WT_HIP(launch_nonfinite(p, n, flag, (hipStream_t)stream));
  return WT_OK;
}

int wt_coef_linear(const float *b, const float *c, int64_t n, float dt, float *A, float *Bm,
                   float *C2, void *stream) {
  if (!b || !c || !A || !Bm || !C2 || n < 0 || !(dt > 0)) {
    set_error("wt_coef_linear: invalid argument");
    return WT_EINVAL;
  }
  float dtm1, dtm2, s2, t1, t4;
  scalars(dt, 1.0f, dtm1, dtm2, s2, t1, t4);
  WT_HIP(launch_coef_linear(b, c, n, dtm1, dtm2, A, Bm, C2, (hipStream_t)stream));
  return WT_OK;
}

int wt_laplacian(const float *u, float *out, int32_t B, int32_t Nx, int32_t Ny, float h,
                 void *stream) {
  if (!u || !out || u == out || B < 0 || Nx <= 0 || Ny <= 0 || !(h > 0) || !std::isfinite(h)) {
    set_error("wt_laplacian: invalid argument");
    return WT_EINVAL;
  }
  float dtm1, dtm2, s2, t1, t4;
  scalars(1.0f, h, dtm1, dtm2, s2, t1, t4);
  WT_HIP(launch_laplacian(u, B, Nx, Ny, t1, t4, out, (hipStream_t)stream));
  return WT_OK;
}

int wt_laplacian_f64(const double *u, double *out, int32_t B, int32_t Nx, int32_t Ny, double h,
                     void *stream) {
  if (!u || !out || u == out || B < 0 || Nx <= 0 || Ny <= 0 || !(h > 0) || !std::isfinite(h)) {
    set_error("wt_laplacian_f64: invalid argument");
    return WT_EINVAL;
  }
  double dtm1, dtm2, s2, t1, t4;
  scalars(1.0, h, dtm1, dtm2, s2, t1, t4);
  WT_HIP(launch_laplacian(u, B, Nx, Ny, t1, t4, out, (hipStream_t)stream));
  return WT_OK;
}

int wt_step_fwd_f64(const double *b, const double *c, int32_t bc_per_sample, const double *y1,
                    const double *y2, int32_t B, int32_t Nx, int32_t Ny, double dt, double h,
                    double *y, void *stream) {
  if (!b || !c || !y1 || !y2 || !y || B < 0 || Nx <= 0 || Ny <= 0 || !(dt > 0) || !(h > 0)) {
    set_error("wt_step_fwd_f64: invalid argument");
    return WT_EINVAL;
  }
  double dtm1, dtm2, s2, t1, t4;
  scalars(dt, h, dtm1, dtm2, s2, t1, t4);
  WT_HIP(launch_step_fwd(h == 1.0, b, c, bc_per_sample, y1, y2, B, Nx, Ny, dtm1, dtm2, s2, t1,
                         t4, y, (hipStream_t)stream));
  return WT_OK;
}

int wt_step_bwd_f64(const double *b, const double *c, int32_t bc_per_sample, const double *y1,
                    const double *y2, const double *g, int32_t B, int32_t Nx, int32_t Ny,
                    double dt, double h, double *grad_b, double *grad_c, double *grad_y1,
                    double *grad_y2, void *stream) {
  if (!b || !c || !y1 || !y2 || !g || B < 0 || Nx <= 0 || Ny <= 0 || !(dt > 0) || !(h > 0)) {
    set_error("wt_step_bwd_f64: invalid argument");
    return WT_EINVAL;
  }
  double dtm1, dtm2, s2, t1, t4;
  scalars(dt, h, dtm1, dtm2, s2, t1, t4);
  WT_HIP(launch_step_bwd(h == 1.0, b, c, bc_per_sample, y1, y2, g, B, Nx, Ny, dt, t1, t4,
                         grad_b, grad_c, grad_y1, grad_y2, (hipStream_t)stream));
  return WT_OK;
}

int wt_step_fwd(const float *b, const float *c, int32_t bc_per_sample, const float *y1,
                const float *y2, int32_t B, int32_t Nx, int32_t Ny, float dt, float h, float *y,
                void *stream) {
  if (!b || !c || !y1 || !y2 || !y || B < 0 || Nx <= 0 || Ny <= 0 || !(dt > 0) || !(h > 0)) {
    set_error("wt_step_fwd: invalid argument");
    return WT_EINVAL;
  }
  float dtm1, dtm2, s2, t1, t4;
  scalars(dt, h, dtm1, dtm2, s2, t1, t4);
  WT_HIP(launch_step_fwd(h == 1.0f, b, c, bc_per_sample, y1, y2, B, Nx, Ny, dtm1, dtm2, s2, t1,
                         t4, y, (hipStream_t)stream));
  return WT_OK;
}

int wt_step_bwd(const float *b, const float *c, int32_t bc_per_sample, const float *y1,
                const float *y2, const float *g, int32_t B, int32_t Nx, int32_t Ny, float dt,
                float h, float *grad_b, float *grad_c, float *grad_y1, float *grad_y2,
                void *stream) {
  if (!b || !c || !y1 || !y2 || !g || B < 0 || Nx <= 0 || Ny <= 0 || !(dt > 0) || !(h > 0)) {
    set_error("wt_step_bwd: invalid argument");
    return WT_EINVAL;
  }
  float dtm1, dtm2, s2, t1, t4;
  scalars(dt, h, dtm1, dtm2, s2, t1, t4);
  WT_HIP(launch_step_bwd(h == 1.0f, b, c, bc_per_sample, y1, y2, g, B, Nx, Ny, dt, t1, t4,
                         grad_b, grad_c, grad_y1, grad_y2, (hipStream_t)stream));
  return WT_OK;
}

}  // extern "C"

static int plan_upload(Plan *p);

// Host half of plan creation: validation, special-pixel tables, tiling and
// variant choice, workspace layout.  No device call except hipGetDevice.
static int plan_build(const wt_config *cfg, int32_t B, int32_t T, float cmax_geom,
                      const wt_tuning *tuning, Plan **plan_out) {
  if (!cfg || !plan_out) {
    set_error("wt_plan_create: null argument");
    return WT_EINVAL;
  }
  *plan_out = nullptr;
  if (B <= 0 || T <= 0 || cfg->Nx < 3 || cfg->Ny < 3 || !(cfg->dt > 0) || !(cfg->h > 0)) {
    set_error("wt_plan_create: invalid shape B=%d T=%d Nx=%d Ny=%d dt=%g h=%g", B, T, cfg->Nx,
              cfg->Ny, cfg->dt, cfg->h);
    return WT_EINVAL;
  }
  if ((int64_t)cfg->Nx * (cfg->Ny + 1) >= (1 << 22)) {
    set_error("wt_plan_create: grid %dx%d too large for int32 cell indexing", cfg->Nx, cfg->Ny);
    return WT_EUNSUPPORTED;
  }
  // CFL check of cell.py:66-73: dt > h / (cmax * sqrt(2)) is rejected.
  if (cmax_geom > 0 && (double)cfg->dt > 1.0 / cmax_geom * cfg->h / std::sqrt(2.0)) {
    set_error("The spatial discretization defined by the geometry `h = %f` and the temporal "
              "discretization defined by the model `dt = %f` do not satisfy the CFL stability "
              "criteria",
              cfg->h, cfg->dt);
    return WT_ECFL;
  }
  if (cfg->nsrc_px < 0 || cfg->ntap < 0 || (cfg->nsrc_px > 0 && !cfg->src_px) ||
      (cfg->ntap > 0 && !cfg->taps)) {
    set_error("wt_plan_create: invalid source/probe tables");
    return WT_EINVAL;
  }
  Plan *p = new Plan();
  p->B = B;
  p->T = T;
  p->Nx = cfg->Nx;
  p->Ny = cfg->Ny;
  p->dt = cfg->dt;
  p->h = cfg->h;
  p->unit_h = cfg->h == 1.0f;
  p->sat_b0 = cfg->sat_b0;
  p->sat_uth = cfg->sat_uth;
  p->c_nl = cfg->c_nl;
  p->cmin = cfg->c_min;
  p->cmax = cfg->c_max;
  int nl = 0;
  if (cfg->sat_b0 > 0) nl |= 1;
  if (cfg->c_nl != 0) nl |= 2;
  if (cfg->c_clamp) nl |= 4;
  if (nl == 4) {  // clamp of a time-invariant c: applied once to c_linear
    p->clamp_only = true;
    nl = 0;
  }
  p->nl = nl;
  scalars(p->dt, p->h, p->dtm1, p->dtm2, p->s2, p->t1, p->t4);
  (void)hipGetDevice(&p->dev);

  // ---- special pixels: sources (with multiplicity) and probe taps
  struct Px { int x, y, mult, slot, tlo, thi; };
  std::vector<Px> px;
  auto find_or_add = [&](int x, int y) -> int {
    for (size_t i = 0; i < px.size(); ++i)
      if (px[i].x == x && px[i].y == y) return (int)i;
    px.push_back(Px{x, y, 0, -1, 0, 0});
    return (int)px.size() - 1;
  };
  for (int k = 0; k < cfg->nsrc_px; ++k) {
    int x = cfg->src_px[3 * k], y = cfg->src_px[3 * k + 1];
    if (x < 0 || x >= p->Nx || y < 0 || y >= p->Ny) {
      set_error("source pixel (%d, %d) outside the %dx%d domain", x, y, p->Nx, p->Ny);
      free_plan(p);
      return WT_EINVAL;
    }
    px[find_or_add(x, y)].mult += 1;
  }
  struct Tp { int x, y, kind, col; };
  std::vector<Tp> tp;
  for (int k = 0; k < cfg->ntap; ++k) {
    int x = cfg->taps[3 * k], y = cfg->taps[3 * k + 1], kind = cfg->taps[3 * k + 2];
    if (x < 0 || x >= p->Nx || y < 0 || y >= p->Ny || (kind != 0 && kind != 1)) {
      set_error("probe tap (%d, %d, kind %d) invalid for the %dx%d domain", x, y, kind, p->Nx,
                p->Ny);
      free_plan(p);
      return WT_EINVAL;
    }
    find_or_add(x, y);
    tp.push_back(Tp{x, y, kind, k});
  }
  std::sort(px.begin(), px.end(),
            [](const Px &a, const Px &b) { return a.x != b.x ? a.x < b.x : a.y < b.y; });
  std::sort(tp.begin(), tp.end(), [](const Tp &a, const Tp &b) {
    return a.x != b.x ? a.x < b.x : (a.y != b.y ? a.y < b.y : a.col < b.col);
  });
  p->nspx = (int)px.size();
  p->ntap = (int)tp.size();
  int nslot = 0;
  for (auto &q : px) {
    if (q.mult > 0) {
      q.slot = nslot++;
      p->h_mult.push_back(q.mult);
    }
    int lo = 0;
    while (lo < (int)tp.size() && (tp[lo].x != q.x || tp[lo].y != q.y)) ++lo;
    int hi = lo;
    while (hi < (int)tp.size() && tp[hi].x == q.x && tp[hi].y == q.y) ++hi;
    q.tlo = lo;
    q.thi = hi;
  }
  p->nslot = nslot;
  for (auto &q : px) {
    int32_t r[SPX_W] = {q.x, q.y, q.mult, q.slot, q.tlo, q.thi, 0, 0};
    p->h_spx.insert(p->h_spx.end(), r, r + SPX_W);
  }
  for (auto &t : tp) {
    int32_t r[TAP_W] = {t.x, t.y, t.kind, t.col};
    p->h_taps.insert(p->h_taps.end(), r, r + TAP_W);
  }

  // ---- tuning
  wt_tuning tun{};
  if (tuning) tun = *tuning;
  p->K = tun.ck_every > 0 ? tun.ck_every : 64;
  if (p->K > T) p->K = T;
  p->nseg = (T + p->K - 1) / p->K;
  bool tiled = !tun.plain && p->unit_h;
  if (tiled) {
    int strips = tun.strips, tb = tun.tblock, nt = tun.nt, cpt = tun.cpt;
    // the nonlinear cell forms its coefficients (three divisions) per cell
    // and step: halo work costs more there, so shallower blocks pay
    if (tb <= 0) tb = p->nl ? 8 : 16;
    if (tb > p->K) tb = p->K;
    if (tun.kind == 0 && tb > 256) tb = 256;  // inputs/seeds of a launch are staged in LDS
    // Heuristic for the column-segment family: s x c tiles per sample, the
    // first in tile_order's order giving one workgroup per CU (B*s*c >= 256)
    // among the tilings whose own rows and columns are at least twice the
    // halo (else the last of those); further along the order only if no
    // block variant holds the tiles.  (Measured on the nonlinear 150x150
    // batch-32 step: 4x2 tiles, 256 workgroups, beat 4x4 square tiles, 512,
    // by 17%.)
    bool square = tun.kind == 0 && strips <= 0 && tun.tcols <= 0;
    std::vector<std::pair<int, int>> order;
    size_t oi = 0;
    if (square) {
      order = tile_order(p->Nx, p->Ny);
      for (size_t i = 1; i < order.size(); ++i) {
        if ((int64_t)B * order[oi].first * order[oi].second >= 256) break;
        if (std::min(p->Nx / order[i].first, p->Ny / order[i].second) >= 2 * tb) oi = i;
      }
      strips = order[oi].first;
    } else if (strips <= 0) {
      // row strips: enough workgroups to cover the 256 CUs
      strips = 1;
      while (strips < p->Nx / 4 && (int64_t)B * strips < 256) ++strips;
    }
    int tcols = (tun.kind == 0 && tun.tcols > 0) ? tun.tcols : 1;
    bool ok = false;
    for (; square ? oi < order.size() : strips <= p->Nx; square ? (void)++oi : (void)++strips) {
      if (square) {
        strips = order[oi].first;
        tcols = order[oi].second;
      }
      build_tiles(*p, strips, tcols, tb);
      if (nt > 0 && cpt > 0) {
        ok = tun.kind == 1 ? tiles_ok(*p, nt, cpt) : tiles_ok_cs(*p, nt, cpt);
        if (!ok && tun.strips > 0) break;
      } else {
        ok = tun.kind == 1 ? pick_variant(*p, nt, cpt) : pick_variant_cs(*p, nt, cpt);
      }
      if (ok) break;
      if (tun.strips > 0) break;
    }
    if (!ok) {
      if (tun.strips > 0 || tun.nt > 0) {
        set_error("wt_plan_create: tuning (strips=%d tblock=%d nt=%d cpt=%d) does not fit",
                  tun.strips, tun.tblock, tun.nt, tun.cpt);
        free_plan(p);
        return WT_EUNSUPPORTED;
      }
      tiled = false;
    } else {
      if (tun.kind == 0 && p->nl == 0 && !(tuning && tuning->tblock > 0) &&
          !(tuning && tuning->nt > 0)) {
        // deepest forward block (fewest launches per checkpoint segment)
        // whose regions still fit 1024 threads at <= 12 rows per thread,
        // while a tile's own rows and columns stay at least twice its halo
        const int own = std::min(p->Nx / strips, p->Ny / tcols);
        // a single tile per sample has no halo: one launch per segment
        const int whole = (strips == 1 && tcols == 1) ? std::min(p->K, 256) : 0;
        for (int d : {whole, 32, 24}) {
          if (d <= tb || d > p->K || (d != whole && own < 2 * d)) continue;
          build_tiles(*p, strips, tcols, d);
          int r = 0;
          if (fit_cs(*p, 12, r)) {
            tb = d;
            cpt = r;
            nt = 1024;
            break;
          }
        }
        build_tiles(*p, strips, tcols, tb);
      }
      tun.strips = strips;
      tun.tcols = tcols;
      tun.tblock = tb;
      tun.nt = nt;
      tun.cpt = cpt;
    }
  }
  if (!tiled) {
    tun.plain = 1;
    tun.tcols = 0;
    tun.strips = 0;
    tun.tblock = 1;
    tun.nt = 256;
    tun.cpt = 1;
  }
  tun.ck_every = p->K;
  if (tiled && tun.kind == 0) {
    // adjoint tiling: defaults to the forward's
    std::vector<int32_t> ft = p->h_tiles;  // forward tiles
    int fnt = p->ntiles, fRH = p->RHmax, fRW = p->RWmax, fLW = p->LWmax;
    int as = tun.a_strips > 0 ? tun.a_strips : tun.strips;
    int ac = tun.a_tcols > 0 ? tun.a_tcols : (tun.tcols > 0 ? tun.tcols : 1);
    // the adjoint holds three fields per row in registers: by default the
    // deepest block whose tiles fit 1024 threads at <= 10 rows per thread
    // (deeper blocks, fewer launches: each launch pays a set-up and a
    // write-back of its region)
    int atb = tun.a_tblock > 0 ? tun.a_tblock
                               : (tuning && tuning->tblock > 0 ? tun.tblock
                                                               : std::min(tun.tblock, 12));
    if (tun.a_tblock <= 0 && !(tuning && tuning->tblock > 0) && tun.a_nt <= 0) {
      const int as0 = tun.a_strips > 0 ? tun.a_strips : tun.strips;
      const int ac0 = tun.a_tcols > 0 ? tun.a_tcols : (tun.tcols > 0 ? tun.tcols : 1);
      const int own = std::min(p->Nx / as0, p->Ny / ac0);
      const int whole = (as0 == 1 && ac0 == 1) ? std::min(p->K, 256) : 0;
      for (int d : {whole, 24, 22, 20, 18, 16, 15, 14, 13, 12, 10, 8, 6, 4, 2, 1}) {
        if (d <= 0 || d > p->K || (d != whole && d > 1 && own < 2 * d)) continue;
        build_tiles(*p, as0, ac0, d);
        int r = 0;
        if (fit_cs(*p, 10, r, false)) {
          atb = d;
          break;
        }
      }
    }
    if (atb > p->K) atb = p->K;
    if (atb > 256) atb = 256;
    int ant = tun.a_nt > 0 ? tun.a_nt : tun.nt;
    int acp = tun.a_cpt > 0 ? tun.a_cpt : tun.cpt;
    bool aok = false;
    if (p->nl) {
      // nonlinear adjoint (k_adj_cs_nl): 4 rows per thread at 1024 threads.
      // Unless given, the block is the deepest of at most 8 steps whose
      // regions fit on the forward's tiling (at least 4 steps: deeper blocks,
      // fewer launches, each with its set-up); failing that, 8-step blocks
      // on the first larger tiling in tile_order that fits.
      ant = 1024;
      acp = 4;
      auto fits = [&](int q, int qc, int d) {
        build_tiles(*p, q, qc, d);
        return tiles_ok_cs(*p, ant, acp, false) &&
               cs_nl_shmem_bytes(acp, p->RHmax, p->RWmax) + 1024 <= LDS_LIMIT;
      };
      const bool fixed = tun.a_strips > 0;
      const int dmax = std::min(8, p->K), dmin = fixed ? 1 : std::min(4, p->K);
      if (tun.a_tblock > 0) {
        aok = fits(as, ac, atb);
      } else {
        for (int d = dmax; d >= dmin && !aok; --d)
          if (fits(as, ac, d)) {
            atb = d;
            aok = true;
          }
        if (!aok) atb = dmax;
      }
      if (!aok && !fixed) {
        for (const auto &t : tile_order(p->Nx, p->Ny)) {
          if (t.first * t.second <= as * ac) continue;
          if (fits(t.first, t.second, atb)) {
            as = t.first;
            ac = t.second;
            aok = true;
            break;
          }
        }
      }
    } else {
      build_tiles(*p, as, ac, atb);
      // an explicit variant (adjoint's own, or the forward's) is tried as is;
      // otherwise the fewest rows per thread that fit
      if (tun.a_nt > 0 || (tuning && tuning->nt > 0)) aok = tiles_ok_cs(*p, ant, acp, false);
      if (!aok && tun.a_nt <= 0) aok = pick_variant_cs(*p, ant, acp, false);
    }
    if (!aok) {
      set_error("wt_plan_create: adjoint tuning (strips=%d tcols=%d tblock=%d nt=%d cpt=%d) "
                "does not fit", as, ac, atb, ant, acp);
      free_plan(p);
      return WT_EUNSUPPORTED;
    }
    p->h_tiles_a = p->h_tiles;
    p->ntiles_a = p->ntiles;
    p->RHmax_a = p->RHmax;
    p->RWmax_a = p->RWmax;
    p->tb_a = atb;
    p->nt_a = ant;
    p->cpt_a = acp;
    tun.a_strips = as; tun.a_tcols = ac; tun.a_tblock = atb; tun.a_nt = ant; tun.a_cpt = acp;
    p->h_tiles = ft;
    p->ntiles = fnt;
    p->RHmax = fRH;
    p->RWmax = fRW;
    p->LWmax = fLW;
  } else {
    p->h_tiles_a = p->h_tiles;
    p->ntiles_a = p->ntiles;
    p->RHmax_a = p->RHmax;
    p->RWmax_a = p->RWmax;
    p->tb_a = tun.tblock;
    p->nt_a = tun.nt;
    p->cpt_a = tun.cpt;
    tun.a_strips = tun.strips; tun.a_tcols = tun.tcols; tun.a_tblock = tun.tblock;
    tun.a_nt = tun.nt; tun.a_cpt = tun.cpt;
  }
  p->tun = tun;
  p->tiled = tiled;
  plan_layout(*p);
  *plan_out = p;
  return WT_OK;
}

extern "C" {

int wt_plan_resolve(const wt_config *cfg, int32_t B, int32_t T, float cmax_geom,
                    const wt_tuning *tuning, wt_tuning *out, int64_t *saved, int64_t *temp) {
  Plan *p = nullptr;
  int rc = plan_build(cfg, B, T, cmax_geom, tuning, &p);
  if (rc) {
    // the plan builder's messages name wt_plan_create; name this entry point
    const std::string create = "wt_plan_create";
    if (g_last_error.compare(0, create.size(), create) == 0)
      g_last_error = "wt_plan_resolve" + g_last_error.substr(create.size());
    return rc;
  }
  if (out) *out = p->tun;
  if (saved) *saved = (int64_t)p->saved_bytes;
  if (temp) *temp = (int64_t)p->temp_bytes;
  free_plan(p);
  return WT_OK;
}

int wt_plan_create(const wt_config *cfg, int32_t B, int32_t T, float cmax_geom,
                   const wt_tuning *tuning, void **plan_out) {
  if (!plan_out) {
    set_error("wt_plan_create: null argument");
    return WT_EINVAL;
  }
  Plan *p = nullptr;
  int rc0 = plan_build(cfg, B, T, cmax_geom, tuning, &p);
  if (rc0) return rc0;
  rc0 = plan_upload(p);
  if (rc0) {
    free_plan(p);
    return rc0;
  }
  *plan_out = p;
  return WT_OK;
}

}  // extern "C"

// Column-segment thread mapping and LDS stride of every tile (descriptor
// words 12, 13) for R rows per thread, padding a tile's columns only within
// the waves it occupies anyway (an extra active wave costs these VALU-bound
// blocks more than the bank conflicts it removes).
static void set_tile_layouts(std::vector<int32_t> &tiles, int R, int RHmax, int RWmax) {
  if (R <= 0) return;
  const int launch_thr = cs_threads(R, RHmax, RWmax);
  for (size_t i = 0; i + TILE_W <= tiles.size(); i += TILE_W) {
    int RWm = 0, CS = 0;
    const int maxthr = std::min(launch_thr, cs_threads(R, tiles[i + 1], tiles[i + 3]));
    cs_layout(tiles[i + 1], tiles[i + 3], R, maxthr, RWm, CS);
    tiles[i + 12] = RWm;
    tiles[i + 13] = CS;
  }
}

// Device copies of a plan's tables (synchronous).
static int plan_upload(Plan *p) {
  if (p->tiled && p->tun.kind == 0) {
    // forward tiles only: the adjoints wait little on LDS (measured: no gain)
    set_tile_layouts(p->h_tiles, p->tun.cpt, p->RHmax, p->RWmax);
  }
  auto up = [&](int32_t **dst, const std::vector<int32_t> &v) -> int {
    size_t n = std::max<size_t>(v.size(), 1) * sizeof(int32_t);
    hipError_t e = hipMalloc((void **)dst, n);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(plan table)");
    if (!v.empty()) {
      e = hipMemcpy(*dst, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice);
      if (e != hipSuccess) return hip_fail(e, "hipMemcpy(plan table)");
    }
    return WT_OK;
  };
  std::vector<int32_t> spmap((size_t)p->Nx * p->Ny, -1);
  for (int k = 0; k < p->nspx; ++k)
    spmap[(size_t)p->h_spx[k * SPX_W] * p->Ny + p->h_spx[k * SPX_W + 1]] = k;
  int rc;
  if ((rc = up(&p->d_spx, p->h_spx)) || (rc = up(&p->d_taps, p->h_taps)) ||
      (rc = up(&p->d_spmap, spmap)) || (rc = up(&p->d_mult, p->h_mult)) ||
      (rc = up(&p->d_tiles, p->h_tiles)) || (rc = up(&p->d_tiles_a, p->h_tiles_a)))
    return rc;
  return WT_OK;
}

namespace wt {
int plan_view(const void *plan, PlanView *out) {
  const Plan *p = (const Plan *)plan;
  if (!p || !out) return WT_EINVAL;
  out->B = p->B; out->T = p->T; out->Nx = p->Nx; out->Ny = p->Ny;
  out->nlf = p->nl | (p->clamp_only ? 4 : 0);
  out->nspx = p->nspx; out->ntap = p->ntap; out->nslot = p->nslot;
  out->spx = p->d_spx; out->taps = p->d_taps; out->spmap = p->d_spmap; out->mult = p->d_mult;
  return WT_OK;
}
}  // namespace wt

extern "C" {

int wt_plan_destroy(void *plan) {
  free_plan((Plan *)plan);
  return WT_OK;
}

int wt_plan_tuning(const void *plan, wt_tuning *out) {
  if (!plan || !out) {
    set_error("wt_plan_tuning: null argument");
    return WT_EINVAL;
  }
  *out = ((const Plan *)plan)->tun;
  return WT_OK;
}

int wt_workspace_bytes(const void *plan, int32_t with_grad, int64_t *saved, int64_t *temp) {
  if (!plan) {
    set_error("wt_workspace_bytes: null plan");
    return WT_EINVAL;
  }
  const Plan *p = (const Plan *)plan;
  if (saved) *saved = with_grad ? (int64_t)p->saved_bytes : 0;
  if (temp) *temp = with_grad ? (int64_t)p->temp_bytes : (int64_t)p->temp_bytes_nograd;
  return WT_OK;
}

// ---------------------------------------------------------------- loops

namespace {

struct Ws {
  float *A, *Bm, *C2, *rhoc;  // coefficients
  float *ck;                  // checkpoints (nseg, 2, B, F)
  float *yraw;
  float *st;                  // 6 state fields
  float *lam;                 // 4 adjoint fields
  float *wl;                  // 2 fields
  float *gacc;                // 1 (linear) or 3 fields
  float *seg;                 // segment record
  float *rec;                 // last segment's record from the forward (or null)
  float *gsrc;
};

Ws layout(const Plan &p, void *saved, void *temp, bool with_grad) {
  Ws w{};
  char *t = (char *)temp;
  w.st = (float *)t;
  t += p.t_state;
  float *coef;
  if (with_grad) {
    char *s = (char *)saved;
    coef = (float *)s;
    s += p.saved_coef;
    w.ck = (float *)s;
    s += p.saved_ck;
    w.yraw = (float *)s;
    s += p.saved_yraw;
    w.rec = p.saved_rec ? (float *)s : nullptr;
    w.lam = (float *)t;
    t += p.t_lam;
    w.wl = (float *)t;
    t += p.t_wl;
    w.gacc = (float *)t;
    t += p.t_gacc;
    w.seg = w.rec ? w.rec : (float *)t;
    t += p.t_seg;
    w.gsrc = (float *)t;
    t += p.t_gsrc;
  } else {
    coef = (float *)t;
  }
  if (p.nl) {
    w.rhoc = coef;
    w.A = w.Bm = w.C2 = nullptr;
  } else {
    w.A = coef;
    w.Bm = coef + p.F;
    w.C2 = coef + 2 * p.F;
    w.rhoc = nullptr;
  }
  return w;
}

FwdArgs base_fwd(const Plan &p, const Ws &w, const float *b, const float *c, const float *rho,
                 const float *x) {
  FwdArgs a{};
  a.B = p.B; a.Nx = p.Nx; a.Ny = p.Ny; a.T = p.T;
  a.ntiles = p.ntiles; a.tiles = p.d_tiles; a.spx = p.d_spx; a.taps = p.d_taps;
  a.ntap = p.ntap;
  a.A = w.A; a.Bm = w.Bm; a.C2 = w.C2;
  a.bgeom = b; a.clin = c; a.rho = rho; a.rhoc = w.rhoc;
  a.s2 = p.s2; a.dtm1 = p.dtm1; a.dtm2 = p.dtm2; a.t1 = p.t1; a.t4 = p.t4;
  a.sat_b0 = p.sat_b0; a.sat_uth = p.sat_uth; a.cmin = p.cmin; a.cmax = p.cmax;
  a.nlf = p.nl;
  a.x = x;
  a.fstride = 1;
  return a;
}

// Runs forward steps [n0, n1) from the state (in1, in2) = (u_n0, u_{n0-1}).
// Tiled: launches of tblock steps; the first reads (in1, in2), intermediate
// states ping-pong through tmp[0..3], and the last launch writes (out1, out2)
// when given.  Plain: one launch per step rotating through tmp[0..2]; the
// final state is copied to (out1, out2) when given.
int run_forward(const Plan &p, FwdArgs a, int n0, int n1, const float *in1, const float *in2,
                float *out1, float *out2, float *tmp[4], bool rec, int prof_cls, hipStream_t s) {
  const int64_t BF = (int64_t)p.B * p.F;
  if (p.tiled) {
    const float *i1 = in1, *i2 = in2;
    int pp = 0;
    int n = n0;
    while (n < n1) {
      int k = std::min(p.tun.tblock, n1 - n);
      bool last = n + k >= n1;
      float *o1 = (last && out1) ? out1 : tmp[2 * pp];
      float *o2 = (last && out2) ? out2 : tmp[2 * pp + 1];
      a.in1 = i1; a.in2 = i2; a.out1 = o1; a.out2 = o2;
      a.n0 = n; a.nsteps = k;
      ProfScope ps(prof_cls, s, (int64_t)p.B * p.Nx * p.Ny * k);
      if (p.tun.kind == 1)
        WT_HIP(launch_fwd_tiled(p.tun.nt, p.tun.cpt, p.unit_h, p.nl, rec, a, p.LWmax, s));
      else
        WT_HIP(launch_fwd_cs(p.tun.nt, p.tun.cpt, p.nl, rec, a, p.RHmax, p.RWmax, s));
      i1 = o1;
      i2 = o2;
      pp ^= 1;
      n += k;
    }
  } else {
    // u_n in r1, u_{n-1} in r2, output into r3; the inputs are copied into
    // the rotation buffers first so the caller's state is never written
    float *r1 = tmp[0], *r2 = tmp[1], *r3 = tmp[2];
    WT_HIP(hipMemcpyAsync(r1, in1, BF * sizeof(float), hipMemcpyDeviceToDevice, s));
    WT_HIP(hipMemcpyAsync(r2, in2, BF * sizeof(float), hipMemcpyDeviceToDevice, s));
    for (int n = n0; n < n1; ++n) {
      a.in1 = r1; a.in2 = r2; a.out1 = r3;
      ProfScope ps(prof_cls, s, (int64_t)p.B * p.Nx * p.Ny);
      WT_HIP(launch_fwd_plain(p.unit_h, p.nl, rec, a, p.d_spmap, n, s));
      float *t = r2; r2 = r1; r1 = r3; r3 = t;
    }
    if (out1) WT_HIP(hipMemcpyAsync(out1, r1, BF * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (out2) WT_HIP(hipMemcpyAsync(out2, r2, BF * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  return WT_OK;
}

int prepare_coef(const Plan &p, const Ws &w, const float *b, const float *c, const float *rho,
                 float *tmp, hipStream_t s) {
  int64_t F = (int64_t)p.F;
  if (p.nl) {
    WT_HIP(launch_mul_scalar(rho, p.c_nl, F, w.rhoc, s));
  } else {
    const float *cc = c;
    if (p.clamp_only) {
      WT_HIP(launch_clamp(c, p.cmin, p.cmax, F, tmp, s));
      cc = tmp;
    }
    WT_HIP(launch_coef_linear(b, cc, F, p.dtm1, p.dtm2, w.A, w.Bm, w.C2, s));
  }
  return WT_OK;
}

// Reverse sweep of the nonlinear cell: per segment, recompute from the
// checkpoint recording u_n and lap(u_n), then one k_loop_adj_nl_plain launch
// per step (the step's coefficients depend on u, so the linear kernels'
// precomputed a1/a2/w do not apply).
int adj_nonlinear(const Plan &p, const float *b, const float *c, const float *rho,
                  const float *x, const float *ybar, const float *fbar, void *saved, void *temp,
                  float *grad_c, float *grad_b, float *grad_rho, float *grad_x, hipStream_t s) {
  Ws w = layout(p, saved, temp, true);
  const int64_t BF = (int64_t)p.B * p.F;
  const int T = p.T, K = p.K;
  const size_t segcells = (size_t)p.B * K * p.Nx * (p.Ny + 1);
  float *st[8];
  for (int i = 0; i < 8; ++i) st[i] = w.st + i * BF;
  // PL, AL (two parity slots), w*lambda ping-pong, three accumulators
  WT_HIP(launch_fill(w.lam, 3 * BF, 0.0f, s));
  WT_HIP(launch_fill(w.wl, 2 * BF, 0.0f, s));
  WT_HIP(launch_fill(w.gacc, 3 * BF, 0.0f, s));
  WT_HIP(launch_fill(w.gsrc, (int64_t)p.B * T * std::max(p.nslot, 1), 0.0f, s));

  FwdArgs fa = base_fwd(p, w, b, c, rho, x);
  fa.rec_lap = w.seg;
  fa.rec_u = w.seg + segcells;
  fa.rec_K = K;
  // the record's layout follows the recompute kernels (column-segment rows of
  // Ny, else Ny + 1); column-segment plans also run the blocked adjoint,
  // except with dense field seeds, which the per-step kernel reads
  const bool cs_rec = p.tiled && p.tun.kind == 0;
  const bool cs = cs_rec && fbar == nullptr;
  NlAdjBufs nb{};
  nb.rec_ld = cs_rec ? p.Ny : p.Ny + 1;
  nb.rec_slice = (int64_t)p.Nx * nb.rec_ld;
  AdjArgs aa{};
  aa.B = p.B; aa.Nx = p.Nx; aa.Ny = p.Ny; aa.T = T;
  aa.spx = p.d_spx; aa.taps = p.d_taps; aa.ntap = p.ntap;
  aa.nsrcslot = std::max(p.nslot, 1);
  aa.bgeom = b; aa.clin = c; aa.rho = rho;
  aa.s2 = p.s2; aa.dtm1 = p.dtm1; aa.dtm2 = p.dtm2; aa.t1 = p.t1; aa.t4 = p.t4;
  aa.sat_b0 = p.sat_b0; aa.sat_uth = p.sat_uth; aa.cmin = p.cmin; aa.cmax = p.cmax;
  aa.c_nl = p.c_nl; aa.nlf = p.nl;
  aa.ybar = ybar; aa.yraw = w.yraw;
  aa.seg_lap = fa.rec_lap; aa.seg_u = fa.rec_u; aa.seg_K = K;
  aa.gacc = w.gacc; aa.gacc_b = w.gacc + BF; aa.gacc_rho = w.gacc + 2 * BF;
  aa.gsrc = w.gsrc;
  if (cs) {
    // temporally blocked (k_adj_cs_nl): lambdas ping-pong in lam[0..3], the
    // carried coefficients (w, P, a2 of two steps) in two sets of four
    WT_HIP(launch_fill(w.lam + 3 * BF, 9 * BF, 0.0f, s));  // (the first 3 are filled above)
    float *l[4], *cA[4], *cB[4];
    for (int i = 0; i < 4; ++i) {
      l[i] = w.lam + i * BF;
      cA[i] = w.lam + (4 + i) * BF;
      cB[i] = w.lam + (8 + i) * BF;
    }
    aa.ntiles = p.ntiles_a;
    aa.tiles = p.d_tiles_a;
    for (int sg = p.nseg - 1; sg >= 0; --sg) {
      const int n0 = sg * K, n1 = std::min(T, n0 + K);
      const float *ck = w.ck + (size_t)sg * 2 * BF;
      fa.rec_n0 = n0;
      float *tmp[4] = {st[0], st[1], st[2], st[3]};
      // the last segment's record came with the forward
      const bool from_fwd = w.rec && sg == p.nseg - 1;
      if (!from_fwd) {
        int rc = run_forward(p, fa, n0, n1, ck, ck + BF, nullptr, nullptr, tmp, true, 1, s);
        if (rc) return rc;
      }
      aa.seg_lap = from_fwd ? w.rec : fa.rec_lap;
      aa.seg_u = from_fwd ? w.rec + segcells : fa.rec_u;
      aa.seg_n0 = n0;
      for (int m = n1; m > n0;) {
        const int k = std::min(p.tb_a, m - n0);
        aa.in1 = l[0]; aa.in2 = l[1]; aa.out1 = l[2]; aa.out2 = l[3];
        aa.m_hi = m;
        aa.nsteps = k;
        NlCarry cy{cA[0], cA[1], cA[2], cA[3], cB[0], cB[1], cB[2], cB[3], ck + BF};
        ProfScope ps(2, s, BF * k);
        WT_HIP(launch_adj_cs_nl(p.nt_a, p.cpt_a, grad_b != nullptr, aa, cy, p.RHmax_a,
                                p.RWmax_a, s));
        std::swap(l[0], l[2]);
        std::swap(l[1], l[3]);
        for (int i = 0; i < 4; ++i) std::swap(cA[i], cB[i]);
        m -= k;
      }
    }
  } else {
  float *wl1 = w.wl, *wl2 = w.wl + BF;
  nb.PL = w.lam;
  for (int sg = p.nseg - 1; sg >= 0; --sg) {
    const int n0 = sg * K, n1 = std::min(T, n0 + K);
    const float *ck = w.ck + (size_t)sg * 2 * BF;
    fa.rec_n0 = n0;
    float *tmp[4] = {st[0], st[1], st[2], st[3]};
    const bool from_fwd = w.rec && sg == p.nseg - 1;
    if (!from_fwd) {
      int rc = run_forward(p, fa, n0, n1, ck, ck + BF, nullptr, nullptr, tmp, true, 1, s);
      if (rc) return rc;
    }
    aa.seg_lap = from_fwd ? w.rec : fa.rec_lap;
    aa.seg_u = from_fwd ? w.rec + segcells : fa.rec_u;
    aa.seg_n0 = n0;
    nb.uprev0 = ck + BF;
    for (int m = n1; m > n0; --m) {
      nb.W_in = wl1;
      nb.W_out = wl2;
      nb.AL = w.lam + (1 + (m & 1)) * BF;
      ProfScope ps(2, s, BF);
      WT_HIP(launch_adj_nl_plain(p.unit_h, p.nl, aa, p.d_spmap, nb, fbar, m, s));
      std::swap(wl1, wl2);
    }
  }
  }
  const int64_t F = (int64_t)p.F;
  if (grad_c) WT_HIP(launch_batch_sum(w.gacc, p.B, F, grad_c, s));
  if (grad_b) WT_HIP(launch_batch_sum(w.gacc + BF, p.B, F, grad_b, s));
  if (grad_rho) WT_HIP(launch_batch_sum(w.gacc + 2 * BF, p.B, F, grad_rho, s));
  if (grad_x) {
    if (p.nslot > 0)
      WT_HIP(launch_finalize_gx(w.gsrc, p.B, T, p.nslot, p.d_mult, grad_x, s));
    else
      WT_HIP(launch_fill(grad_x, (int64_t)p.B * T, 0.0f, s));
  }
  return WT_OK;
}

}  // namespace

int wt_loop_fwd(void *plan, const float *b, const float *c, const float *rho, const float *x,
                float *y, float *fields, int32_t field_stride, int32_t with_grad, void *saved,
                void *temp, void *stream) {
  Plan *p = (Plan *)plan;
  if (!p || !b || !c || !x || !temp || (with_grad && !saved) || (p->nl && !rho) ||
      field_stride <= 0) {
    set_error("wt_loop_fwd: invalid argument");
    return WT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  Ws w = layout(*p, saved, temp, with_grad != 0);
  const int64_t BF = (int64_t)p->B * p->F;
  float *bufs[8];
  for (int i = 0; i < 8; ++i) bufs[i] = w.st + i * BF;
  // clamp-only: the clamped c goes into a state buffer before the loop uses it
  int rc = prepare_coef(*p, w, b, c, rho, bufs[7], s);
  if (rc) return rc;
  FwdArgs a = base_fwd(*p, w, b, c, rho, x);
  a.y = y;
  a.yraw = with_grad ? w.yraw : nullptr;
  a.fields = fields;
  a.fstride = field_stride;
  // state pairs: bufs[0..3] intermediate ping-pong, bufs[4..7] segment
  // input/output when no checkpoints are kept.  u_0 = u_{-1} = 0
  // (rnn.py:39-41) lives in checkpoint 0 when recording.
  float *tmp[4] = {bufs[0], bufs[1], bufs[2], bufs[3]};
  float *z = with_grad ? w.ck : bufs[4];
  WT_HIP(launch_fill(z, 2 * BF, 0.0f, s));
  const float *cur1 = z, *cur2 = z + BF;
  for (int sg = 0; sg < p->nseg; ++sg) {
    int n0 = sg * p->K, n1 = std::min(p->T, n0 + p->K);
    float *o1;
    if (with_grad && sg + 1 < p->nseg)
      o1 = w.ck + (size_t)(sg + 1) * 2 * BF;
    else
      o1 = (cur1 == bufs[4]) ? bufs[6] : bufs[4];
    if (with_grad && w.rec && sg == p->nseg - 1) {
      FwdArgs ar = a;
      ar.rec_lap = w.rec;
      ar.rec_u = p->nl ? w.rec + (size_t)p->B * p->K * p->Nx * (p->Ny + 1) : nullptr;
      ar.rec_K = p->K;
      ar.rec_n0 = n0;
      rc = run_forward(*p, ar, n0, n1, cur1, cur2, o1, o1 + BF, tmp, true, 0, s);
    } else {
      rc = run_forward(*p, a, n0, n1, cur1, cur2, o1, o1 + BF, tmp, false, 0, s);
    }
    if (rc) return rc;
    cur1 = o1;
    cur2 = o1 + BF;
  }
  return WT_OK;
}

int wt_loop_adj(void *plan, const float *b, const float *c, const float *rho, const float *x,
                const float *ybar, const float *fbar, void *saved, void *temp, float *grad_c,
                float *grad_b, float *grad_rho, float *grad_x, void *stream) {
  Plan *p = (Plan *)plan;
  if (!p || !b || !c || !x || !saved || !temp || (p->nl && !rho)) {
    set_error("wt_loop_adj: invalid argument");
    return WT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  if (p->nl)
    return adj_nonlinear(*p, b, c, rho, x, ybar, fbar, saved, temp, grad_c, grad_b, grad_rho,
                         grad_x, s);
  Ws w = layout(*p, saved, temp, true);
  const int64_t BF = (int64_t)p->B * p->F;
  const int T = p->T, K = p->K;
  bool use_tiled = p->tiled && fbar == nullptr;
  float *st[8];
  for (int i = 0; i < 8; ++i) st[i] = w.st + i * BF;
  float *lam[4];
  for (int i = 0; i < 4; ++i) lam[i] = w.lam + i * BF;
  WT_HIP(launch_fill(w.lam, 4 * BF, 0.0f, s));  // lambda_{T+1} = lambda_{T+2} = 0
  WT_HIP(launch_fill(w.wl, 2 * BF, 0.0f, s));
  WT_HIP(launch_fill(w.gacc, BF, 0.0f, s));
  WT_HIP(launch_fill(w.gsrc, (int64_t)p->B * T * std::max(p->nslot, 1), 0.0f, s));

  FwdArgs fa = base_fwd(*p, w, b, c, rho, x);
  fa.rec_lap = w.seg;
  fa.rec_K = K;
  AdjArgs aa{};
  aa.B = p->B; aa.Nx = p->Nx; aa.Ny = p->Ny; aa.T = T;
  aa.ntiles = p->ntiles_a; aa.tiles = p->d_tiles_a; aa.spx = p->d_spx; aa.taps = p->d_taps;
  aa.ntap = p->ntap; aa.nsrcslot = std::max(p->nslot, 1);
  aa.A = w.A; aa.Bm = w.Bm; aa.C2 = w.C2;
  aa.s2 = p->s2; aa.dtm1 = p->dtm1; aa.dtm2 = p->dtm2; aa.t1 = p->t1; aa.t4 = p->t4;
  aa.ybar = ybar; aa.yraw = w.yraw;
  aa.seg_lap = w.seg; aa.seg_K = K;
  aa.gacc = w.gacc; aa.gsrc = w.gsrc;

  // adjoint state: (l1, l2) = (lambda_{m+1}, lambda_{m+2})
  float *l1 = lam[0], *l2 = lam[1], *l3 = lam[2], *l4 = lam[3];
  float *wl1 = w.wl, *wl2 = w.wl + BF;
  for (int sg = p->nseg - 1; sg >= 0; --sg) {
    int n0 = sg * K, n1 = std::min(T, n0 + K);
    // recompute the segment from its checkpoint, recording lap(u_n)
    const float *ck = w.ck + (size_t)sg * 2 * BF;
    fa.rec_n0 = n0;
    float *tmp[4] = {st[0], st[1], st[2], st[3]};
    // the last segment's record came with the forward when the forward ran
    // the kernel family this sweep reads (same record layout)
    const bool from_fwd = w.rec && sg == p->nseg - 1 && use_tiled == p->tiled;
    if (!from_fwd) {
      int rc;
      if (use_tiled) {
        rc = run_forward(*p, fa, n0, n1, ck, ck + BF, nullptr, nullptr, tmp, true, 1, s);
      } else {
        Plan q = *p;
        q.tiled = false;
        rc = run_forward(q, fa, n0, n1, ck, ck + BF, nullptr, nullptr, tmp, true, 1, s);
      }
      if (rc) return rc;
    }
    aa.seg_lap = from_fwd ? w.rec : w.seg;
    aa.seg_n0 = n0;
    // adjoint over m = n1 .. n0+1
    int m = n1;
    if (use_tiled) {
      while (m > n0) {
        int k = std::min(p->tb_a, m - n0);
        aa.in1 = l1; aa.in2 = l2; aa.out1 = l3; aa.out2 = l4;
        aa.m_hi = m; aa.nsteps = k;
        ProfScope ps(2, s, (int64_t)p->B * p->Nx * p->Ny * k);
        if (p->tun.kind == 1)
          WT_HIP(launch_adj_tiled(p->tun.nt, p->tun.cpt, p->unit_h, aa, p->LWmax, s));
        else
          WT_HIP(launch_adj_cs(p->nt_a, p->cpt_a, aa, p->RHmax_a, p->RWmax_a, s));
        std::swap(l1, l3);
        std::swap(l2, l4);
        m -= k;
      }
    } else {
      for (; m > n0; --m) {
        aa.in1 = l1; aa.in2 = l2; aa.out1 = l3;
        ProfScope ps(2, s, (int64_t)p->B * p->Nx * p->Ny);
        WT_HIP(launch_adj_plain(p->unit_h, aa, p->d_spmap, wl1, wl2, fbar, m, s));
        // rotate: (l1, l2, l3) <- (l3, l1, l2)
        float *t = l2; l2 = l1; l1 = l3; l3 = t;
        std::swap(wl1, wl2);
      }
      // the tiled path uses the pair (l3, l4) as outputs; keep the plain
      // rotation consistent with it: l4 unused here
    }
  }
  if (grad_c) {
    // dJ/dc = 2 c A sum_b gacc (the clamp-only case: c is time-invariant and
    // clamped before the loop; gradient through the clamp is masked by the
    // caller)
    WT_HIP(launch_finalize_gc(w.gacc, p->B, (int64_t)p->F, c, w.A, grad_c, s));
  }
  if (grad_b) WT_HIP(launch_fill(grad_b, (int64_t)p->F, 0.0f, s));
  if (grad_rho) WT_HIP(launch_fill(grad_rho, (int64_t)p->F, 0.0f, s));
  if (grad_x) {
    if (p->nslot > 0)
      WT_HIP(launch_finalize_gx(w.gsrc, p->B, T, p->nslot, p->d_mult, grad_x, s));
    else
      WT_HIP(launch_fill(grad_x, (int64_t)p->B * T, 0.0f, s));
  }
  return WT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batches
//
// Many independent linear WaveRNN problems ("instances") of uneven grid size
// and length.  Instances are grouped by configuration (grid, PML, sources,
// probes); each group is a plan over its instances with one c per instance
// and a per-instance step count.  Every launch covers the work items of all
// groups that still have steps to run (k_fwd_cs_batch / k_adj_cs_batch), so
// small and large grids fill the GPU together and the whole forward +
// adjoint is one fixed sequence of launches (capturable in one graph).

namespace {

struct BatchPlan {
  int ngrp = 0;
  std::vector<Plan *> g;
  std::vector<int32_t *> d_Tlim;  // per group, per instance step count (device)
  std::vector<int> Tg;            // per group longest instance
  std::vector<std::vector<int>> len;  // per group, per instance step count (host)
  std::vector<char> desc;         // group's lengths non-increasing: launches cover
                                  // only the prefix of instances with steps left
  std::vector<int32_t *> d_soff;  // one segment: per instance first step slot of
                                  // the record, packed by length (device), or null
  std::vector<int> ord_f, ord_a;  // groups in a launch's table: largest regions
                                  // first (forward, adjoint tilings)
  int T = 0, K = 64, nseg = 1;
  int tb = 24, nt = 1024, R = 10, tb_a = 24, nt_a = 1024, R_a = 10;
  int RHmax = 0, RWmax = 0, RHmax_a = 0, RWmax_a = 0;
  std::vector<size_t> soff, toff;
  size_t saved_bytes = 0, temp_bytes = 0;
};

void free_batch(BatchPlan *bp) {
  if (!bp) return;
  for (Plan *p : bp->g) free_plan(p);
  for (int32_t *t : bp->d_Tlim)
    if (t) (void)hipFree(t);
  for (int32_t *t : bp->d_soff)
    if (t) (void)hipFree(t);
  delete bp;
}

// Instances of group g that a launch must cover when it runs for instances
// whose length exceeds `after` (all B unless the group's lengths are in
// non-increasing order, when those instances are a prefix), and the useful
// cell-steps of steps [lo, hi) over the group.
int batch_active(const BatchPlan &bp, int g, int after) {
  const std::vector<int> &L = bp.len[g];
  if (!bp.desc[g]) return (int)L.size();
  int n = 0;
  while (n < (int)L.size() && L[n] > after) ++n;
  return n;
}
int64_t batch_work(const BatchPlan &bp, int g, int lo, int hi) {
  int64_t w = 0;
  for (int l : bp.len[g]) w += std::max(0, std::min(l, hi) - lo);
  return w * (int64_t)bp.g[g]->F;
}

AdjArgs base_adj(const Plan &p, const Ws &w, const float *ybar) {
  AdjArgs aa{};
  aa.B = p.B; aa.Nx = p.Nx; aa.Ny = p.Ny; aa.T = p.T;
  aa.ntiles = p.ntiles_a; aa.tiles = p.d_tiles_a; aa.spx = p.d_spx; aa.taps = p.d_taps;
  aa.ntap = p.ntap; aa.nsrcslot = std::max(p.nslot, 1);
  aa.A = w.A; aa.Bm = w.Bm; aa.C2 = w.C2;
  aa.s2 = p.s2; aa.dtm1 = p.dtm1; aa.dtm2 = p.dtm2; aa.t1 = p.t1; aa.t4 = p.t4;
  aa.ybar = ybar; aa.yraw = w.yraw;
  aa.seg_lap = w.seg; aa.seg_K = p.K;
  aa.gacc = w.gacc; aa.gsrc = w.gsrc;
  return aa;
}

int batch_forward(const BatchPlan &bp, const wt_batch_group_io *io, bool with_grad, char *saved,
                  char *temp, hipStream_t s) {
  const int G = bp.ngrp;
  std::vector<Ws> W(G);
  std::vector<std::array<float *, 8>> bufs(G);
  std::vector<const float *> c1(G), c2(G), i1(G), i2(G);
  std::vector<float *> o1(G);
  for (int g = 0; g < G; ++g) {
    const Plan &p = *bp.g[g];
    const int64_t F = (int64_t)p.F, BF = (int64_t)p.B * F;
    W[g] = layout(p, saved ? saved + bp.soff[g] : nullptr, temp + bp.toff[g], with_grad);
    // A, Bm from the group's b; C2 of every instance from its own c
    WT_HIP(launch_coef_linear(io[g].b, io[g].c, F, p.dtm1, p.dtm2, W[g].A, W[g].Bm, W[g].C2, s));
    WT_HIP(launch_square(io[g].c, BF, W[g].C2, s));
    for (int i = 0; i < 8; ++i) bufs[g][i] = W[g].st + i * BF;
    float *z = with_grad ? W[g].ck : bufs[g][4];
    WT_HIP(launch_fill(z, 2 * BF, 0.0f, s));
    c1[g] = z;
    c2[g] = z + BF;
  }
  for (int sg = 0; sg < bp.nseg; ++sg) {
    const int n0 = sg * bp.K, n1 = std::min(bp.T, n0 + bp.K);
    for (int g = 0; g < G; ++g) {
      const int64_t BF = (int64_t)bp.g[g]->B * bp.g[g]->F;
      o1[g] = (with_grad && sg + 1 < bp.nseg) ? W[g].ck + (size_t)(sg + 1) * 2 * BF
                                              : (c1[g] == bufs[g][4] ? bufs[g][6] : bufs[g][4]);
      i1[g] = c1[g];
      i2[g] = c2[g];
    }
    // the last segment also records lap(u) of its steps into the saved
    // workspace (the reverse sweep starts there without a recompute)
    const bool rec = with_grad && sg + 1 == bp.nseg;
    int pp = 0;
    for (int n = n0; n < n1;) {
      const int k = std::min(bp.tb, n1 - n);
      const bool last = n + k >= n1;
      BatchFwd bl{};
      int64_t work = 0;
      std::vector<float *> out1(G), out2(G);
      for (int j = 0; j < G; ++j) {
        const int g = bp.ord_f[j];
        out1[g] = last ? o1[g] : bufs[g][2 * pp];
        out2[g] = last ? o1[g] + (int64_t)bp.g[g]->B * bp.g[g]->F : bufs[g][2 * pp + 1];
        const int nact = batch_active(bp, g, n);
        if (nact == 0) continue;
        const Plan &p = *bp.g[g];
        FwdArgs a = base_fwd(p, W[g], io[g].b, io[g].c, nullptr, io[g].x);
        a.y = io[g].y;
        a.yraw = with_grad ? W[g].yraw : nullptr;
        if (rec) {
          a.rec_lap = W[g].rec; a.rec_K = bp.K; a.rec_n0 = n0; a.rec_soff = bp.d_soff[g];
        }
        a.in1 = i1[g]; a.in2 = i2[g]; a.out1 = out1[g]; a.out2 = out2[g];
        a.n0 = n; a.nsteps = k;
        a.Tlim = bp.d_Tlim[g];
        a.c2_bstride = (int)p.F;
        bl.grp[bl.ngrp] = a;
        bl.start[bl.ngrp + 1] = bl.start[bl.ngrp] + nact * p.ntiles;
        ++bl.ngrp;
        work += batch_work(bp, g, n, n + k);
      }
      {
        ProfScope ps(rec ? 1 : 0, s, work);
        WT_HIP(launch_fwd_cs_batch(bp.nt, bp.R, rec, bl, bp.RHmax, bp.RWmax, s));
      }
      for (int g = 0; g < G; ++g) {
        i1[g] = out1[g];
        i2[g] = out2[g];
      }
      pp ^= 1;
      n += k;
    }
    for (int g = 0; g < G; ++g) {
      c1[g] = o1[g];
      c2[g] = o1[g] + (int64_t)bp.g[g]->B * bp.g[g]->F;
    }
  }
  return WT_OK;
}

int batch_adjoint(const BatchPlan &bp, const wt_batch_group_io *io, char *saved, char *temp,
                  hipStream_t s) {
  const int G = bp.ngrp;
  std::vector<Ws> W(G);
  std::vector<std::array<float *, 4>> lam(G), st(G);
  for (int g = 0; g < G; ++g) {
    const Plan &p = *bp.g[g];
    const int64_t BF = (int64_t)p.B * p.F;
    W[g] = layout(p, saved + bp.soff[g], temp + bp.toff[g], true);
    for (int i = 0; i < 4; ++i) {
      lam[g][i] = W[g].lam + i * BF;
      st[g][i] = W[g].st + i * BF;
    }
    WT_HIP(launch_fill(W[g].lam, 4 * BF, 0.0f, s));  // lambda_{T+1} = lambda_{T+2} = 0
    WT_HIP(launch_fill(W[g].gacc, BF, 0.0f, s));
    WT_HIP(launch_fill(W[g].gsrc, (int64_t)p.B * p.T * std::max(p.nslot, 1), 0.0f, s));
  }
  for (int sg = bp.nseg - 1; sg >= 0; --sg) {
    const int n0 = sg * bp.K, n1 = std::min(bp.T, n0 + bp.K);
    // recompute the segment from its checkpoint, recording lap(u_n) (the
    // last segment's record is the forward's)
    const int nrec = sg + 1 == bp.nseg ? n0 : n1;
    std::vector<const float *> i1(G), i2(G);
    for (int g = 0; g < G; ++g) {
      const int64_t BF = (int64_t)bp.g[g]->B * bp.g[g]->F;
      i1[g] = W[g].ck + (size_t)sg * 2 * BF;
      i2[g] = i1[g] + BF;
    }
    int pp = 0;
    for (int n = n0; n < nrec;) {
      const int k = std::min(bp.tb, n1 - n);
      BatchFwd bl{};
      int64_t work = 0;
      for (int j = 0; j < G; ++j) {
        const int g = bp.ord_f[j];
        const int nact = batch_active(bp, g, n);
        if (nact == 0) continue;
        const Plan &p = *bp.g[g];
        FwdArgs a = base_fwd(p, W[g], io[g].b, io[g].c, nullptr, io[g].x);
        a.rec_lap = W[g].seg; a.rec_K = bp.K; a.rec_n0 = n0;
        a.in1 = i1[g]; a.in2 = i2[g]; a.out1 = st[g][2 * pp]; a.out2 = st[g][2 * pp + 1];
        a.n0 = n; a.nsteps = k;
        a.Tlim = bp.d_Tlim[g];
        a.c2_bstride = (int)p.F;
        bl.grp[bl.ngrp] = a;
        bl.start[bl.ngrp + 1] = bl.start[bl.ngrp] + nact * p.ntiles;
        ++bl.ngrp;
        work += batch_work(bp, g, n, n + k);
      }
      {
        ProfScope ps(1, s, work);
        WT_HIP(launch_fwd_cs_batch(bp.nt, bp.R, true, bl, bp.RHmax, bp.RWmax, s));
      }
      for (int g = 0; g < G; ++g) {
        i1[g] = st[g][2 * pp];
        i2[g] = st[g][2 * pp + 1];
      }
      pp ^= 1;
      n += k;
    }
    // adjoint over m = n1 .. n0+1
    for (int m = n1; m > n0;) {
      const int k = std::min(bp.tb_a, m - n0);
      BatchAdj bl{};
      int64_t work = 0;
      for (int j = 0; j < G; ++j) {
        const int g = bp.ord_a[j];
        // reverse steps m - k + 1 .. m of instances with length >= m - k + 1
        const int nact = batch_active(bp, g, m - k);
        if (nact == 0) continue;
        const Plan &p = *bp.g[g];
        AdjArgs aa = base_adj(p, W[g], io[g].ybar);
        aa.seg_soff = bp.d_soff[g];  // one segment: the forward's packed record
        aa.in1 = lam[g][0]; aa.in2 = lam[g][1]; aa.out1 = lam[g][2]; aa.out2 = lam[g][3];
        aa.m_hi = m; aa.nsteps = k; aa.seg_n0 = n0;
        aa.Tlim = bp.d_Tlim[g];
        aa.c2_bstride = (int)p.F;
        bl.grp[bl.ngrp] = aa;
        bl.start[bl.ngrp + 1] = bl.start[bl.ngrp] + nact * p.ntiles_a;
        ++bl.ngrp;
        work += batch_work(bp, g, m - k, m);
      }
      {
        ProfScope ps(2, s, work);
        WT_HIP(launch_adj_cs_batch(bp.nt_a, bp.R_a, bl, bp.RHmax_a, bp.RWmax_a, s));
      }
      for (int g = 0; g < G; ++g) {
        std::swap(lam[g][0], lam[g][2]);
        std::swap(lam[g][1], lam[g][3]);
      }
      m -= k;
    }
  }
  for (int g = 0; g < G; ++g) {
    const Plan &p = *bp.g[g];
    if (io[g].grad_c)
      WT_HIP(launch_finalize_gc_ps(W[g].gacc, p.B, (int64_t)p.F, io[g].c, W[g].A, io[g].grad_c,
                                   s));
    if (io[g].grad_x) {
      if (p.nslot > 0)
        WT_HIP(launch_finalize_gx(W[g].gsrc, p.B, p.T, p.nslot, p.d_mult, io[g].grad_x, s));
      else
        WT_HIP(launch_fill(io[g].grad_x, (int64_t)p.B * p.T, 0.0f, s));
    }
  }
  return WT_OK;
}

}  // namespace

extern "C" {

// One batch plan over the instances given (the body of
// wt_batch_plan_create); T_force > 0 sets the plan's length T (the row stride
// of x, y, ybar and grad_x) to that of the whole batch when this plan is a
// part of it.
static int batch_build(int32_t ngroups, const wt_config *cfgs, const int32_t *B,
                       const int32_t *lengths, const float *cmax, const wt_tuning *tunings,
                       int T_force, BatchPlan **plan_out) {
  *plan_out = nullptr;
  BatchPlan *bp = new BatchPlan();
  bp->ngrp = ngroups;
  const wt_tuning *t0 = tunings;
  if (t0) {
    bp->tb = t0->tblock > 0 ? t0->tblock : bp->tb;
    bp->nt = t0->nt > 0 ? t0->nt : bp->nt;
    bp->R = t0->cpt > 0 ? t0->cpt : bp->R;
    bp->tb_a = t0->a_tblock > 0 ? t0->a_tblock : bp->tb_a;
    bp->nt_a = t0->a_nt > 0 ? t0->a_nt : bp->nt_a;
    bp->R_a = t0->a_cpt > 0 ? t0->a_cpt : bp->R_a;
    bp->K = t0->ck_every > 0 ? t0->ck_every : bp->K;
  }
  // lengths: concatenated per group
  int off = 0;
  for (int g = 0; g < ngroups; ++g) {
    if (B[g] <= 0) {
      set_error("wt_batch_plan_create: group %d has no instances", g);
      free_batch(bp);
      return WT_EINVAL;
    }
    int tg = 0;
    for (int i = 0; i < B[g]; ++i) {
      if (lengths[off + i] <= 0) {
        set_error("wt_batch_plan_create: instance %d of group %d has length %d", i, g,
                  lengths[off + i]);
        free_batch(bp);
        return WT_EINVAL;
      }
      tg = std::max(tg, (int)lengths[off + i]);
    }
    bp->Tg.push_back(tg);
    bp->len.emplace_back(lengths + off, lengths + off + B[g]);
    bp->desc.push_back(std::is_sorted(bp->len.back().rbegin(), bp->len.back().rend()));
    bp->T = std::max(bp->T, tg);
    off += B[g];
  }
  if (T_force > bp->T) bp->T = T_force;
  bp->K = std::min(bp->K, bp->T);
  bp->nseg = (bp->T + bp->K - 1) / bp->K;
  bp->tb = std::min(bp->tb, bp->K);
  bp->tb_a = std::min(bp->tb_a, bp->K);
  off = 0;
  size_t soff = 0, toff = 0;
  for (int g = 0; g < ngroups; ++g) {
    const wt_config &cf = cfgs[g];
    if (cf.sat_b0 > 0 || cf.c_nl != 0 || cf.c_clamp || cf.h != 1.0f) {
      set_error("wt_batch_plan_create: group %d: batched instances take the linear cell with "
                "h = 1", g);
      free_batch(bp);
      return WT_EUNSUPPORTED;
    }
    // the largest square tiles whose regions fit both block variants
    // (unless the group's tuning names its tiling)
    Plan *p = nullptr;
    int rc = WT_EUNSUPPORTED;
    const bool given = tunings && tunings[g].strips > 0;
    const int gmax = given ? 1 : std::max(cf.Nx, cf.Ny);
    for (int q = 1; q <= gmax && rc; ++q) {
      wt_tuning tu{};
      tu.kind = 0;
      tu.strips = given ? tunings[g].strips : q;
      tu.tcols = given ? std::max(tunings[g].tcols, 1) : q;
      tu.a_strips = given && tunings[g].a_strips > 0 ? tunings[g].a_strips : tu.strips;
      tu.a_tcols = given && tunings[g].a_tcols > 0 ? tunings[g].a_tcols : tu.tcols;
      tu.tblock = bp->tb; tu.nt = bp->nt; tu.cpt = bp->R;
      tu.a_tblock = bp->tb_a; tu.a_nt = bp->nt_a; tu.a_cpt = bp->R_a;
      tu.ck_every = bp->K;
      rc = plan_build(&cf, B[g], bp->T, cmax[g], &tu, &p);
      if (rc == WT_EINVAL || rc == WT_ECFL) break;
    }
    if (rc) {
      if (rc == WT_EUNSUPPORTED)
        set_error("wt_batch_plan_create: group %d (%dx%d) fits no tiling of the block variants "
                  "(%d threads x %d rows, %d x %d)", g, cf.Nx, cf.Ny, bp->nt, bp->R, bp->nt_a,
                  bp->R_a);
      free_batch(bp);
      return rc;
    }
    p->c2ps = true;
    // one segment: the forward records each instance's own steps only, so the
    // record packs them by length (instance i's steps start at slot rslot[i])
    std::vector<int32_t> rslot;
    if (bp->nseg == 1) {
      int64_t acc = 0;
      for (int i = 0; i < B[g]; ++i) {
        rslot.push_back((int32_t)acc);
        acc += std::min(lengths[off + i], bp->K);
      }
      p->rec_steps = acc;
    }
    plan_layout(*p);
    bp->g.push_back(p);
    if ((rc = plan_upload(p))) {
      free_batch(bp);
      return rc;
    }
    int32_t *d = nullptr;
    hipError_t e = hipMalloc((void **)&d, (size_t)B[g] * sizeof(int32_t));
    if (e == hipSuccess)
      e = hipMemcpy(d, lengths + off, (size_t)B[g] * sizeof(int32_t), hipMemcpyHostToDevice);
    bp->d_Tlim.push_back(d);
    if (e != hipSuccess) {
      rc = hip_fail(e, "wt_batch_plan_create: lengths table");
      free_batch(bp);
      return rc;
    }
    int32_t *ds = nullptr;
    if (!rslot.empty()) {
      e = hipMalloc((void **)&ds, rslot.size() * sizeof(int32_t));
      if (e == hipSuccess)
        e = hipMemcpy(ds, rslot.data(), rslot.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    }
    bp->d_soff.push_back(ds);
    if (e != hipSuccess) {
      rc = hip_fail(e, "wt_batch_plan_create: record slot table");
      free_batch(bp);
      return rc;
    }
    off += B[g];
    bp->RHmax = std::max(bp->RHmax, p->RHmax);
    bp->RWmax = std::max(bp->RWmax, p->RWmax);
    bp->RHmax_a = std::max(bp->RHmax_a, p->RHmax_a);
    bp->RWmax_a = std::max(bp->RWmax_a, p->RWmax_a);
    bp->soff.push_back(soff);
    bp->toff.push_back(toff);
    soff += align_up(p->saved_bytes);
    toff += align_up(p->temp_bytes);
  }
  bp->saved_bytes = soff;
  bp->temp_bytes = toff;
  // a launch's workgroups start in table order: the groups with the largest
  // regions (the longest-running work items) first, so that the small ones
  // fill the CUs at the launch's end
  for (int g = 0; g < ngroups; ++g) {
    bp->ord_f.push_back(g);
    bp->ord_a.push_back(g);
  }
  auto area = [bp](int g, bool adj) {
    const Plan *p = bp->g[g];
    return adj ? (int64_t)p->RHmax_a * p->RWmax_a : (int64_t)p->RHmax * p->RWmax;
  };
  std::stable_sort(bp->ord_f.begin(), bp->ord_f.end(),
                   [&](int a, int b) { return area(a, false) > area(b, false); });
  std::stable_sort(bp->ord_a.begin(), bp->ord_a.end(),
                   [&](int a, int b) { return area(a, true) > area(b, true); });
  *plan_out = bp;
  return WT_OK;
}


// The handle of wt_batch_plan_create: the batch as `parts` concurrent batch
// plans, each over a contiguous range of every group's instances (balanced by
// steps), all with the batch's length T, so that a part's arrays are the
// caller's group arrays at an offset.  With one part it runs on the caller's
// stream; with several, each part runs on a stream of the handle's own,
// forked from the caller's stream and joined back to it with events (a
// captured graph holds both).  Kernel launches of the parts overlap: one
// part's work items fill the CUs that another's launch leaves idle at its
// end (a launch of ~2400 work items keeps ~81% of the CU time busy).
struct BatchHandle {
  int ngrp = 0, T = 0;
  std::vector<BatchPlan *> parts;
  std::vector<std::vector<int>> grp;  // per part: original group of each part group
  std::vector<std::vector<int>> lo;   // per part: first instance in that group
  std::vector<int> ntap;              // per original group
  std::vector<int64_t> F;             // per original group: Nx * Ny
  std::vector<size_t> soff, toff;
  size_t saved_bytes = 0, temp_bytes = 0;
  std::vector<hipStream_t> streams;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
};

static void free_handle(BatchHandle *h) {
  if (!h) return;
  for (BatchPlan *bp : h->parts) free_batch(bp);
  for (hipStream_t s : h->streams)
    if (s) (void)hipStreamDestroy(s);
  if (h->fork) (void)hipEventDestroy(h->fork);
  for (hipEvent_t e : h->join)
    if (e) (void)hipEventDestroy(e);
  delete h;
}

// Library choice of the number of parts: one per ~512 work items of the
// first launch (two rounds of 256 workgroups), at most 4.  Measured on a
// batch of 256 instances (sides 64-256, 1000-4000 steps, ~2400 work items,
// instances longest first), graph replay of forward + adjoint: 1, 2, 3, 4,
// 5, 6, 8 parts took 125.0, 118.2, 115.7, 113.2, 115.3, 116.0, 116.4 ms;
// a 256-work-item batch on two streams was slower than on one.
static int default_parts(const BatchPlan &bp) {
  int64_t items = 0;
  for (const Plan *p : bp.g) items += (int64_t)p->B * p->ntiles;
  return (int)std::max<int64_t>(1, std::min<int64_t>(4, items / 512));
}

int wt_batch_plan_create(int32_t ngroups, const wt_config *cfgs, const int32_t *B,
                         const int32_t *lengths, const float *cmax, const wt_tuning *tunings,
                         void **plan_out) {
  if (!plan_out || !cfgs || !B || !lengths || !cmax || ngroups <= 0 ||
      ngroups > WT_MAX_GROUPS) {
    set_error("wt_batch_plan_create: invalid argument (1 <= ngroups <= %d)", WT_MAX_GROUPS);
    return WT_EINVAL;
  }
  *plan_out = nullptr;
  BatchPlan *whole = nullptr;
  int rc = batch_build(ngroups, cfgs, B, lengths, cmax, tunings, 0, &whole);
  if (rc) return rc;
  BatchHandle *h = new BatchHandle();
  h->ngrp = ngroups;
  h->T = whole->T;
  for (const Plan *p : whole->g) {
    h->ntap.push_back(p->ntap);
    h->F.push_back((int64_t)p->F);
  }
  const int nparts = (tunings && tunings[0].streams > 0) ? std::min(tunings[0].streams, 8)
                                                         : default_parts(*whole);
  if (nparts <= 1) {
    h->parts.push_back(whole);
    h->grp.emplace_back();
    h->lo.emplace_back();
    for (int g = 0; g < ngroups; ++g) {
      h->grp[0].push_back(g);
      h->lo[0].push_back(0);
    }
  } else {
    free_batch(whole);
    // group g's instances [cut[g][k], cut[g][k+1]) go to part k, cut at the
    // instance boundary nearest to k/nparts of the group's total steps
    std::vector<std::vector<int>> cut(ngroups);
    std::vector<int> goff(ngroups, 0);
    for (int g = 0; g < ngroups; ++g) {
      if (g > 0) goff[g] = goff[g - 1] + B[g - 1];
      int64_t tot = 0;
      for (int i = 0; i < B[g]; ++i) tot += lengths[goff[g] + i];
      cut[g].push_back(0);
      int64_t run = 0;
      for (int i = 0, k = 1; i < B[g] && k < nparts; ++i) {
        const int64_t r0 = run * nparts;
        run += lengths[goff[g] + i];
        const int64_t r1 = run * nparts;
        while (k < nparts && r1 >= tot * k) {
          const int at = (r1 - tot * k <= tot * k - r0) ? i + 1 : i;
          cut[g].push_back(std::max(at, cut[g].back()));
          ++k;
        }
      }
      while ((int)cut[g].size() < nparts + 1) cut[g].push_back(B[g]);
    }
    for (int k = 0; k < nparts; ++k) {
      std::vector<wt_config> pc;
      std::vector<int32_t> pB, plen;
      std::vector<float> pcm;
      std::vector<wt_tuning> ptu;
      std::vector<int> pg, plo;
      for (int g = 0; g < ngroups; ++g) {
        const int a = cut[g][k], b = cut[g][k + 1];
        if (b <= a) continue;
        pc.push_back(cfgs[g]);
        pB.push_back(b - a);
        for (int i = a; i < b; ++i) plen.push_back(lengths[goff[g] + i]);
        pcm.push_back(cmax[g]);
        if (tunings) ptu.push_back(tunings[g]);
        pg.push_back(g);
        plo.push_back(a);
      }
      if (pg.empty()) continue;
      if (tunings) {  // the block choice common to every launch comes from entry 0
        const wt_tuning &t0 = tunings[0];
        ptu[0].tblock = t0.tblock; ptu[0].nt = t0.nt; ptu[0].cpt = t0.cpt;
        ptu[0].a_tblock = t0.a_tblock; ptu[0].a_nt = t0.a_nt; ptu[0].a_cpt = t0.a_cpt;
        ptu[0].ck_every = t0.ck_every;
      }
      BatchPlan *bp = nullptr;
      rc = batch_build((int32_t)pg.size(), pc.data(), pB.data(), plen.data(), pcm.data(),
                       tunings ? ptu.data() : nullptr, h->T, &bp);
      if (rc) {
        free_handle(h);
        return rc;
      }
      h->parts.push_back(bp);
      h->grp.push_back(pg);
      h->lo.push_back(plo);
    }
  }
  for (BatchPlan *bp : h->parts) {
    h->soff.push_back(h->saved_bytes);
    h->toff.push_back(h->temp_bytes);
    h->saved_bytes += align_up(bp->saved_bytes);
    h->temp_bytes += align_up(bp->temp_bytes);
  }
  if (h->parts.size() > 1) {
    hipError_t e = hipEventCreateWithFlags(&h->fork, hipEventDisableTiming);
    for (size_t k = 0; e == hipSuccess && k < h->parts.size(); ++k) {
      hipStream_t st = nullptr;
      hipEvent_t ev = nullptr;
      e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
      if (e == hipSuccess) h->streams.push_back(st);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e == hipSuccess) h->join.push_back(ev);
    }
    if (e != hipSuccess) {
      rc = hip_fail(e, "wt_batch_plan_create: part streams");
      free_handle(h);
      return rc;
    }
  }
  *plan_out = h;
  return WT_OK;
}

int wt_batch_plan_destroy(void *plan) {
  free_handle((BatchHandle *)plan);
  return WT_OK;
}

int wt_batch_plan_tuning(const void *plan, int32_t group, wt_tuning *out) {
  const BatchHandle *h = (const BatchHandle *)plan;
  if (!h || !out || group < 0 || group >= h->ngrp) {
    set_error("wt_batch_plan_tuning: invalid argument");
    return WT_EINVAL;
  }
  for (size_t k = 0; k < h->parts.size(); ++k)
    for (size_t j = 0; j < h->grp[k].size(); ++j)
      if (h->grp[k][j] == group) {
        *out = h->parts[k]->g[j]->tun;
        out->streams = (int32_t)h->parts.size();
        return WT_OK;
      }
  set_error("wt_batch_plan_tuning: group %d has no plan", group);
  return WT_EINVAL;
}

int wt_batch_workspace_bytes(const void *plan, int64_t *saved, int64_t *temp) {
  const BatchHandle *h = (const BatchHandle *)plan;
  if (!h) {
    set_error("wt_batch_workspace_bytes: null plan");
    return WT_EINVAL;
  }
  if (saved) *saved = (int64_t)h->saved_bytes;
  if (temp) *temp = (int64_t)h->temp_bytes;
  return WT_OK;
}

}  // extern "C"

namespace {

bool handle_io_ok(const BatchHandle *h, const wt_batch_group_io *io, bool adj) {
  if (!h || !io) return false;
  for (int g = 0; g < h->ngrp; ++g) {
    if (!io[g].b || !io[g].c || !io[g].x) return false;
    if (!adj && !io[g].y) return false;
    if (adj && !io[g].ybar) return false;
  }
  return true;
}

// The group arrays of part k: the caller's at the part's first instance.
std::vector<wt_batch_group_io> part_io(const BatchHandle &h, size_t k,
                                       const wt_batch_group_io *io) {
  std::vector<wt_batch_group_io> v;
  for (size_t j = 0; j < h.grp[k].size(); ++j) {
    const int g = h.grp[k][j];
    const int64_t a = h.lo[k][j], F = h.F[g], T = h.T, P = h.ntap[g];
    wt_batch_group_io q = io[g];
    q.c = io[g].c + a * F;
    q.x = io[g].x + a * T;
    if (q.y) q.y = io[g].y + a * T * P;
    if (q.ybar) q.ybar = io[g].ybar + a * T * P;
    if (q.grad_c) q.grad_c = io[g].grad_c + a * F;
    if (q.grad_x) q.grad_x = io[g].grad_x + a * T;
    v.push_back(q);
  }
  return v;
}

// Runs fn(part plan, part io, saved, temp, stream) for every part: on the
// caller's stream for one part, else forked onto the parts' streams and
// joined back.
template <typename Fn>
int run_parts(const BatchHandle &h, const wt_batch_group_io *io, char *saved, char *temp,
              hipStream_t s, Fn fn) {
  if (h.parts.size() == 1)
    return fn(*h.parts[0], io, saved ? saved + h.soff[0] : nullptr, temp + h.toff[0], s);
  WT_HIP(hipEventRecord(h.fork, s));
  for (size_t k = 0; k < h.parts.size(); ++k) {
    WT_HIP(hipStreamWaitEvent(h.streams[k], h.fork, 0));
    const std::vector<wt_batch_group_io> pio = part_io(h, k, io);
    const int rc = fn(*h.parts[k], pio.data(), saved ? saved + h.soff[k] : nullptr,
                      temp + h.toff[k], h.streams[k]);
    if (rc) return rc;
    WT_HIP(hipEventRecord(h.join[k], h.streams[k]));
  }
  for (size_t k = 0; k < h.parts.size(); ++k) WT_HIP(hipStreamWaitEvent(s, h.join[k], 0));
  return WT_OK;
}

}  // namespace

extern "C" {

int wt_batch_fwd(void *plan, const wt_batch_group_io *io, int32_t with_grad, void *saved,
                 void *temp, void *stream) {
  const BatchHandle *h = (const BatchHandle *)plan;
  if (!handle_io_ok(h, io, false) || !temp || (with_grad && !saved)) {
    set_error("wt_batch_fwd: invalid argument");
    return WT_EINVAL;
  }
  const bool wg = with_grad != 0;
  return run_parts(*h, io, (char *)saved, (char *)temp, (hipStream_t)stream,
                   [wg](const BatchPlan &bp, const wt_batch_group_io *pio, char *sv, char *tp,
                        hipStream_t st) { return batch_forward(bp, pio, wg, sv, tp, st); });
}

int wt_batch_adj(void *plan, const wt_batch_group_io *io, void *saved, void *temp,
                 void *stream) {
  const BatchHandle *h = (const BatchHandle *)plan;
  if (!handle_io_ok(h, io, true) || !saved || !temp) {
    set_error("wt_batch_adj: invalid argument");
    return WT_EINVAL;
  }
  return run_parts(*h, io, (char *)saved, (char *)temp, (hipStream_t)stream,
                   [](const BatchPlan &bp, const wt_batch_group_io *pio, char *sv, char *tp,
                      hipStream_t st) { return batch_adjoint(bp, pio, sv, tp, st); });
}

int wt_batch_fwd_adj(void *plan, const wt_batch_group_io *io, void *saved, void *temp,
                     void *stream) {
  const BatchHandle *h = (const BatchHandle *)plan;
  if (!handle_io_ok(h, io, false) || !handle_io_ok(h, io, true) || !saved || !temp) {
    set_error("wt_batch_fwd_adj: invalid argument");
    return WT_EINVAL;
  }
  return run_parts(*h, io, (char *)saved, (char *)temp, (hipStream_t)stream,
                   [](const BatchPlan &bp, const wt_batch_group_io *pio, char *sv, char *tp,
                      hipStream_t st) {
                     const int rc = batch_forward(bp, pio, true, sv, tp, st);
                     return rc ? rc : batch_adjoint(bp, pio, sv, tp, st);
                   });
}

// ---------------------------------------------------------------- profiling

int wt_profile(int32_t enable) {
  std::lock_guard<std::mutex> g(g_prof.mu);
  g_prof.on = enable != 0;
  return WT_OK;
}

int wt_profile_read(double *ms, int64_t *launches, int64_t *cell_steps) {
  std::lock_guard<std::mutex> g(g_prof.mu);
  for (int c = 0; c < 3; ++c) {
    double tot = 0.0;
    int64_t w = 0;
    for (size_t i = 0; i < g_prof.ev[c].size(); ++i) {
      auto &pr = g_prof.ev[c][i];
      hipError_t e = hipEventSynchronize(pr.second);
      if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize(profile)");
      float t = 0.f;
      e = hipEventElapsedTime(&t, pr.first, pr.second);
      if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime(profile)");
      tot += t;
      w += g_prof.work[c][i];
      g_prof.pool.push_back(pr.first);
      g_prof.pool.push_back(pr.second);
    }
    if (ms) ms[c] = tot;
    if (launches) launches[c] = (int64_t)g_prof.ev[c].size();
    if (cell_steps) cell_steps[c] = w;
    g_prof.ev[c].clear();
    g_prof.work[c].clear();
  }
  return WT_OK;
}

// ---------------------------------------------------------------- graphs

struct Graph {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
};

int wt_graph_begin(void *stream) {
  WT_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return WT_OK;
}

int wt_graph_end(void *stream, void **graph_out) {
  if (!graph_out) {
    set_error("wt_graph_end: null output");
    return WT_EINVAL;
  }
  Graph *G = new Graph();
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &G->g);
  if (e != hipSuccess) {
    delete G;
    return hip_fail(e, "hipStreamEndCapture");
  }
  e = hipGraphInstantiate(&G->e, G->g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(G->g);
    delete G;
    return hip_fail(e, "hipGraphInstantiate");
  }
  *graph_out = G;
  return WT_OK;
}

int wt_graph_launch(void *graph, void *stream) {
  if (!graph) {
    set_error("wt_graph_launch: null graph");
    return WT_EINVAL;
  }
  WT_HIP(hipGraphLaunch(((Graph *)graph)->e, (hipStream_t)stream));
  return WT_OK;
}

int wt_graph_destroy(void *graph) {
  Graph *G = (Graph *)graph;
  if (!G) return WT_OK;
  if (G->e) (void)hipGraphExecDestroy(G->e);
  if (G->g) (void)hipGraphDestroy(G->g);
  delete G;
  return WT_OK;
}

}  // extern "C"
