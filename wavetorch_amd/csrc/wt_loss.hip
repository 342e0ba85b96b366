// Loss head of the reference's training step, fused: the probe series of a
// batch -> integrated powers -> normalised powers -> cross-entropy, and the
// gradient the adjoint seeds from.
//
// Reference: train.py:61-62 (`criterion(normalize_power(model(xb).sum(dim=1)),
// yb.argmax(dim=1))` with `criterion = torch.nn.CrossEntropyLoss()`),
// utils.py:35-36 (normalize_power).  Per sample b with label l_b:
//   I[b,k] = sum_t y[b,t,k]        S_b = sum_k I[b,k]        p = I / S_b
//   loss_b = logsumexp_k p - p[l_b]              J = (1/B) sum_b loss_b
// and, with g_k = softmax(p)_k - [k == l_b],
//   dJ/dI[b,k] = (g_k - sum_j g_j p_j) / (B S_b)
// (p_j = I_j / S: dp_j/dI_k = ([j == k] - p_j) / S), which is dJ/dy[b,t,k]
// for every t.  Sums and the softmax are formed in double and rounded once.
//
// Three launches, no atomics and no memsets (the sequence can be captured in
// a HIP graph):
//  1. partial sums: B x C workgroups, workgroup (c, b) sums steps
//     [T c / C, T (c+1) / C) of sample b.  A thread of the first
//     floor(256/P)*P threads owns probe tid % P and strides through the flat
//     chunk by that count, so its loads stay on one probe and a wave's loads
//     are contiguous.  C spreads a batch over >= 512 workgroups (at most 16
//     chunks, at least 64 steps each): a sample per workgroup leaves the
//     reduction latency-bound on B CUs.
//  2. per sample: the C partials -> I, the softmax, loss_b and dJ/dI.
//  3. the batch mean.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "wt_internal.h"

namespace wt {

constexpr int LOSS_NT = 256;
constexpr int LOSS_MAX_P = 256;
constexpr int LOSS_MAX_CHUNKS = 16;

static int loss_chunks(int B, int T) {
  int c = (512 + B - 1) / B;
  c = std::min(c, LOSS_MAX_CHUNKS);
  c = std::min(c, std::max(1, T / 64));
  return std::max(c, 1);
}

__global__ __launch_bounds__(LOSS_NT) void k_power_ce_partial(const float *__restrict__ y,
                                                              int64_t T, int P,
                                                              double *__restrict__ part) {
  __shared__ double red[LOSS_NT];
  const int c = blockIdx.x, C = gridDim.x, b = blockIdx.y, tid = threadIdx.x;
  const int64_t t0 = T * c / C, t1 = T * (c + 1) / C;
  const float *ys = y + ((int64_t)b * T + t0) * P;
  const int64_t n = (t1 - t0) * P;
  const int nact = (LOSS_NT / P) * P;
  double acc = 0.0;
  if (tid < nact)
    for (int64_t i = tid; i < n; i += nact) acc += (double)ys[i];
  red[tid] = acc;
  __syncthreads();
  if (tid < P) {
    double s = 0.0;
    for (int j = tid; j < nact; j += P) s += red[j];
    part[((int64_t)b * C + c) * P + tid] = s;
  }
}

__global__ __launch_bounds__(LOSS_NT) void k_power_ce_sample(
    const double *__restrict__ part, int C, const int64_t *__restrict__ labels, int P,
    double invB, float *__restrict__ grad_I, double *__restrict__ loss_b) {
  __shared__ double pk[LOSS_MAX_P];
  const int b = blockIdx.x, tid = threadIdx.x;
  if (tid < P) {
    double s = 0.0;
    for (int c = 0; c < C; ++c) s += part[((int64_t)b * C + c) * P + tid];
    pk[tid] = s;
  }
  __syncthreads();
  if (tid != 0) return;
  const int64_t lab = labels[b];
  double S = 0.0;
  for (int k = 0; k < P; ++k) S += pk[k];
  if (lab < 0 || lab >= P) {  // no class: NaN loss and gradient for the sample
    loss_b[b] = NAN;
    for (int k = 0; k < P; ++k) grad_I[(int64_t)b * P + k] = NAN;
    return;
  }
  double m = -INFINITY;
  for (int k = 0; k < P; ++k) {
    pk[k] /= S;
    m = fmax(m, pk[k]);
  }
  double se = 0.0;
  for (int k = 0; k < P; ++k) se += exp(pk[k] - m);
  const double lse = m + log(se);
  loss_b[b] = lse - pk[lab];
  double gp = 0.0;
  for (int k = 0; k < P; ++k) gp += (exp(pk[k] - lse) - (k == lab ? 1.0 : 0.0)) * pk[k];
  const double scale = invB / S;
  for (int k = 0; k < P; ++k) {
    const double g = exp(pk[k] - lse) - (k == lab ? 1.0 : 0.0);
    grad_I[(int64_t)b * P + k] = (float)((g - gp) * scale);
  }
}

__global__ __launch_bounds__(LOSS_NT) void k_power_ce_mean(const double *__restrict__ loss_b,
                                                           int B, float *__restrict__ loss) {
  __shared__ double red[LOSS_NT];
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int i = tid; i < B; i += LOSS_NT) s += loss_b[i];
  red[tid] = s;
  __syncthreads();
  for (int w = LOSS_NT / 2; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) loss[0] = (float)(red[0] / B);
}

}  // namespace wt

using namespace wt;

static bool loss_shape_ok(const char *what, int32_t B, int32_t T, int32_t P) {
  if (B <= 0 || T <= 0 || P <= 0 || B > 65535) {
    set_error("%s: invalid shape (B %d, T %d, P %d; 1 <= B <= 65535)", what, B, T, P);
    return false;
  }
  return true;
}

extern "C" int wt_loss_power_ce_workspace_bytes(int32_t B, int32_t T, int32_t P,
                                                int64_t *bytes) {
  if (!bytes) {
    set_error("wt_loss_power_ce_workspace_bytes: invalid argument (null pointer)");
    return WT_EINVAL;
  }
  if (!loss_shape_ok("wt_loss_power_ce_workspace_bytes", B, T, P)) return WT_EINVAL;
  const int64_t C = loss_chunks(B, T);
  *bytes = ((int64_t)B * C * P + B) * (int64_t)sizeof(double);
  return WT_OK;
}

// The three launches on `work` (B*C*P + B doubles: partials, then per-sample
// losses).  Every partial and per-sample loss is written before it is read,
// so `work` needs no clearing.
static int power_ce_launch(const float *y, const int64_t *labels, int32_t B, int32_t T,
                           int32_t P, float *loss, float *grad_I, double *work,
                           hipStream_t s) {
  const int C = loss_chunks(B, T);
  double *part = work;
  double *loss_b = part + (int64_t)B * C * P;
  hipLaunchKernelGGL(k_power_ce_partial, dim3(C, B), dim3(LOSS_NT), 0, s, y, (int64_t)T, P,
                     part);
  WT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_power_ce_sample, dim3(B), dim3(LOSS_NT), 0, s, part, C, labels, P,
                     1.0 / (double)B, grad_I, loss_b);
  WT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_power_ce_mean, dim3(1), dim3(LOSS_NT), 0, s, loss_b, B, loss);
  WT_HIP(hipGetLastError());
  return WT_OK;
}

// Argument checks shared by both entry points.
static int power_ce_check(const char *what, const float *y, const int64_t *labels, int32_t B,
                          int32_t T, int32_t P, const float *loss, const float *grad_I) {
  if (!y || !labels || !loss || !grad_I) {
    set_error("%s: invalid argument (null pointer)", what);
    return WT_EINVAL;
  }
  if (!loss_shape_ok(what, B, T, P)) return WT_EINVAL;
  if (P > LOSS_MAX_P) {
    set_error("%s: %d probes (at most %d)", what, P, LOSS_MAX_P);
    return WT_EUNSUPPORTED;
  }
  return WT_OK;
}

extern "C" int wt_loss_power_ce(const float *y, const int64_t *labels, int32_t B, int32_t T,
                                int32_t P, float *loss, float *grad_I, void *work,
                                void *stream) {
  if (!work) {
    set_error("wt_loss_power_ce: invalid argument (null pointer)");
    return WT_EINVAL;
  }
  const int st = power_ce_check("wt_loss_power_ce", y, labels, B, T, P, loss, grad_I);
  if (st != WT_OK) return st;
  return power_ce_launch(y, labels, B, T, P, loss, grad_I, (double *)work, (hipStream_t)stream);
}

// ------------------------------------------------ the allocation-free variant
//
// wt_loss_power_ce_static runs the same three kernels on a fixed device array
// of this file instead of a caller-provided workspace, for callers that cannot
// allocate device memory (code under stream capture, fixed-shape serving
// loops).  Each device the library runs on holds its own copy of the array
// until the process ends.  Its device address is looked up once per device
// (hipGetSymbolAddress with that device current) and kept here; host code
// never passes the array's own name to a kernel, since in host code that name
// is a host-side stand-in, not a device address.  Calls that overlap on one
// device share the array and race; calls on one stream are ordered.
constexpr int64_t LOSS_STATIC_DOUBLES = 131072;  // 1 MiB
static __device__ __align__(256) double scratch[LOSS_STATIC_DOUBLES];

constexpr int LOSS_MAX_DEVICES = 64;
static std::atomic<double *> scratch_ptr[LOSS_MAX_DEVICES];
static std::mutex scratch_mu;

// Device address of this device's copy of `scratch`, looked up on the first
// call on the device (the current device).
static int scratch_of_current_device(const char *what, double **out) {
  int dev = 0;
  WT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= LOSS_MAX_DEVICES) {
    set_error("%s: device %d (at most %d devices)", what, dev, LOSS_MAX_DEVICES);
    return WT_EUNSUPPORTED;
  }
  double *p = scratch_ptr[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lk(scratch_mu);
    p = scratch_ptr[dev].load(std::memory_order_relaxed);
    if (!p) {
      void *addr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !addr) {
        set_error("%s: hipGetSymbolAddress of the static workspace failed on device %d: %s",
                  what, dev, hipGetErrorString(e));
        return WT_EHIP;
      }
      p = (double *)addr;
      scratch_ptr[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return WT_OK;
}

extern "C" int wt_loss_power_ce_static(const float *y, const int64_t *labels, int32_t B,
                                       int32_t T, int32_t P, float *loss, float *grad_I,
                                       void *stream) {
  const char *what = "wt_loss_power_ce_static";
  int st = power_ce_check(what, y, labels, B, T, P, loss, grad_I);
  if (st != WT_OK) return st;
  const int64_t need = (int64_t)B * loss_chunks(B, T) * P + B;
  if (need > LOSS_STATIC_DOUBLES) {
    set_error("%s: (B %d, T %d, P %d) needs %lld doubles of workspace, the static workspace "
              "holds %lld (B * chunks * P + B; use wt_loss_power_ce with a workspace)",
              what, B, T, P, (long long)need, (long long)LOSS_STATIC_DOUBLES);
    return WT_ENOMEM;
  }
  double *work = nullptr;
  st = scratch_of_current_device(what, &work);
  if (st != WT_OK) return st;
  return power_ce_launch(y, labels, B, T, P, loss, grad_I, work, (hipStream_t)stream);
}
