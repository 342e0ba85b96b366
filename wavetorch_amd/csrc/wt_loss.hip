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
// One workgroup per sample reads its T x P series with coalesced loads: a
// thread of the first floor(256/P)*P threads owns probe tid % P and strides
// through the flat series by that count, so its loads stay on one probe.
// A second one-workgroup launch takes the batch mean.  No atomics and no
// memsets: the pair can be captured in a HIP graph.
#include <cmath>

#include "wt_internal.h"

namespace wt {

constexpr int LOSS_NT = 256;
constexpr int LOSS_MAX_P = 256;

__global__ __launch_bounds__(LOSS_NT) void k_power_ce_sample(
    const float *__restrict__ y, const int64_t *__restrict__ labels, int64_t T, int P,
    double invB, float *__restrict__ grad_I, double *__restrict__ loss_b) {
  __shared__ double red[LOSS_NT];
  __shared__ double pk[LOSS_MAX_P];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int64_t n = T * P;
  const float *ys = y + (int64_t)b * n;
  const int nact = (LOSS_NT / P) * P;
  double acc = 0.0;
  if (tid < nact)
    for (int64_t i = tid; i < n; i += nact) acc += (double)ys[i];
  red[tid] = acc;
  __syncthreads();
  if (tid < P) {
    double s = 0.0;
    for (int j = tid; j < nact; j += P) s += red[j];
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

extern "C" int wt_loss_power_ce(const float *y, const int64_t *labels, int32_t B, int32_t T,
                                int32_t P, float *loss, float *grad_I, double *work,
                                void *stream) {
  if (!y || !labels || !loss || !grad_I || !work || B <= 0 || T <= 0 || P <= 0) {
    set_error("wt_loss_power_ce: invalid argument (null pointer or non-positive size)");
    return WT_EINVAL;
  }
  if (P > LOSS_MAX_P) {
    set_error("wt_loss_power_ce: %d probes (at most %d)", P, LOSS_MAX_P);
    return WT_EUNSUPPORTED;
  }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_power_ce_sample, dim3(B), dim3(LOSS_NT), 0, s, y, labels, (int64_t)T, P,
                     1.0 / (double)B, grad_I, work);
  WT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_power_ce_mean, dim3(1), dim3(LOSS_NT), 0, s, work, B, loss);
  WT_HIP(hipGetLastError());
  return WT_OK;
}
