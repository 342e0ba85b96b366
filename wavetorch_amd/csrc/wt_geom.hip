// wt_geom.hip -- WaveGeometryFreeForm parameterisation on gfx950
// (geom.py:207-233): c = c0 + (c1-c0) * P(Blur^N(rho)), and its backward
// grad_rho = (Blur^T)^N((c1-c0) * P'(Blur^N rho) * grad_c).
//
// Blur is conv2d with a (2r+1)^2 kernel and zero padding (geom.py:213),
// accumulated over the kernel in row-major order.  P is the tanh projection
// of geom.py:217-222 with its two constant tanh terms evaluated on the host
// in double precision, as numpy evaluates them in the reference.
#include <cmath>

#include "wt_internal.h"

namespace wt {

// out[i,j] = sum_{a,b} k[a,b] * in[i+a-r, j+b-r]   (cross-correlation, as conv2d)
// flip != 0 applies the transposed operator (kernel rotated by 180 degrees).
__global__ void k_blur(const float *__restrict__ in, const float *__restrict__ k, int r,
                       int flip, int Nx, int Ny, float *__restrict__ out) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Nx * Ny) return;
  int i = q / Ny, j = q - (q / Ny) * Ny;
  int w = 2 * r + 1;
  float s = 0.0f;
  for (int a = 0; a < w; ++a) {
    int ii = i + a - r;
    for (int bb = 0; bb < w; ++bb) {
      int jj = j + bb - r;
      float kv = flip ? k[(w - 1 - a) * w + (w - 1 - bb)] : k[a * w + bb];
      float v = (ii >= 0 && ii < Nx && jj >= 0 && jj < Ny) ? in[ii * Ny + jj] : 0.0f;
      s = s + kv * v;
    }
  }
  out[q] = s;
}

// rho_proj = (T1 + tanh(beta*(x - eta))) / D;  c = c0 + dc * rho_proj
__global__ void k_proj_c(const float *__restrict__ xb, int n, float T1, float D, float beta,
                         float eta, float c0, float dc, float *__restrict__ c,
                         float *__restrict__ proj) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  float t = tanhf(beta * (xb[q] - eta));
  float p = (T1 + t) / D;
  if (proj) proj[q] = p;
  c[q] = c0 + dc * p;
}

// g_x = grad_c * dc * beta * (1 - tanh^2) / D
__global__ void k_proj_c_bwd(const float *__restrict__ gc, const float *__restrict__ xb, int n,
                             float D, float beta, float eta, float dc,
                             float *__restrict__ gx) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  float t = tanhf(beta * (xb[q] - eta));
  gx[q] = gc[q] * dc * beta * (1.0f - t * t) / D;
}

__global__ void k_mul_scalar(const float *__restrict__ in, float v, int64_t n,
                             float *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] * v;
}

__global__ void k_clamp(const float *__restrict__ in, float lo, float hi, int64_t n,
                        float *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = fminf(fmaxf(in[i], lo), hi);
}

static inline unsigned nb(int64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_mul_scalar(const float *in, float v, int64_t n, float *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mul_scalar, dim3(nb(n)), dim3(256), 0, s, in, v, n, out);
  return hipGetLastError();
}

hipError_t launch_clamp(const float *in, float lo, float hi, int64_t n, float *out,
                        hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_clamp, dim3(nb(n)), dim3(256), 0, s, in, lo, hi, n, out);
  return hipGetLastError();
}

}  // namespace wt

using namespace wt;

static void proj_consts(float eta, float beta, float &T1, float &D) {
  // np.tanh(beta*eta) and np.tanh(beta*eta) + np.tanh(beta*(1-eta)) in
  // double (geom.py:221-222), then used as fp32 scalars by torch.
  double be = (double)beta * (double)eta;
  double t1 = std::tanh(be);
  double d = t1 + std::tanh((double)beta * (1.0 - (double)eta));
  T1 = (float)t1;
  D = (float)d;
}

extern "C" int wt_geom_freeform_c(const float *rho, const float *kernel, int32_t radius,
                                  int32_t blur_N, int32_t Nx, int32_t Ny, float eta, float beta,
                                  float c0, float c1, float *c, float *rho_proj,
                                  float *blur_out, float *work, void *stream) {
  if (!rho || !kernel || !c || !work || radius < 0 || blur_N < 0 || Nx <= 0 || Ny <= 0) {
    set_error("wt_geom_freeform_c: invalid argument");
    return WT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  int n = Nx * Ny;
  const float *cur = rho;
  float *w0 = work, *w1 = work + n;
  for (int it = 0; it < blur_N; ++it) {
    float *dst = (it & 1) ? w1 : w0;
    hipLaunchKernelGGL(k_blur, dim3(nb(n)), dim3(256), 0, s, cur, kernel, radius, 0, Nx, Ny,
                       dst);
    cur = dst;
  }
  WT_HIP(hipGetLastError());
  if (blur_out && blur_out != cur)
    WT_HIP(hipMemcpyAsync(blur_out, cur, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice,
                          s));
  float T1, D;
  proj_consts(eta, beta, T1, D);
  // c0 + (c1 - c0) * rho with the Python-float difference rounded to fp32
  float dc = (float)((double)c1 - (double)c0);
  hipLaunchKernelGGL(k_proj_c, dim3(nb(n)), dim3(256), 0, s, cur, n, T1, D, beta, eta, c0, dc,
                     c, rho_proj);
  WT_HIP(hipGetLastError());
  return WT_OK;
}

extern "C" int wt_geom_freeform_c_bwd(const float *grad_c, const float *blur_in,
                                      const float *kernel, int32_t radius, int32_t blur_N,
                                      int32_t Nx, int32_t Ny, float eta, float beta, float c0,
                                      float c1, float *grad_rho, float *work, void *stream) {
  if (!grad_c || !blur_in || !kernel || !grad_rho || !work || radius < 0 || blur_N < 0 ||
      Nx <= 0 || Ny <= 0) {
    set_error("wt_geom_freeform_c_bwd: invalid argument");
    return WT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  int n = Nx * Ny;
  float T1, D;
  proj_consts(eta, beta, T1, D);
  float dc = (float)((double)c1 - (double)c0);
  float *w0 = work, *w1 = work + n;
  float *g = blur_N == 0 ? grad_rho : w0;
  hipLaunchKernelGGL(k_proj_c_bwd, dim3(nb(n)), dim3(256), 0, s, grad_c, blur_in, n, D, beta,
                     eta, dc, g);
  WT_HIP(hipGetLastError());
  const float *cur = g;
  for (int it = 0; it < blur_N; ++it) {
    float *dst = (it == blur_N - 1) ? grad_rho : ((cur == w0) ? w1 : w0);
    hipLaunchKernelGGL(k_blur, dim3(nb(n)), dim3(256), 0, s, cur, kernel, radius, 1, Nx, Ny,
                       dst);
    cur = dst;
  }
  WT_HIP(hipGetLastError());
  return WT_OK;
}
