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

// Device math by type: the fp32 instantiations call the fp32 functions
// (unchanged bits), the fp64 ones the double functions.
__device__ __forceinline__ float d_tanh(float x) { return tanhf(x); }
__device__ __forceinline__ double d_tanh(double x) { return tanh(x); }
__device__ __forceinline__ float d_exp(float x) { return expf(x); }
__device__ __forceinline__ double d_exp(double x) { return exp(x); }
__device__ __forceinline__ float d_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double d_sqrt(double x) { return sqrt(x); }

// out[i,j] = sum_{a,b} k[a,b] * in[i+a-r, j+b-r]   (cross-correlation, as conv2d)
// flip != 0 applies the transposed operator (kernel rotated by 180 degrees).
template <typename T>
__global__ void k_blur(const T *__restrict__ in, const T *__restrict__ k, int r, int flip,
                       int Nx, int Ny, T *__restrict__ out) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Nx * Ny) return;
  int i = q / Ny, j = q - (q / Ny) * Ny;
  int w = 2 * r + 1;
  T s = T(0);
  for (int a = 0; a < w; ++a) {
    int ii = i + a - r;
    for (int bb = 0; bb < w; ++bb) {
      int jj = j + bb - r;
      T kv = flip ? k[(w - 1 - a) * w + (w - 1 - bb)] : k[a * w + bb];
      T v = (ii >= 0 && ii < Nx && jj >= 0 && jj < Ny) ? in[ii * Ny + jj] : T(0);
      s = s + kv * v;
    }
  }
  out[q] = s;
}

// rho_proj = (T1 + tanh(beta*(x - eta))) / D;  c = c0 + dc * rho_proj
template <typename T>
__global__ void k_proj_c(const T *__restrict__ xb, int n, T T1, T D, T beta, T eta, T c0, T dc,
                         T *__restrict__ c, T *__restrict__ proj) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  T t = d_tanh(beta * (xb[q] - eta));
  T p = (T1 + t) / D;
  if (proj) proj[q] = p;
  c[q] = c0 + dc * p;
}

// g_x = grad_c * dc * beta * (1 - tanh^2) / D
template <typename T>
__global__ void k_proj_c_bwd(const T *__restrict__ gc, const T *__restrict__ xb, int n, T D,
                             T beta, T eta, T dc, T *__restrict__ gx) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  T t = d_tanh(beta * (xb[q] - eta));
  gx[q] = gc[q] * dc * beta * (T(1) - t * t) / D;
}

// ---------------------------------------------------------------- Holey
//
// WaveGeometryHoley (geom.py:109-135): the density before projection is
//   raw[i,j] = sum_h exp(-d_h / r_h),  d_h = sqrt((i - x_h)^2 + (j - y_h)^2)
// accumulated in the compute type in hole order with the reference's
// operation order ((i - x)^2 + (j - y)^2, sqrt, (-d) / r, exp).  The
// projection and c reuse k_proj_c / k_proj_c_bwd.  Hole parameters are read
// with uniform loads.
template <typename T>
__global__ void k_holey_raw(const T *__restrict__ hx, const T *__restrict__ hy,
                            const T *__restrict__ hr, int H, int Nx, int Ny,
                            T *__restrict__ raw) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Nx * Ny) return;
  const T fi = (T)(q / Ny), fj = (T)(q - (q / Ny) * Ny);
  T s = T(0);
  for (int h = 0; h < H; ++h) {
    const T dx = fi - hx[h], dy = fj - hy[h];
    const T d = d_sqrt(dx * dx + dy * dy);
    s = s + d_exp(-d / hr[h]);
  }
  raw[q] = s;
}

// Hole gradients from g = dJ/d(raw): with e = exp(-d/r),
//   dJ/dx_h = sum g e (i - x_h) / (r_h d),  dJ/dy_h = sum g e (j - y_h) / (r_h d),
//   dJ/dr_h = sum g e d / r_h^2
// (autograd through sqrt gives the same 0/0 = NaN at a centre that sits on
// a grid point).  Block (c, h) sums cells [c*per, (c+1)*per) for hole h in
// double; k_holey_grad_final adds the C partials.
constexpr int HOLEY_NT = 256;
constexpr int HOLEY_CELLS = 2048;  // cells per partial-sum workgroup

template <typename T>
__global__ __launch_bounds__(HOLEY_NT) void k_holey_grad_partial(
    const T *__restrict__ g, const T *__restrict__ hx, const T *__restrict__ hy,
    const T *__restrict__ hr, int Nx, int Ny, double *__restrict__ part) {
  __shared__ double red[3][HOLEY_NT];
  const int c = blockIdx.x, C = gridDim.x, h = blockIdx.y, tid = threadIdx.x;
  const int n = Nx * Ny;
  const int q0 = c * HOLEY_CELLS, q1 = min(n, q0 + HOLEY_CELLS);
  const T xh = hx[h], yh = hy[h], rh = hr[h];
  double ax = 0.0, ay = 0.0, ar = 0.0;
  for (int q = q0 + tid; q < q1; q += HOLEY_NT) {
    const T dx = (T)(q / Ny) - xh, dy = (T)(q - (q / Ny) * Ny) - yh;
    const T d = d_sqrt(dx * dx + dy * dy);
    const T e = d_exp(-d / rh);
    const double ge = (double)g[q] * (double)e;
    ax += ge * ((double)dx / ((double)rh * (double)d));
    ay += ge * ((double)dy / ((double)rh * (double)d));
    ar += ge * ((double)d / ((double)rh * (double)rh));
  }
  red[0][tid] = ax; red[1][tid] = ay; red[2][tid] = ar;
  __syncthreads();
  for (int w = HOLEY_NT / 2; w > 0; w >>= 1) {
    if (tid < w) {
      red[0][tid] += red[0][tid + w];
      red[1][tid] += red[1][tid + w];
      red[2][tid] += red[2][tid + w];
    }
    __syncthreads();
  }
  if (tid < 3) part[((int64_t)h * C + c) * 3 + tid] = red[tid][0];
}

template <typename T>
__global__ void k_holey_grad_final(const double *__restrict__ part, int H, int C,
                                   T *__restrict__ gx, T *__restrict__ gy, T *__restrict__ gr) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= H) return;
  double sx = 0.0, sy = 0.0, sr = 0.0;
  for (int c = 0; c < C; ++c) {
    const double *p = part + ((int64_t)h * C + c) * 3;
    sx += p[0]; sy += p[1]; sr += p[2];
  }
  if (gx) gx[h] = (T)sx;
  if (gy) gy[h] = (T)sy;
  if (gr) gr[h] = (T)sr;
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

// The projection's constant terms: np.tanh(beta*eta) and
// np.tanh(beta*eta) + np.tanh(beta*(1-eta)) in double (geom.py:221-222),
// then used as scalars of the compute type by torch.
template <typename T>
static void proj_consts(T eta, T beta, T &T1, T &D) {
  double be = (double)beta * (double)eta;
  double t1 = std::tanh(be);
  double d = t1 + std::tanh((double)beta * (1.0 - (double)eta));
  T1 = (T)t1;
  D = (T)d;
}

// c0 + (c1 - c0) * rho: the difference of the two Python floats in double,
// as a scalar of the compute type
template <typename T>
static T proj_dc(T c0, T c1) {
  return (T)((double)c1 - (double)c0);
}

template <typename T>
static int freeform_c(const char *what, const T *rho, const T *kernel, int32_t radius,
                      int32_t blur_N, int32_t Nx, int32_t Ny, T eta, T beta, T c0, T c1, T *c,
                      T *rho_proj, T *blur_out, T *work, void *stream) {
  if (!rho || !kernel || !c || !work || radius < 0 || blur_N < 0 || Nx <= 0 || Ny <= 0) {
    set_error("%s: invalid argument", what);
    return WT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  int n = Nx * Ny;
  const T *cur = rho;
  T *w0 = work, *w1 = work + n;
  for (int it = 0; it < blur_N; ++it) {
    T *dst = (it & 1) ? w1 : w0;
    hipLaunchKernelGGL(k_blur<T>, dim3(nb(n)), dim3(256), 0, s, cur, kernel, radius, 0, Nx, Ny,
                       dst);
    cur = dst;
  }
  WT_HIP(hipGetLastError());
  if (blur_out && blur_out != cur)
    WT_HIP(hipMemcpyAsync(blur_out, cur, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, s));
  T T1, D;
  proj_consts(eta, beta, T1, D);
  hipLaunchKernelGGL(k_proj_c<T>, dim3(nb(n)), dim3(256), 0, s, cur, n, T1, D, beta, eta, c0,
                     proj_dc(c0, c1), c, rho_proj);
  WT_HIP(hipGetLastError());
  return WT_OK;
}

template <typename T>
static int freeform_c_bwd(const char *what, const T *grad_c, const T *blur_in, const T *kernel,
                          int32_t radius, int32_t blur_N, int32_t Nx, int32_t Ny, T eta,
                          T beta, T c0, T c1, T *grad_rho, T *work, void *stream) {
  if (!grad_c || !blur_in || !kernel || !grad_rho || !work || radius < 0 || blur_N < 0 ||
      Nx <= 0 || Ny <= 0) {
    set_error("%s: invalid argument", what);
    return WT_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  int n = Nx * Ny;
  T T1, D;
  proj_consts(eta, beta, T1, D);
  T *w0 = work, *w1 = work + n;
  T *g = blur_N == 0 ? grad_rho : w0;
  hipLaunchKernelGGL(k_proj_c_bwd<T>, dim3(nb(n)), dim3(256), 0, s, grad_c, blur_in, n, D,
                     beta, eta, proj_dc(c0, c1), g);
  WT_HIP(hipGetLastError());
  const T *cur = g;
  for (int it = 0; it < blur_N; ++it) {
    T *dst = (it == blur_N - 1) ? grad_rho : ((cur == w0) ? w1 : w0);
    hipLaunchKernelGGL(k_blur<T>, dim3(nb(n)), dim3(256), 0, s, cur, kernel, radius, 1, Nx, Ny,
                       dst);
    cur = dst;
  }
  WT_HIP(hipGetLastError());
  return WT_OK;
}

static int holey_chunks(int Nx, int Ny) {
  return (int)(((int64_t)Nx * Ny + HOLEY_CELLS - 1) / HOLEY_CELLS);
}

static bool holey_args_ok(const char *what, int32_t H, int32_t Nx, int32_t Ny) {
  if (H < 0 || H > 65535 || Nx <= 0 || Ny <= 0 || (int64_t)Nx * Ny > (int64_t)1 << 30) {
    set_error("%s: invalid shape (holes %d, %d x %d)", what, H, Nx, Ny);
    return false;
  }
  return true;
}

// dJ/d(raw) (n values of the compute type, padded to 8 bytes), then the
// partial sums
static int holey_workspace(const char *what, int32_t nholes, int32_t Nx, int32_t Ny,
                           int64_t elem, int64_t *bytes) {
  if (!bytes) {
    set_error("%s: invalid argument (null pointer)", what);
    return WT_EINVAL;
  }
  if (!holey_args_ok(what, nholes, Nx, Ny)) return WT_EINVAL;
  const int64_t n = (int64_t)Nx * Ny;
  *bytes = (n * elem + 7) / 8 * 8 + (int64_t)nholes * holey_chunks(Nx, Ny) * 3 * 8;
  return WT_OK;
}

template <typename T>
static int holey_c(const char *what, const T *hx, const T *hy, const T *hr, int32_t nholes,
                   int32_t Nx, int32_t Ny, T eta, T beta, T c0, T c1, T *c, T *raw_out,
                   void *stream) {
  if (!c || !raw_out || (nholes > 0 && (!hx || !hy || !hr))) {
    set_error("%s: invalid argument (null pointer)", what);
    return WT_EINVAL;
  }
  if (!holey_args_ok(what, nholes, Nx, Ny)) return WT_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const int n = Nx * Ny;
  hipLaunchKernelGGL(k_holey_raw<T>, dim3(nb(n)), dim3(256), 0, s, hx, hy, hr, nholes, Nx, Ny,
                     raw_out);
  WT_HIP(hipGetLastError());
  T T1, D;
  proj_consts(eta, beta, T1, D);
  hipLaunchKernelGGL(k_proj_c<T>, dim3(nb(n)), dim3(256), 0, s, raw_out, n, T1, D, beta, eta,
                     c0, proj_dc(c0, c1), c, (T *)nullptr);
  WT_HIP(hipGetLastError());
  return WT_OK;
}

template <typename T>
static int holey_c_bwd(const char *what, const T *grad_c, const T *raw, const T *hx,
                       const T *hy, const T *hr, int32_t nholes, int32_t Nx, int32_t Ny, T eta,
                       T beta, T c0, T c1, T *grad_x, T *grad_y, T *grad_r, void *work,
                       void *stream) {
  if (!grad_c || !raw || !work || (nholes > 0 && (!hx || !hy || !hr))) {
    set_error("%s: invalid argument (null pointer)", what);
    return WT_EINVAL;
  }
  if (!holey_args_ok(what, nholes, Nx, Ny)) return WT_EINVAL;
  if (nholes == 0) return WT_OK;
  hipStream_t s = (hipStream_t)stream;
  const int n = Nx * Ny;
  T T1, D;
  proj_consts(eta, beta, T1, D);
  T *g = (T *)work;
  double *part = (double *)((char *)work + ((int64_t)n * (int64_t)sizeof(T) + 7) / 8 * 8);
  hipLaunchKernelGGL(k_proj_c_bwd<T>, dim3(nb(n)), dim3(256), 0, s, grad_c, raw, n, D, beta,
                     eta, proj_dc(c0, c1), g);
  WT_HIP(hipGetLastError());
  const int C = holey_chunks(Nx, Ny);
  hipLaunchKernelGGL(k_holey_grad_partial<T>, dim3(C, nholes), dim3(HOLEY_NT), 0, s, g, hx, hy,
                     hr, Nx, Ny, part);
  WT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_holey_grad_final<T>, dim3((nholes + 255) / 256), dim3(256), 0, s, part,
                     nholes, C, grad_x, grad_y, grad_r);
  WT_HIP(hipGetLastError());
  return WT_OK;
}

extern "C" {

int wt_geom_freeform_c(const float *rho, const float *kernel, int32_t radius, int32_t blur_N,
                       int32_t Nx, int32_t Ny, float eta, float beta, float c0, float c1,
                       float *c, float *rho_proj, float *blur_out, float *work, void *stream) {
  return freeform_c<float>("wt_geom_freeform_c", rho, kernel, radius, blur_N, Nx, Ny, eta,
                           beta, c0, c1, c, rho_proj, blur_out, work, stream);
}

int wt_geom_freeform_c_bwd(const float *grad_c, const float *blur_in, const float *kernel,
                           int32_t radius, int32_t blur_N, int32_t Nx, int32_t Ny, float eta,
                           float beta, float c0, float c1, float *grad_rho, float *work,
                           void *stream) {
  return freeform_c_bwd<float>("wt_geom_freeform_c_bwd", grad_c, blur_in, kernel, radius,
                               blur_N, Nx, Ny, eta, beta, c0, c1, grad_rho, work, stream);
}

int wt_geom_freeform_c_f64(const double *rho, const double *kernel, int32_t radius,
                           int32_t blur_N, int32_t Nx, int32_t Ny, double eta, double beta,
                           double c0, double c1, double *c, double *rho_proj, double *blur_out,
                           double *work, void *stream) {
  return freeform_c<double>("wt_geom_freeform_c_f64", rho, kernel, radius, blur_N, Nx, Ny, eta,
                            beta, c0, c1, c, rho_proj, blur_out, work, stream);
}

int wt_geom_freeform_c_bwd_f64(const double *grad_c, const double *blur_in,
                               const double *kernel, int32_t radius, int32_t blur_N, int32_t Nx,
                               int32_t Ny, double eta, double beta, double c0, double c1,
                               double *grad_rho, double *work, void *stream) {
  return freeform_c_bwd<double>("wt_geom_freeform_c_bwd_f64", grad_c, blur_in, kernel, radius,
                                blur_N, Nx, Ny, eta, beta, c0, c1, grad_rho, work, stream);
}

int wt_geom_holey_workspace_bytes(int32_t nholes, int32_t Nx, int32_t Ny, int64_t *bytes) {
  return holey_workspace("wt_geom_holey_workspace_bytes", nholes, Nx, Ny, 4, bytes);
}

int wt_geom_holey_workspace_bytes_f64(int32_t nholes, int32_t Nx, int32_t Ny, int64_t *bytes) {
  return holey_workspace("wt_geom_holey_workspace_bytes_f64", nholes, Nx, Ny, 8, bytes);
}

int wt_geom_holey_c(const float *hx, const float *hy, const float *hr, int32_t nholes,
                    int32_t Nx, int32_t Ny, float eta, float beta, float c0, float c1, float *c,
                    float *raw_out, void *stream) {
  return holey_c<float>("wt_geom_holey_c", hx, hy, hr, nholes, Nx, Ny, eta, beta, c0, c1, c,
                        raw_out, stream);
}

int wt_geom_holey_c_f64(const double *hx, const double *hy, const double *hr, int32_t nholes,
                        int32_t Nx, int32_t Ny, double eta, double beta, double c0, double c1,
                        double *c, double *raw_out, void *stream) {
  return holey_c<double>("wt_geom_holey_c_f64", hx, hy, hr, nholes, Nx, Ny, eta, beta, c0, c1,
                         c, raw_out, stream);
}

int wt_geom_holey_c_bwd(const float *grad_c, const float *raw, const float *hx, const float *hy,
                        const float *hr, int32_t nholes, int32_t Nx, int32_t Ny, float eta,
                        float beta, float c0, float c1, float *grad_x, float *grad_y,
                        float *grad_r, void *work, void *stream) {
  return holey_c_bwd<float>("wt_geom_holey_c_bwd", grad_c, raw, hx, hy, hr, nholes, Nx, Ny, eta,
                            beta, c0, c1, grad_x, grad_y, grad_r, work, stream);
}

int wt_geom_holey_c_bwd_f64(const double *grad_c, const double *raw, const double *hx,
                            const double *hy, const double *hr, int32_t nholes, int32_t Nx,
                            int32_t Ny, double eta, double beta, double c0, double c1,
                            double *grad_x, double *grad_y, double *grad_r, void *work,
                            void *stream) {
  return holey_c_bwd<double>("wt_geom_holey_c_bwd_f64", grad_c, raw, hx, hy, hr, nholes, Nx, Ny,
                             eta, beta, c0, c1, grad_x, grad_y, grad_r, work, stream);
}

}  // extern "C"
