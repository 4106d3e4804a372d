// Fitting-net element-wise stages of the FP64 engine (gfx950).
//
// Fitting layer (deepmd/pt/model/network/mlp.py:188-219): z = x W (GEMM),
//   y = tanh(z + b), out = y * idt + (x | [x, x] | 0).
#include "dpmi_common.h"
#include "kernels.h"

namespace dpmi {

namespace {

__global__ void k_fit_act(const double* __restrict__ z, const double* __restrict__ b, const double* __restrict__ idt,
                          const double* __restrict__ x, int act, int res_mode, long long n, int n_in, int n_out,
                          double* __restrict__ y, double* __restrict__ xn) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n * n_out) return;
  const long long r = tid / n_out;
  const int c = (int)(tid - r * n_out);
  double v = z[tid] + b[c];
  if (act) v = tanh(v);
  y[tid] = v;
  double o = idt ? v * idt[c] : v;
  if (res_mode == 1)
    o += x[r * n_in + c];
  else if (res_mode == 2)
    o += x[r * n_in + (c < n_in ? c : c - n_in)];
  xn[tid] = o;
}

__global__ void k_fit_act_bwd(const double* __restrict__ g_out, const double* __restrict__ y,
                              const double* __restrict__ idt, int act, long long n, int n_out,
                              double* __restrict__ g_z) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n * n_out) return;
  const int c = (int)(tid % n_out);
  double g = g_out[tid];
  if (idt) g *= idt[c];
  if (act) {
    const double v = y[tid];
    g *= (1. - v * v);
  }
  g_z[tid] = g;
}

__global__ void k_fit_res_bwd(const double* __restrict__ g_out, int res_mode, long long n, int n_in, int n_out,
                              double* __restrict__ g_in) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n * n_in) return;
  const long long r = tid / n_in;
  const int c = (int)(tid - r * n_in);
  double add = 0.;
  if (res_mode == 1)
    add = g_out[r * n_out + c];
  else if (res_mode == 2)
    add = g_out[r * n_out + c] + g_out[r * n_out + c + n_in];
  g_in[tid] += add;
}

__global__ void k_fit_out(const double* __restrict__ z, double b_out, double bias, const int* __restrict__ perm,
                          long long p0, long long n, double* __restrict__ atom_e) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  atom_e[perm[p0 + p]] = z[p] + b_out + bias;
}

__global__ void k_fill_rows(double* __restrict__ g, const double* __restrict__ w, long long n, int width) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n * width) return;
  g[tid] = w[tid % width];
}

__global__ void k_fill_int(int* __restrict__ p, int v, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_double(double* __restrict__ p, double v, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_float(float* __restrict__ p, float v, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_copy_double(double* __restrict__ dst, const double* __restrict__ src, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

void launch_fit_act(const double* z, const double* b, const double* idt, const double* x, int act, int res_mode,
                    long long n, int n_in, int n_out, double* y, double* x_next, hipStream_t st) {
  const long long t = n * n_out;
  if (t <= 0) return;
  hipLaunchKernelGGL(k_fit_act, dim3(ceil_div(t, 256)), dim3(256), 0, st, z, b, idt, x, act, res_mode, n, n_in, n_out,
                     y, x_next);
  DPMI_LAUNCH_CHECK();
}

void launch_fit_act_bwd(const double* g_out, const double* y, const double* idt, int act, long long n, int n_out,
                        double* g_z, hipStream_t st) {
  const long long t = n * n_out;
  if (t <= 0) return;
  hipLaunchKernelGGL(k_fit_act_bwd, dim3(ceil_div(t, 256)), dim3(256), 0, st, g_out, y, idt, act, n, n_out, g_z);
  DPMI_LAUNCH_CHECK();
}

void launch_fit_res_bwd(const double* g_out, int res_mode, long long n, int n_in, int n_out, double* g_in,
                        hipStream_t st) {
  if (res_mode == 0) return;
  const long long t = n * n_in;
  if (t <= 0) return;
  hipLaunchKernelGGL(k_fit_res_bwd, dim3(ceil_div(t, 256)), dim3(256), 0, st, g_out, res_mode, n, n_in, n_out, g_in);
  DPMI_LAUNCH_CHECK();
}

void launch_fit_out(const double* z, double b_out, double bias, const int* perm, long long p0, long long n,
                    double* atom_e, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fit_out, dim3(ceil_div(n, 256)), dim3(256), 0, st, z, b_out, bias, perm, p0, n, atom_e);
  DPMI_LAUNCH_CHECK();
}

void launch_fill_rows(double* g, const double* w, long long n, int width, hipStream_t st) {
  const long long t = n * width;
  if (t <= 0) return;
  hipLaunchKernelGGL(k_fill_rows, dim3(ceil_div(t, 256)), dim3(256), 0, st, g, w, n, width);
  DPMI_LAUNCH_CHECK();
}

void launch_fill_int(int* p, int v, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_int, dim3(ceil_div(n, 256)), dim3(256), 0, st, p, v, n);
  DPMI_LAUNCH_CHECK();
}
void launch_fill_double(double* p, double v, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_double, dim3(ceil_div(n, 256)), dim3(256), 0, st, p, v, n);
  DPMI_LAUNCH_CHECK();
}

void launch_fill_float(float* p, float v, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_float, dim3(ceil_div(n, 256)), dim3(256), 0, st, p, v, n);
  DPMI_LAUNCH_CHECK();
}

void launch_copy_double(double* dst, const double* src, long long n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy_double, dim3(ceil_div(n, 256)), dim3(256), 0, st, dst, src, n);
  DPMI_LAUNCH_CHECK();
}

}  // namespace dpmi
