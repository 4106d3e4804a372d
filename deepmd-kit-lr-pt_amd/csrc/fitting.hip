// Fill and copy kernels shared by the engines (gfx950).
#include "dpmi_common.h"
#include "kernels.h"

namespace dpmi {

namespace {

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
