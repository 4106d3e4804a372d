// Descriptor contraction and fitting-net element-wise stages (gfx950), FP64.
//
// Descriptor (deepmd/pt/model/descriptor/se_a.py:844-851):
//   T = xyz / nnei  [4, M];  D[m, a] = sum_c T[c, m] T[c, a]  (a < A)
// written row-permuted so that atoms of one type form a contiguous block
// for the per-type fitting GEMMs.
// Fitting layer (deepmd/pt/model/network/mlp.py:188-219): z = x W (GEMM),
//   y = tanh(z + b), out = y * idt + (x | [x, x] | 0).
#include "dpmi_common.h"
#include "kernels.h"

namespace dpmi {

namespace {

__global__ void k_descriptor(const double* __restrict__ xyz, const int* __restrict__ pos, long long nrows, int M,
                             int A, double inv_nnei, double* __restrict__ D) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nrows * M) return;
  const long long g = tid / M;
  const int m = (int)(tid - g * M);
  const double* x = xyz + (size_t)g * 4 * M;
  const double t0 = x[m] * inv_nnei, t1 = x[M + m] * inv_nnei, t2 = x[2 * M + m] * inv_nnei,
               t3 = x[3 * M + m] * inv_nnei;
  double* out = D + ((size_t)pos[g] * M + m) * A;
  for (int a = 0; a < A; ++a) {
    const double v = t0 * (x[a] * inv_nnei) + t1 * (x[M + a] * inv_nnei) + t2 * (x[2 * M + a] * inv_nnei) +
                     t3 * (x[3 * M + a] * inv_nnei);
    out[a] = v;
  }
}

// gT[c, m] = sum_a G[m, a] T[c, a] + [m < A] sum_m' G[m', m] T[c, m']
__global__ void k_descriptor_bwd(const double* __restrict__ xyz, const double* __restrict__ gD,
                                 const int* __restrict__ pos, long long nrows, int M, int A, double inv_nnei,
                                 double* __restrict__ gxyz) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nrows * M) return;
  const long long g = tid / M;
  const int m = (int)(tid - g * M);
  const double* x = xyz + (size_t)g * 4 * M;
  const double* G = gD + (size_t)pos[g] * M * A;
  double acc[4] = {0., 0., 0., 0.};
  for (int a = 0; a < A; ++a) {
    const double w = G[(size_t)m * A + a];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] += w * x[c * M + a];
  }
  if (m < A) {
    for (int mp = 0; mp < M; ++mp) {
      const double w = G[(size_t)mp * A + m];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] += w * x[c * M + mp];
    }
  }
  // d/dxyz = inv_nnei * d/dT ; T itself carries one inv_nnei factor
#pragma unroll
  for (int c = 0; c < 4; ++c) gxyz[((size_t)g * 4 + c) * M + m] = acc[c] * inv_nnei * inv_nnei;
}

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

}  // namespace

void launch_descriptor(const double* xyz, const int* pos, long long nrows, int M, int A, double inv_nnei, double* D,
                       hipStream_t st) {
  const long long n = nrows * M;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_descriptor, dim3(ceil_div(n, 256)), dim3(256), 0, st, xyz, pos, nrows, M, A, inv_nnei, D);
  DPMI_LAUNCH_CHECK();
}

void launch_descriptor_bwd(const double* xyz, const double* gD, const int* pos, long long nrows, int M, int A,
                           double inv_nnei, double* gxyz, hipStream_t st) {
  const long long n = nrows * M;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_descriptor_bwd, dim3(ceil_div(n, 256)), dim3(256), 0, st, xyz, gD, pos, nrows, M, A, inv_nnei,
                     gxyz);
  DPMI_LAUNCH_CHECK();
}

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

}  // namespace dpmi
