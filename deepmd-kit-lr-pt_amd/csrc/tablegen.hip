// Compressed embedding table of se_e2_a, built on the device from the
// embedding net's weights (gfx950).
//
// Follows deepmd/utils/tabulate.py:67-370 (DPTabulate.build, _build_lower)
// and deepmd/utils/tabulate_math.py:353-470 (_make_data):
//   knots   xx = arange(lower, upper, stride0) ++ arange(upper, ext*upper,
//           stride1) ++ [ext*upper]  (numpy arange semantics: x_0 = start,
//           x_1 = start + step, x_i = start + i ((start + step) - start));
//   values  y, dy/dx, d2y/dx2 of the net at every knot: per layer
//           z = x W + b, y = tanh(z) (or z), dy = tanh'(z) dz,
//           d2y = tanh''(z) dz^2 + tanh'(z) d2z, times idt, plus the residual
//           x (equal widths) or [x, x] (doubled width), carried through both
//           derivatives;
//   table   per interval s of width tt (stride0 below upper, else stride1)
//           and column m: a0 = y_s, a1 = y'_s, a2 = y''_s / 2 and a3..a5 the
//           quintic Hermite terms matching y, y', y'' at the interval's far
//           knot, in the reference's expression order:
//             a3 = 1/(2 tt^3) (20 h - (8 y'_1 + 12 y'_0) tt - (3 y''_0 - y''_1) tt^2)
//             a4 = 1/(2 tt^4) (-30 h + (14 y'_1 + 16 y'_0) tt + (3 y''_0 - 2 y''_1) tt^2)
//             a5 = 1/(2 tt^5) (12 h - 6 (y'_1 + y'_0) tt + (y''_1 - y''_0) tt^2)
//           with h = y_1 - y_0; row-major [nspline][M][6].
// One workgroup per knot walks the layers with the three vectors of the
// current layer in LDS (threads over output neurons); a second kernel forms
// the coefficients with one thread per (interval, column).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "dpmi_common.h"
#include "kernels.h"

namespace dpmi {

namespace {

constexpr int kTabMaxWidth = 1024;

struct TabLayer {
  const double* w;    // [n_in, n_out]
  const double* b;    // [n_out]
  const double* idt;  // [n_out] or null
  int n_in, n_out;
};

__global__ void __launch_bounds__(256) k_tab_knots(const double* __restrict__ xx, int nknots,
                                                   const TabLayer* __restrict__ layers, int nlayers,
                                                   double* __restrict__ vv, double* __restrict__ dd,
                                                   double* __restrict__ d2, int M) {
#pragma clang fp contract(off)
  __shared__ double buf[2][3][kTabMaxWidth];
  const int k = blockIdx.x;
  if (k >= nknots) return;
  if (threadIdx.x == 0) {
    buf[0][0][0] = xx[k];
    buf[0][1][0] = 1.0;
    buf[0][2][0] = 0.0;
  }
  __syncthreads();
  int cur = 0;
  for (int l = 0; l < nlayers; ++l) {
    const TabLayer L = layers[l];
    const double* h = buf[cur][0];
    const double* dh = buf[cur][1];
    const double* d2h = buf[cur][2];
    for (int o = threadIdx.x; o < L.n_out; o += blockDim.x) {
      double z = 0.0, dz = 0.0, d2z = 0.0;
      for (int i = 0; i < L.n_in; ++i) {
        const double w = L.w[(size_t)i * L.n_out + o];
        z += h[i] * w;
        dz += dh[i] * w;
        d2z += d2h[i] * w;
      }
      z += L.b[o];
      const double y = tanh(z);
      const double g1 = 1.0 - y * y;
      const double g2 = -2.0 * y * g1;
      double yo = y, dyo = g1 * dz, d2yo = g2 * dz * dz + g1 * d2z;
      if (L.idt) {
        yo *= L.idt[o];
        dyo *= L.idt[o];
        d2yo *= L.idt[o];
      }
      if (L.n_out == L.n_in) {
        yo += h[o];
        dyo += dh[o];
        d2yo += d2h[o];
      } else if (L.n_out == 2 * L.n_in) {
        const int q = o % L.n_in;
        yo += h[q];
        dyo += dh[q];
        d2yo += d2h[q];
      }
      buf[cur ^ 1][0][o] = yo;
      buf[cur ^ 1][1][o] = dyo;
      buf[cur ^ 1][2][o] = d2yo;
    }
    __syncthreads();
    cur ^= 1;
  }
  for (int m = threadIdx.x; m < M; m += blockDim.x) {
    vv[(size_t)k * M + m] = buf[cur][0][m];
    dd[(size_t)k * M + m] = buf[cur][1][m];
    d2[(size_t)k * M + m] = buf[cur][2][m];
  }
}

__global__ void k_tab_coef(const double* __restrict__ vv, const double* __restrict__ dd,
                           const double* __restrict__ d2, int nspline, int M, int n0, double stride0,
                           double stride1, double* __restrict__ table) {
#pragma clang fp contract(off)
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nspline * M) return;
  const int s = (int)(t / M), m = (int)(t - (long long)s * M);
  const double tt = s < n0 ? stride0 : stride1;
  const size_t a = (size_t)s * M + m, b = a + M;
  const double y0 = vv[a], y1 = vv[b], g0 = dd[a], g1 = dd[b], c0 = d2[a], c1 = d2[b];
  const double hh = y1 - y0;
  double* out = table + (size_t)s * 6 * M + (size_t)m * 6;
  out[0] = y0;
  out[1] = g0;
  out[2] = 0.5 * c0;
  out[3] = (1 / (2 * tt * tt * tt)) * (20 * hh - (8 * g1 + 12 * g0) * tt - (3 * c0 - c1) * tt * tt);
  out[4] = (1 / (2 * tt * tt * tt * tt)) * (-30 * hh + (14 * g1 + 16 * g0) * tt + (3 * c0 - 2 * c1) * tt * tt);
  out[5] = (1 / (2 * tt * tt * tt * tt * tt)) * (12 * hh - 6 * (g1 + g0) * tt + (c1 - c0) * tt * tt);
}

// numpy.arange(start, stop, step) for float64 (length ceil((stop - start) /
// step); second element start + step, the rest start + i * delta with
// delta = (start + step) - start).
void np_arange(std::vector<double>& out, double start, double stop, double step) {
#pragma clang fp contract(off)
  const double len = std::ceil((stop - start) / step);
  if (!(len > 0)) return;
  const long long n = (long long)len;
  const double x1 = start + step;
  const double delta = x1 - start;
  out.push_back(start);
  if (n > 1) out.push_back(x1);
  for (long long i = 2; i < n; ++i) out.push_back(start + (double)i * delta);
}

}  // namespace

int tab_nspline(double lower, double upper, double extrapolate, double stride0, double stride1) {
#pragma clang fp contract(off)
  return (int)((upper - lower) / stride0 + (extrapolate * upper - upper) / stride1);
}

void tab_knots(std::vector<double>& xx, double lower, double upper, double extrapolate, double stride0,
               double stride1) {
#pragma clang fp contract(off)
  xx.clear();
  np_arange(xx, lower, upper, stride0);
  np_arange(xx, upper, extrapolate * upper, stride1);
  xx.push_back(extrapolate * upper);
}

void build_table_device(const TabNet& net, double lower, double upper, double extrapolate, double stride0,
                        double stride1, std::vector<double>& table, int& nspline, hipStream_t st) {
  const int nl = (int)net.width.size() - 1;
  if (nl < 1 || net.width[0] != 1) throw std::invalid_argument("the embedding net maps one input to M outputs");
  for (int w : net.width)
    if (w <= 0 || w > kTabMaxWidth)
      throw std::invalid_argument("embedding widths must be in [1, " + std::to_string(kTabMaxWidth) + "]");
  if (!(stride0 > 0) || !(stride1 > 0) || !(upper > lower) || !(extrapolate > 1))
    throw std::invalid_argument("table range: need lower < upper, extrapolate > 1 and positive strides");
  const int M = net.width[nl];
  std::vector<double> xx;
  tab_knots(xx, lower, upper, extrapolate, stride0, stride1);
  nspline = tab_nspline(lower, upper, extrapolate, stride0, stride1);
  if (nspline <= 0 || nspline + 1 > (int)xx.size()) throw std::invalid_argument("table range gives no intervals");
  const int n0 = (int)((upper - lower) / stride0);
  const int nk = nspline + 1;
  // device copies of the weights (one block per array)
  std::vector<void*> mem;
  auto up = [&](const double* h, size_t n) {
    void* d = nullptr;
    DPMI_CHECK(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(double)));
    mem.push_back(d);
    DPMI_CHECK(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, st));
    return static_cast<double*>(d);
  };
  auto done = [&]() {
    (void)hipStreamSynchronize(st);
    for (void* p : mem) (void)hipFree(p);
  };
  try {
    std::vector<TabLayer> L(nl);
    for (int l = 0; l < nl; ++l) {
      L[l].n_in = net.width[l];
      L[l].n_out = net.width[l + 1];
      L[l].w = up(net.w[l], (size_t)L[l].n_in * L[l].n_out);
      L[l].b = up(net.b[l], (size_t)L[l].n_out);
      L[l].idt = net.idt[l] ? up(net.idt[l], (size_t)L[l].n_out) : nullptr;
    }
    double* d_xx = up(xx.data(), nk);
    void* d_layers = nullptr;
    DPMI_CHECK(hipMalloc(&d_layers, sizeof(TabLayer) * nl));
    mem.push_back(d_layers);
    DPMI_CHECK(hipMemcpyAsync(d_layers, L.data(), sizeof(TabLayer) * nl, hipMemcpyHostToDevice, st));
    double* d_v = nullptr;
    DPMI_CHECK(hipMalloc((void**)&d_v, sizeof(double) * 3 * (size_t)nk * M));
    mem.push_back(d_v);
    double* d_tab = nullptr;
    DPMI_CHECK(hipMalloc((void**)&d_tab, sizeof(double) * 6 * (size_t)nspline * M));
    mem.push_back(d_tab);
    hipLaunchKernelGGL(k_tab_knots, dim3(nk), dim3(256), 0, st, d_xx, nk, static_cast<const TabLayer*>(d_layers), nl,
                       d_v, d_v + (size_t)nk * M, d_v + 2 * (size_t)nk * M, M);
    DPMI_LAUNCH_CHECK();
    const long long nt = (long long)nspline * M;
    hipLaunchKernelGGL(k_tab_coef, dim3(ceil_div(nt, 256)), dim3(256), 0, st, d_v, d_v + (size_t)nk * M,
                       d_v + 2 * (size_t)nk * M, nspline, M, n0, stride0, stride1, d_tab);
    DPMI_LAUNCH_CHECK();
    table.resize((size_t)nspline * 6 * M);
    DPMI_CHECK(hipMemcpyAsync(table.data(), d_tab, sizeof(double) * table.size(), hipMemcpyDeviceToHost, st));
    DPMI_CHECK(hipStreamSynchronize(st));
  } catch (...) {
    done();
    throw;
  }
  done();
}

}  // namespace dpmi
