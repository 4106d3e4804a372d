// Compressed embedding (quintic table) fused with the G^T R contraction,
// and its first and second adjoints (gfx950).
//
// out[i, m, k] = sum_j em[i, j, m] * P_k(em_x[i, j])                       (fwd)
// dy_dem_x[i, j] = sum_k P'_k * sum_m em[i,j,m] dy[i,m,k]
// dy_dem[i, j, m] = sum_k P_k * dy[i, m, k]                                (grad)
// dz_dy[i, m, k] = sum_j P_k dz_dem[i,j,m] + dz_dx[i,j] P'_k em[i,j,m]     (grad_grad)
// with P_k the table quintic of segment locate(em_x) plus its linear tail
// outside [lower, max), and the sorted-padding fold of the reference
// (tabulate.cc:162-447): the first neighbour whose em_x equals the row's
// last em_x and whose direction part is zero stands for the whole padded
// tail and is weighted by the number of remaining slots.
//
// Layout: one wavefront per centre atom; lane l owns output columns
// k = l, l + 64, ...  Neighbour data (segment, offset, tail delta, em) are
// loaded 64 at a time, one neighbour per lane, and broadcast to the wave
// with readlane, so the per-neighbour table row (6*M coefficients) is read
// as one contiguous, coalesced span.  The grad kernel reduces its per-lane
// partials across the wave with a butterfly.
#include <cmath>

#include "dpmi_common.h"
#include "kernels.h"
#include "tabulate_core.h"

namespace dpmi {

namespace {

using namespace tab;

template <typename T, int ND, int KQ>
__global__ void __launch_bounds__(256) k_tab_fwd(T* __restrict__ out, const T* __restrict__ table, Info<T> f,
                                                 const T* __restrict__ em_x, const T* __restrict__ em, int nloc,
                                                 int nnei, int M, int sorted) {
  const int ii = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (ii >= nloc) return;
  T acc[KQ][ND];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
    for (int m = 0; m < ND; ++m) acc[kq][m] = 0;
  fwd_block<T, ND, KQ>(acc, table, f, em_x + (size_t)ii * nnei, 1, em + (size_t)ii * nnei * ND, nnei, M, sorted != 0);
  const int lane = lane_id();
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
    if (k < M)
#pragma unroll
      for (int m = 0; m < ND; ++m) out[((size_t)ii * ND + m) * M + k] = acc[kq][m];
  }
}

template <typename T, int ND, int KQ>
__device__ __forceinline__ void grad_block(T* __restrict__ dx_out, T* __restrict__ dem_out, int dem_stride,
                                           const T (&dyr)[KQ][ND], const T* __restrict__ table, const Info<T>& f,
                                           const T* __restrict__ em_x, int xs, const T* __restrict__ em, int nnei,
                                           int M, bool sorted, bool add_x_into_dem0) {
  if (nnei <= 0) return;
  const int lane = lane_id();
  const T last = em_x[(size_t)(nnei - 1) * xs];
  bool done = false;
  for (int c0 = 0; c0 < nnei; c0 += kWave) {
    T my_x = 0;
    T my_m[ND];
#pragma unroll
    for (int m = 0; m < ND; ++m) my_m[m] = 0;
    if (!done) {
      Nbr<T, ND> r;
      const bool tail = load_nbr(r, f, em_x + (size_t)c0 * xs, xs, em + (size_t)c0 * ND, lane, nnei - c0, last,
                                 sorted);
      const unsigned long long tm = __ballot(tail);
      int jend = min(kWave, nnei - c0);
      if (tm) jend = __ffsll((long long)tm);
      for (int q = 0; q < jend; ++q) {
        const int idx = bcast(r.idx, q);
        const T x = bcast(r.x, q), delta = bcast(r.delta, q);
        T ll[ND];
#pragma unroll
        for (int m = 0; m < ND; ++m) ll[m] = bcast(r.ll[m], q);
        const T scale = (tm && q == jend - 1) ? (T)(nnei - (c0 + q)) : (T)1;
        const T* row = table + (size_t)idx * M * 6;
        T px = 0;
        T pm[ND];
#pragma unroll
        for (int m = 0; m < ND; ++m) pm[m] = 0;
#pragma unroll
        for (int kq = 0; kq < KQ; ++kq) {
          const int k = lane + kq * kWave;
          if (k < M) {
            T a[6];
            load6(row + 6 * k, a);
            T v, d;
            quintic(a, x, delta, v, d);
            T proj = 0;
#pragma unroll
            for (int m = 0; m < ND; ++m) proj += ll[m] * dyr[kq][m];
            px += d * proj;
#pragma unroll
            for (int m = 0; m < ND; ++m) pm[m] += v * dyr[kq][m];
          }
        }
        px = wave_sum(px) * scale;
#pragma unroll
        for (int m = 0; m < ND; ++m) pm[m] = wave_sum(pm[m]) * scale;
        if (lane == q) {
          my_x = px;
#pragma unroll
          for (int m = 0; m < ND; ++m) my_m[m] = pm[m];
        }
      }
      if (tm) done = true;
    }
    const int j = c0 + lane;
    if (j < nnei) {
      if (dx_out) dx_out[j] = my_x;
      if (add_x_into_dem0) my_m[0] += my_x;
#pragma unroll
      for (int m = 0; m < ND; ++m) dem_out[(size_t)j * dem_stride + m] = my_m[m];
    }
  }
}

template <typename T, int ND, int KQ>
__global__ void __launch_bounds__(256) k_tab_grad(T* __restrict__ dy_dem_x, T* __restrict__ dy_dem,
                                                  const T* __restrict__ table, Info<T> f, const T* __restrict__ em_x,
                                                  const T* __restrict__ em, const T* __restrict__ dy, int nloc,
                                                  int nnei, int M, int sorted) {
  const int ii = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (ii >= nloc) return;
  const int lane = lane_id();
  T dyr[KQ][ND];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
#pragma unroll
    for (int m = 0; m < ND; ++m) dyr[kq][m] = k < M ? dy[((size_t)ii * ND + m) * M + k] : (T)0;
  }
  grad_block<T, ND, KQ>(dy_dem_x + (size_t)ii * nnei, dy_dem + (size_t)ii * nnei * ND, ND, dyr, table, f,
                        em_x + (size_t)ii * nnei, 1, em + (size_t)ii * nnei * ND, nnei, M, sorted != 0, false);
}

template <typename T, int ND, int KQ>
__global__ void __launch_bounds__(256) k_tab_grad_grad(T* __restrict__ dz_dy, const T* __restrict__ table, Info<T> f,
                                                       const T* __restrict__ em_x, const T* __restrict__ em,
                                                       const T* __restrict__ dz_dx, const T* __restrict__ dz_dem,
                                                       int nloc, int nnei, int M, int sorted) {
  const int ii = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (ii >= nloc) return;
  const int lane = lane_id();
  T acc[KQ][ND];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
    for (int m = 0; m < ND; ++m) acc[kq][m] = 0;
  if (nnei > 0) {
    const T* ex = em_x + (size_t)ii * nnei;
    const T* e = em + (size_t)ii * nnei * ND;
    const T* zx_row = dz_dx + (size_t)ii * nnei;
    const T* zm_row = dz_dem + (size_t)ii * nnei * ND;
    const T last = ex[nnei - 1];
    for (int c0 = 0; c0 < nnei; c0 += kWave) {
      Nbr<T, ND> r;
      const bool tail = load_nbr(r, f, ex + c0, 1, e + (size_t)c0 * ND, lane, nnei - c0, last, sorted != 0);
      T zx = 0, zm[ND];
      const int j = c0 + lane;
#pragma unroll
      for (int m = 0; m < ND; ++m) zm[m] = j < nnei ? zm_row[(size_t)j * ND + m] : (T)0;
      if (j < nnei) zx = zx_row[j];
      const unsigned long long tm = __ballot(tail);
      int jend = min(kWave, nnei - c0);
      if (tm) jend = __ffsll((long long)tm);
      for (int q = 0; q < jend; ++q) {
        const int idx = bcast(r.idx, q);
        const T x = bcast(r.x, q), delta = bcast(r.delta, q), zxq = bcast(zx, q);
        T ll[ND], hh[ND];
#pragma unroll
        for (int m = 0; m < ND; ++m) {
          ll[m] = bcast(r.ll[m], q);
          hh[m] = bcast(zm[m], q);
        }
        const T scale = (tm && q == jend - 1) ? (T)(nnei - (c0 + q)) : (T)1;
        const T* row = table + (size_t)idx * M * 6;
#pragma unroll
        for (int kq = 0; kq < KQ; ++kq) {
          const int k = lane + kq * kWave;
          if (k < M) {
            T a[6];
            load6(row + 6 * k, a);
            T v, d;
            quintic(a, x, delta, v, d);
            const T zd = zxq * d;
#pragma unroll
            for (int m = 0; m < ND; ++m) acc[kq][m] += scale * (v * hh[m] + zd * ll[m]);
          }
        }
      }
      if (tm) break;
    }
  }
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
    if (k < M)
#pragma unroll
      for (int m = 0; m < ND; ++m) dz_dy[((size_t)ii * ND + m) * M + k] = acc[kq][m];
  }
}

template <typename T>
int check_nd_m(int nd, int M) {
  if (!(nd == 4 || nd == 9 || nd == 16 || nd == 25)) return 1;
  if (M <= 0 || M > 1024) return 2;
  return 0;
}


}  // namespace


template <typename T>
int launch_tab_fwd(T* out, const T* table, const T* info, const T* em_x, const T* em, int nloc, int nnei, int M,
                   int sorted, int nd, hipStream_t st) {
  if (int e = check_nd_m<T>(nd, M)) return e;
  if (nloc <= 0) return 0;
  if (nnei <= 0) {
    DPMI_CHECK(hipMemsetAsync(out, 0, sizeof(T) * (size_t)nloc * nd * M, st));
    return 0;
  }
  const Info<T> f = make_info(info);
  const dim3 grid(ceil_div((long long)nloc * kWave, 256));
  DPMI_ND_DISPATCH(nd, DPMI_KQ_DISPATCH(M, hipLaunchKernelGGL((k_tab_fwd<T, ND, KQ>), grid, dim3(256), 0, st, out,
                                                             table, f, em_x, em, nloc, nnei, M, sorted)));
  DPMI_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int launch_tab_grad(T* dy_dem_x, T* dy_dem, const T* table, const T* info, const T* em_x, const T* em, const T* dy,
                    int nloc, int nnei, int M, int sorted, int nd, hipStream_t st) {
  if (int e = check_nd_m<T>(nd, M)) return e;
  if (nloc <= 0 || nnei <= 0) return 0;
  const Info<T> f = make_info(info);
  const dim3 grid(ceil_div((long long)nloc * kWave, 256));
  DPMI_ND_DISPATCH(nd, DPMI_KQ_DISPATCH(M, hipLaunchKernelGGL((k_tab_grad<T, ND, KQ>), grid, dim3(256), 0, st,
                                                             dy_dem_x, dy_dem, table, f, em_x, em, dy, nloc, nnei,
                                                             M, sorted)));
  DPMI_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int launch_tab_grad_grad(T* dz_dy, const T* table, const T* info, const T* em_x, const T* em, const T* dz_dx,
                         const T* dz_dem, int nloc, int nnei, int M, int sorted, int nd, hipStream_t st) {
  if (int e = check_nd_m<T>(nd, M)) return e;
  if (nloc <= 0) return 0;
  if (nnei <= 0) {
    DPMI_CHECK(hipMemsetAsync(dz_dy, 0, sizeof(T) * (size_t)nloc * nd * M, st));
    return 0;
  }
  const Info<T> f = make_info(info);
  const dim3 grid(ceil_div((long long)nloc * kWave, 256));
  DPMI_ND_DISPATCH(nd, DPMI_KQ_DISPATCH(M, hipLaunchKernelGGL((k_tab_grad_grad<T, ND, KQ>), grid, dim3(256), 0, st,
                                                             dz_dy, table, f, em_x, em, dz_dx, dz_dem, nloc, nnei, M,
                                                             sorted)));
  DPMI_LAUNCH_CHECK();
  return 0;
}

template int launch_tab_fwd<double>(double*, const double*, const double*, const double*, const double*, int, int,
                                    int, int, int, hipStream_t);
template int launch_tab_fwd<float>(float*, const float*, const float*, const float*, const float*, int, int, int,
                                   int, int, hipStream_t);
template int launch_tab_grad<double>(double*, double*, const double*, const double*, const double*, const double*,
                                     const double*, int, int, int, int, int, hipStream_t);
template int launch_tab_grad<float>(float*, float*, const float*, const float*, const float*, const float*,
                                    const float*, int, int, int, int, int, hipStream_t);
template int launch_tab_grad_grad<double>(double*, const double*, const double*, const double*, const double*,
                                          const double*, const double*, int, int, int, int, int, hipStream_t);
template int launch_tab_grad_grad<float>(float*, const float*, const float*, const float*, const float*,
                                         const float*, const float*, int, int, int, int, int, hipStream_t);

}  // namespace dpmi
