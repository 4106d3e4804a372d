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

namespace dpmi {

namespace {

template <typename T>
struct Info {
  T lower, upper, mx, s0, s1;
  int n0, ntail;  // int((upper-lower)/s0) and the high-tail segment index
};

template <typename T>
__device__ __forceinline__ void locate(const Info<T>& f, T& x, int& idx, T& delta) {
#pragma clang fp contract(off)
  // tabulate.cc:46-73
  const T x0 = x;
  delta = (T)0;
  if (x < f.lower) {
    idx = 0;
    x = (T)0;
    delta = x0 - f.lower;
  } else if (x < f.upper) {
    idx = (int)((x - f.lower) / f.s0);
    x -= (idx * f.s0 + f.lower);
  } else if (x < f.mx) {
    // min(): a quotient rounded up to the tail count must not address the
    // row past the last segment
    idx = min(f.n0 + (int)((x - f.upper) / f.s1), f.ntail);
    x -= ((idx - f.n0) * f.s1 + f.upper);
  } else {
    idx = f.ntail;
    x = f.mx - ((idx - f.n0) * f.s1 + f.upper);
    delta = x0 - f.mx;
  }
}

__device__ __forceinline__ double bcast(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float bcast(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ int bcast(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

template <typename T>
__device__ __forceinline__ void load6(const T* p, T* a) {
#pragma unroll
  for (int q = 0; q < 6; ++q) a[q] = p[q];
}

template <typename T>
__device__ __forceinline__ void quintic(const T* a, T x, T delta, T& v, T& d) {
  // tabulate.cc:122-158
  d = a[1] + ((T)2 * a[2] + ((T)3 * a[3] + ((T)4 * a[4] + (T)5 * a[5] * x) * x) * x) * x;
  v = a[0] + (a[1] + (a[2] + (a[3] + (a[4] + a[5] * x) * x) * x) * x) * x;
  v += d * delta;
}

// Per-lane neighbour record of a 64-neighbour chunk.
template <typename T, int ND>
struct Nbr {
  T x, delta;
  int idx;
  T ll[ND];
};

template <typename T, int ND>
__device__ __forceinline__ bool load_nbr(Nbr<T, ND>& r, const Info<T>& f, const T* em_x, int xs, const T* em, int j,
                                         int nnei, T last, bool sorted) {
  if (j >= nnei) {
    r.x = 0;
    r.delta = 0;
    r.idx = 0;
#pragma unroll
    for (int m = 0; m < ND; ++m) r.ll[m] = 0;
    return false;
  }
  T x = em_x[(size_t)j * xs];
#pragma unroll
  for (int m = 0; m < ND; ++m) r.ll[m] = em[(size_t)j * ND + m];
  const bool tail = sorted && last == x && r.ll[1] == (T)0 && r.ll[2] == (T)0 && r.ll[3] == (T)0;
  locate(f, x, r.idx, r.delta);
  r.x = x;
  return tail;
}

// Accumulates one row block (nnei neighbours) into acc[KQ][ND].
template <typename T, int ND, int KQ>
__device__ __forceinline__ void fwd_block(T (&acc)[KQ][ND], const T* __restrict__ table, const Info<T>& f,
                                          const T* __restrict__ em_x, int xs, const T* __restrict__ em, int nnei,
                                          int M, bool sorted) {
  if (nnei <= 0) return;
  const int lane = lane_id();
  const T last = em_x[(size_t)(nnei - 1) * xs];
  for (int c0 = 0; c0 < nnei; c0 += kWave) {
    Nbr<T, ND> r;
    const bool tail = load_nbr(r, f, em_x + (size_t)c0 * xs, xs, em + (size_t)c0 * ND, lane, nnei - c0, last, sorted);
    const unsigned long long tm = __ballot(tail);
    int jend = min(kWave, nnei - c0);
    if (tm) jend = __ffsll((long long)tm);  // up to and including the first tail neighbour
    for (int q = 0; q < jend; ++q) {
      const int idx = bcast(r.idx, q);
      const T x = bcast(r.x, q), delta = bcast(r.delta, q);
      T ll[ND];
#pragma unroll
      for (int m = 0; m < ND; ++m) ll[m] = bcast(r.ll[m], q);
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
          const T w = v * scale;
#pragma unroll
          for (int m = 0; m < ND; ++m) acc[kq][m] += w * ll[m];
        }
      }
    }
    if (tm) break;
  }
}

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
Info<T> make_info(const T* ti) {
  Info<T> f;
  f.lower = ti[0];
  f.upper = ti[1];
  f.mx = ti[2];
  f.s0 = ti[3];
  f.s1 = ti[4];
  f.n0 = int((f.upper - f.lower) / f.s0);
  // tabulate.cc:21-30: the last segment below max
  const T edge = std::nextafter(f.mx, f.lower);
  f.ntail = f.n0 + int((edge - f.upper) / f.s1);
  return f;
}

// ------------------------------------------------------- engine variant --
// One wave per packed centre row; loops over neighbour-type blocks, each
// with the table of its (centre type, neighbour type) net; accumulates all
// blocks into xyz[g, 4, M] (se_a.py:785-842 before the /nnei).
template <typename T, int KQ>
__global__ void __launch_bounds__(256) k_tab_fwd_rows(T* __restrict__ xyz, const T* __restrict__ tables,
                                                      const long long* __restrict__ tab_off,
                                                      const Info<T>* __restrict__ infos,
                                                      const int* __restrict__ net_of, const T* __restrict__ em,
                                                      const int* __restrict__ row_type,
                                                      const int* __restrict__ sec, int ntypes, long long nrows,
                                                      int nnei, int M) {
  const long long g = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (g >= nrows) return;
  const int ti = row_type[g];
  T acc[KQ][4];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[kq][m] = 0;
  if (ti >= 0) {
    for (int tj = 0; tj < ntypes; ++tj) {
      const int net = net_of[ti * ntypes + tj];
      const int b = sec[tj], nb = sec[tj + 1] - sec[tj];
      fwd_block<T, 4, KQ>(acc, tables + tab_off[net], infos[net], em + ((size_t)g * nnei + b) * 4, 4,
                          em + ((size_t)g * nnei + b) * 4, nb, M, true);
    }
  }
  const int lane = lane_id();
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
    if (k < M)
#pragma unroll
      for (int m = 0; m < 4; ++m) xyz[((size_t)g * 4 + m) * M + k] = acc[kq][m];
  }
}

// Backward of the above into net_deriv[g, nnei, 4] = dE/d(em) with the
// em_x gradient added to component 0 (em_x is column 0 of em).
template <typename T, int KQ>
__global__ void __launch_bounds__(256) k_tab_grad_rows(T* __restrict__ net_deriv, const T* __restrict__ gxyz,
                                                       const T* __restrict__ tables,
                                                       const long long* __restrict__ tab_off,
                                                       const Info<T>* __restrict__ infos,
                                                       const int* __restrict__ net_of, const T* __restrict__ em,
                                                       const int* __restrict__ row_type,
                                                       const int* __restrict__ sec, int ntypes, long long nrows,
                                                       int nnei, int M) {
  const long long g = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (g >= nrows) return;
  const int lane = lane_id();
  const int ti = row_type[g];
  if (ti < 0) {
    for (int s = lane; s < nnei * 4; s += kWave) net_deriv[(size_t)g * nnei * 4 + s] = 0;
    return;
  }
  T dyr[KQ][4];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
#pragma unroll
    for (int m = 0; m < 4; ++m) dyr[kq][m] = k < M ? gxyz[((size_t)g * 4 + m) * M + k] : (T)0;
  }
  for (int tj = 0; tj < ntypes; ++tj) {
    const int net = net_of[ti * ntypes + tj];
    const int b = sec[tj], nb = sec[tj + 1] - sec[tj];
    grad_block<T, 4, KQ>(nullptr, net_deriv + ((size_t)g * nnei + b) * 4, 4, dyr, tables + tab_off[net],
                         infos[net], em + ((size_t)g * nnei + b) * 4, 4, em + ((size_t)g * nnei + b) * 4, nb, M,
                         true, true);
  }
}

template <typename T>
int check_nd_m(int nd, int M) {
  if (!(nd == 4 || nd == 9 || nd == 16 || nd == 25)) return 1;
  if (M <= 0 || M > 1024) return 2;
  return 0;
}

#define DPMI_KQ_DISPATCH(M, ...)                                      \
  do {                                                                \
    const int kq_ = ((M) + kWave - 1) / kWave;                         \
    if (kq_ <= 1) { constexpr int KQ = 1; __VA_ARGS__; }              \
    else if (kq_ == 2) { constexpr int KQ = 2; __VA_ARGS__; }         \
    else if (kq_ <= 4) { constexpr int KQ = 4; __VA_ARGS__; }         \
    else if (kq_ <= 8) { constexpr int KQ = 8; __VA_ARGS__; }         \
    else { constexpr int KQ = 16; __VA_ARGS__; }                      \
  } while (0)

#define DPMI_ND_DISPATCH(nd, ...)                                     \
  do {                                                                \
    if ((nd) == 4) { constexpr int ND = 4; __VA_ARGS__; }             \
    else if ((nd) == 9) { constexpr int ND = 9; __VA_ARGS__; }        \
    else if ((nd) == 16) { constexpr int ND = 16; __VA_ARGS__; }      \
    else { constexpr int ND = 25; __VA_ARGS__; }                      \
  } while (0)

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

// ------------------------------------------------------------ engine API --
void* make_table_infos(const double* infos_host, int nnets, bool fp32) {
  if (fp32) {
    auto* p = new Info<float>[nnets];
    for (int n = 0; n < nnets; ++n) {
      float t[6];
      for (int q = 0; q < 6; ++q) t[q] = (float)infos_host[n * 6 + q];
      p[n] = make_info(t);
    }
    return p;
  }
  auto* p = new Info<double>[nnets];
  for (int n = 0; n < nnets; ++n) p[n] = make_info(infos_host + n * 6);
  return p;
}
size_t table_info_bytes(bool fp32) { return fp32 ? sizeof(Info<float>) : sizeof(Info<double>); }
void free_table_infos(void* p, bool fp32) {
  if (fp32)
    delete[] static_cast<Info<float>*>(p);
  else
    delete[] static_cast<Info<double>*>(p);
}

void launch_tab_fwd_rows(const TabRowsArgs& a, hipStream_t st) {
  const dim3 grid(ceil_div(a.nrows * kWave, 256));
  DPMI_KQ_DISPATCH(a.M, hipLaunchKernelGGL((k_tab_fwd_rows<double, KQ>), grid, dim3(256), 0, st, (double*)a.xyz,
                                           (const double*)a.tables, a.tab_off, (const Info<double>*)a.infos,
                                           a.net_of, (const double*)a.em, a.row_type, a.sec, a.ntypes, a.nrows,
                                           a.nnei, a.M));
  DPMI_LAUNCH_CHECK();
}

void launch_tab_grad_rows(const TabRowsArgs& a, hipStream_t st) {
  const dim3 grid(ceil_div(a.nrows * kWave, 256));
  DPMI_KQ_DISPATCH(a.M, hipLaunchKernelGGL((k_tab_grad_rows<double, KQ>), grid, dim3(256), 0, st,
                                           (double*)a.net_deriv, (const double*)a.gxyz, (const double*)a.tables,
                                           a.tab_off, (const Info<double>*)a.infos, a.net_of, (const double*)a.em,
                                           a.row_type, a.sec, a.ntypes, a.nrows, a.nnei, a.M));
  DPMI_LAUNCH_CHECK();
}

}  // namespace dpmi
