// Table lookup helpers shared by the tabulate kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "dpmi_common.h"

namespace dpmi {
namespace tab {

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

// Coefficients a0..a5 of column `col` of table row `idx`.  Plain tables hold
// 6 values per column ([row][M][6]); packed FP32 tables (PK) hold one 16-byte
// record per column, a0 and a1 in FP32 and a2..a5 in FP16 ([row][pk_stride(M)]
// records, rows padded with zero records to a multiple of 16, so that every
// row starts on a 256-byte boundary and 16-lane groups read whole spans; the
// terms a_p t^p, p >= 2, are at most a_p stride^2 in size, so FP16 keeps them
// to FP32-level absolute accuracy of the value), read with one load.
typedef _Float16 tab_half2 __attribute__((ext_vector_type(2)));
__host__ __device__ constexpr int pk_stride(int M) { return (M + 15) / 16 * 16; }
template <typename T, bool PK>
__device__ __forceinline__ void load_coef(const T* __restrict__ table, int idx, int col, int M, T* a) {
  if constexpr (PK) {
    static_assert(sizeof(T) == 4, "packed tables are FP32");
    const float4 r = reinterpret_cast<const float4*>(table)[(size_t)idx * pk_stride(M) + col];
    const float w2 = r.z, w3 = r.w;  // scalars for the bit casts (as load_coef_pk)
    const tab_half2 h23 = __builtin_bit_cast(tab_half2, w2), h45 = __builtin_bit_cast(tab_half2, w3);
    a[0] = r.x;
    a[1] = r.y;
    a[2] = (T)h23.x;
    a[3] = (T)h23.y;
    a[4] = (T)h45.x;
    a[5] = (T)h45.y;
  } else {
    load6(table + ((size_t)idx * M + col) * 6, a);
  }
}

// Packed records through a buffer resource: the row offset idx * stride * 16 is
// wave-uniform and goes in the scalar offset, the lane's column offset
// col * 16 in the vector offset, so a record read needs no per-lane 64-bit
// address arithmetic.  The resource's record count is 2^31 - 1 bytes; a lane
// with vector offset kDeadOffset (2^31) is out of range and reads zeros
// (buffer range checking), which masks lanes without a branch.
typedef unsigned int tab_u32x4 __attribute__((ext_vector_type(4)));
constexpr int kDeadOffset = (int)0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pk_rsrc(const void* table) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(table), (short)0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ void load_coef_pk(__amdgpu_buffer_rsrc_t rs, int row_off, int col_off, float* a) {
  const tab_u32x4 r = __builtin_amdgcn_raw_buffer_load_b128(rs, col_off, row_off, 0);
  // elements copied to scalars first: clang's __builtin_bit_cast of an
  // ext_vector element expression reads from the start of the vector
  const unsigned int w0 = r.x, w1 = r.y, w2 = r.z, w3 = r.w;
  const tab_half2 h23 = __builtin_bit_cast(tab_half2, w2), h45 = __builtin_bit_cast(tab_half2, w3);
  a[0] = __uint_as_float(w0);
  a[1] = __uint_as_float(w1);
  a[2] = (float)h23.x;
  a[3] = (float)h23.y;
  a[4] = (float)h45.x;
  a[5] = (float)h45.y;
}

// Reads the coefficients of a lane's KQ columns (k = lane + 64 kq) of one
// table row.  Packed tables: lanes past the last column (k >= M) read zeros
// through an out-of-range offset.  Plain tables: such lanes read column M - 1
// (finite values; callers give those columns zero weight).
template <typename T, int KQ, bool PK>
struct RowReader {
  const T* table;
  __amdgpu_buffer_rsrc_t rs;
  int M, col[KQ];
  __device__ __forceinline__ RowReader(const T* t, int M_, int lane) : table(t), M(M_) {
    rs = pk_rsrc(t);
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) {
      const int k = lane + kq * kWave;
      col[kq] = PK ? (k < M ? k * 16 : kDeadOffset) : min(k, M - 1);
    }
  }
  __device__ __forceinline__ void load(int idx, T (&a)[KQ][6]) const {
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) {
      if constexpr (PK) load_coef_pk(rs, idx * pk_stride(M) * 16, col[kq], a[kq]);
      else load_coef<T, PK>(table, idx, col[kq], M, a[kq]);
    }
  }
};

template <typename T>
__device__ __forceinline__ void quintic(const T* a, T x, T delta, T& v, T& d) {
  // tabulate.cc:122-158, evaluated as Horner's scheme with the derivative
  // carried along (synthetic division): b_i = b_(i+1) x + a_i gives the
  // value b_0, c_i = c_(i+1) x + b_(i+1) the slope c_0 -- nine FMAs and no
  // multiplies by the constants 2..5
  T b = a[5] * x + a[4];
  T c = a[5] * x + b;
  b = b * x + a[3];
  c = c * x + b;
  b = b * x + a[2];
  c = c * x + b;
  b = b * x + a[1];
  c = c * x + b;
  d = c;
  v = b * x + a[0];
  v += d * delta;
}

// Value and derivative of the quintic separately (the forward needs the
// derivative only for the linear extension outside the table, delta != 0).
template <typename T>
__device__ __forceinline__ T quintic_value(const T* a, T x) {
  return a[0] + (a[1] + (a[2] + (a[3] + (a[4] + a[5] * x) * x) * x) * x) * x;
}
template <typename T>
__device__ __forceinline__ T quintic_slope(const T* a, T x) {
  return a[1] + ((T)2 * a[2] + ((T)3 * a[3] + ((T)4 * a[4] + (T)5 * a[5] * x) * x) * x) * x;
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

// acc += w P(x_q) ll_q for neighbour q of the chunk record r (the slope
// term only outside the table, delta != 0: a uniform branch).  em_q, when
// given, is neighbour q's em row (read with uniform, scalar loads instead of
// broadcasts from r).
template <typename T, int ND, int KQ, bool UNIT = false>
__device__ __forceinline__ void fwd_neighbour(T (&acc)[KQ][ND], const T (&a)[KQ][6], const Nbr<T, ND>& r, int q,
                                              T w, const T* __restrict__ em_q = nullptr) {
  const T x = bcast(r.x, q), delta = bcast(r.delta, q);
  T ll[ND];
  if (em_q) {
#pragma unroll
    for (int m = 0; m < ND; ++m) ll[m] = em_q[m];
  } else {
#pragma unroll
    for (int m = 0; m < ND; ++m) ll[m] = bcast(r.ll[m], q);
  }
  T v[KQ];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) v[kq] = quintic_value(a[kq], x);
  if (delta != (T)0) {
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) v[kq] += quintic_slope(a[kq], x) * delta;
  }
  if (!UNIT && w != (T)1) {  // UNIT: the caller guarantees w = 1
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) v[kq] *= w;
  }
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
    for (int m = 0; m < ND; ++m) acc[kq][m] += v[kq] * ll[m];
}

// Accumulates one row block (nnei neighbours) into acc[KQ][ND].
template <typename T, int ND, int KQ, bool PK = false>
__device__ __forceinline__ void fwd_block(T (&acc)[KQ][ND], const T* __restrict__ table, const Info<T>& f,
                                          const T* __restrict__ em_x, int xs, const T* __restrict__ em, int nnei,
                                          int M, bool sorted) {
  if (nnei <= 0) return;
  const int lane = lane_id();
  const T last = em_x[(size_t)(nnei - 1) * xs];
  const RowReader<T, KQ, PK> rows(table, M, lane);
  for (int c0 = 0; c0 < nnei; c0 += kWave) {
    Nbr<T, ND> r;
    const bool tail = load_nbr(r, f, em_x + (size_t)c0 * xs, xs, em + (size_t)c0 * ND, lane, nnei - c0, last, sorted);
    const unsigned long long tm = __ballot(tail);
    int jend = min(kWave, nnei - c0);
    if (tm) jend = __ffsll((long long)tm);  // up to and including the first tail neighbour
    // U neighbours per step: every table row of the step is requested before
    // any is used, so the row loads overlap instead of forming a load-use
    // chain per neighbour.  Full steps hold unit-weight neighbours only; the
    // remaining ones and the first tail neighbour (weight nnei - jt) follow
    // one at a time.  The order of accumulation is the neighbour order.
    constexpr int U = KQ <= 2 ? 4 : (KQ <= 4 ? 2 : 1);
    const int jmain = tm ? jend - 1 : jend;
    int q0 = 0;
    for (; q0 + U <= jmain; q0 += U) {
      T a[U][KQ][6];
#pragma unroll
      for (int u = 0; u < U; ++u) rows.load(bcast(r.idx, q0 + u), a[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) fwd_neighbour<T, ND, KQ>(acc, a[u], r, q0 + u, (T)1);
    }
    for (int q = q0; q < jend; ++q) {
      T a[KQ][6];
      rows.load(bcast(r.idx, q), a);
      fwd_neighbour<T, ND, KQ>(acc, a, r, q, (tm && q == jend - 1) ? (T)(nnei - (c0 + q)) : (T)1);
    }
    if (tm) break;
  }
}

// fwd_block reading each table row once per run of consecutive neighbours
// that share it (a block is sorted by distance, so its rows come in runs;
// 2.7 neighbours per row for water with unit statistics, 9 for fcc Cu).
// Run starts are the neighbours whose row differs from the previous one;
// each step requests the rows of U runs before using them.  Neighbours are
// accumulated in the same order and with the same weights as fwd_block.
template <typename T, int ND, int KQ, bool PK>
__device__ __forceinline__ void fwd_block_runs(T (&acc)[KQ][ND], const T* __restrict__ table, const Info<T>& f,
                                               const T* __restrict__ em_x, int xs, const T* __restrict__ em, int nnei,
                                               int M, bool sorted) {
  if (nnei <= 0) return;
  const int lane = lane_id();
  const T last = em_x[(size_t)(nnei - 1) * xs];
  const RowReader<T, KQ, PK> rows(table, M, lane);
  for (int c0 = 0; c0 < nnei; c0 += kWave) {
    Nbr<T, ND> r;
    const bool tail = load_nbr(r, f, em_x + (size_t)c0 * xs, xs, em + (size_t)c0 * ND, lane, nnei - c0, last, sorted);
    const unsigned long long tm = __ballot(tail);
    int jend = min(kWave, nnei - c0);
    if (tm) jend = __ffsll((long long)tm);  // up to and including the first tail neighbour
    const int prev = __shfl_up(r.idx, 1, kWave);
    const unsigned long long live = jend >= kWave ? ~0ull : ((1ull << jend) - 1ull);
    unsigned long long starts = __ballot(lane == 0 || r.idx != prev) & live;
    constexpr int U = KQ <= 2 ? 4 : (KQ <= 4 ? 2 : 1);
    // unit-weight neighbours [0, jmain); the first tail neighbour (weight
    // nnei - jt) follows on its own
    const int jmain = tm ? jend - 1 : jend;
    while (starts) {
      int s[U + 1];
#pragma unroll
      for (int u = 0; u <= U; ++u) {
        s[u] = starts ? __ffsll((long long)starts) - 1 : jend;
        if (u < U) starts &= starts - 1ull;
      }
      T a[U][KQ][6];
#pragma unroll
      for (int u = 0; u < U; ++u) rows.load(bcast(r.idx, min(s[u], jend - 1)), a[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int qe = min(s[u + 1], jmain);
        for (int q = s[u]; q < qe; ++q)
          fwd_neighbour<T, ND, KQ, true>(acc, a[u], r, q, (T)1, em + (size_t)(c0 + q) * ND);
      }
    }
    if (tm) {
      const int q = jend - 1;
      T a[KQ][6];
      rows.load(bcast(r.idx, q), a);
      fwd_neighbour<T, ND, KQ>(acc, a, r, q, (T)(nnei - (c0 + q)), em + (size_t)(c0 + q) * ND);
      break;
    }
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


}  // namespace tab
}  // namespace dpmi
