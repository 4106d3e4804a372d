// Engine kernels of the descriptor (gfx950): compressed embedding + G^T R
// contraction + descriptor, and the fused adjoint back to net_deriv.
//
// Forward, one wavefront per packed centre row g (lanes own columns k):
//   xyz[g, c, k] = sum over neighbour blocks tj of sum_j em[g,j,c] P_k(em[g,j,0])
//   (table of net (t_g, tj); the sorted-padding fold per block)
//   T = xyz / nnei;  D[pos[g], k, a] = sum_c T[c, k] T[c, a]        (a < A)
// reference: deepmd/pt/model/descriptor/se_a.py:785-851, tabulate.cc:162-232.
//
// Backward, one wavefront per row, from G = dE/dD (row pos[g]):
//   dE/dT[c, k] = sum_a G[k, a] T[c, a] + [k < A] sum_m G[m, k] T[c, m]
//   dy = dE/dxyz = dE/dT / nnei
//   net_deriv[g, j, :] = dE/d(em[g, j, :]) through the table (tabulate.cc
//   :234-341) with the em_x gradient added to component 0.
// The adjoint keeps the forward's lane layout (lanes over columns, one
// coalesced table row per neighbour) and reduces over columns with a
// butterfly transpose over groups of neighbours.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>

#include "dpmi_common.h"
#include "kernels.h"
#include "tabulate_core.h"

namespace dpmi {

namespace {

using namespace tab;

typedef _Float16 dhalf8 __attribute__((ext_vector_type(8)));

template <typename TD>
__device__ __forceinline__ TD to_out(float v);
template <>
__device__ __forceinline__ float to_out<float>(float v) {
  return v;
}
template <>
__device__ __forceinline__ _Float16 to_out<_Float16>(float v) {
  return (_Float16)v;
}
template <typename TD, typename T>
__device__ __forceinline__ TD cvt(T v) {
  return (TD)v;
}

template <typename T, typename TD, int KQ, bool PK>
__global__ void __launch_bounds__(256) k_embed_fwd(T* __restrict__ xyz, TD* __restrict__ D, const int* __restrict__ pos,
                                                   const T* __restrict__ tables, const long long* __restrict__ tab_off,
                                                   const Info<T>* __restrict__ infos, const int* __restrict__ net_of,
                                                   const T* __restrict__ em, const int* __restrict__ row_type,
                                                   const int* __restrict__ sec, int ntypes, long long nrows, int nnei,
                                                   int M, int A, int ldD, T inv_nnei) {
  const long long g = wave_index();
  if (g >= nrows) return;
  const int lane = lane_id();
  const int ti = row_type[g];
  T acc[KQ][4];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[kq][c] = 0;
  if (ti >= 0) {
    for (int tj = 0; tj < ntypes; ++tj) {
      const int net = net_of[ti * ntypes + tj];
      const int b = sec[tj], nb = sec[tj + 1] - sec[tj];
      const T* e = em + ((size_t)g * nnei + b) * 4;
      fwd_block_runs<T, 4, KQ, PK>(acc, tables + tab_off[net], infos[net], e, 4, e, nb, M, true);
    }
  }
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
    if (k < M)
#pragma unroll
      for (int c = 0; c < 4; ++c) xyz[((size_t)g * 4 + c) * M + k] = acc[kq][c];
  }
  // descriptor row block: columns a < A live in lanes 0..A-1 (kq = 0)
  TD* out = D + (size_t)pos[g] * ldD;
  if constexpr (sizeof(TD) == 2) {
    if (A == 16) {
      // FP16 rows of 16: a lane's values of row k are contiguous, so they
      // leave as two 16-byte stores instead of sixteen 2-byte ones
      dhalf8 h[KQ][2];
#pragma unroll
      for (int a = 0; a < 16; ++a) {
        T ta[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) ta[c] = bcast(acc[0][c], a) * inv_nnei;
#pragma unroll
        for (int kq = 0; kq < KQ; ++kq) {
          T v = 0;
#pragma unroll
          for (int c = 0; c < 4; ++c) v += (acc[kq][c] * inv_nnei) * ta[c];
          h[kq][a >> 3][a & 7] = cvt<TD>(v);
        }
      }
#pragma unroll
      for (int kq = 0; kq < KQ; ++kq) {
        const int k = lane + kq * kWave;
        if (k < M) {
          *reinterpret_cast<dhalf8*>(out + (size_t)k * 16) = h[kq][0];
          *reinterpret_cast<dhalf8*>(out + (size_t)k * 16 + 8) = h[kq][1];
        }
      }
      return;
    }
  } else if constexpr (sizeof(TD) == 8) {
    if (A == 16) {
      // FP64 rows of 16 (128 contiguous bytes per lane and row k): pairs
      // of columns leave as 16-byte stores, eight instead of sixteen 8-byte
      // ones (one pair at a time keeps the broadcast values few)
      for (int a2 = 0; a2 < 8; ++a2) {
        T t0[4], t1[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          t0[c] = bcast(acc[0][c], 2 * a2) * inv_nnei;
          t1[c] = bcast(acc[0][c], 2 * a2 + 1) * inv_nnei;
        }
#pragma unroll
        for (int kq = 0; kq < KQ; ++kq) {
          const int k = lane + kq * kWave;
          T v0 = 0, v1 = 0;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            v0 += (acc[kq][c] * inv_nnei) * t0[c];
            v1 += (acc[kq][c] * inv_nnei) * t1[c];
          }
          if (k < M) *reinterpret_cast<double2*>(out + (size_t)k * 16 + 2 * a2) = make_double2(cvt<TD>(v0), cvt<TD>(v1));
        }
      }
      return;
    }
  }
  for (int a = 0; a < A; ++a) {
    T ta[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) ta[c] = bcast(acc[0][c], a) * inv_nnei;
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) {
      const int k = lane + kq * kWave;
      if (k < M) {
        T v = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) v += (acc[kq][c] * inv_nnei) * ta[c];
        out[(size_t)k * A + a] = cvt<TD>(v);
      }
    }
  }
}

// Sum over the 64 lanes of NV per-lane values w[v] (NV a power of two,
// 4..64): butterfly transpose -- at step s each lane keeps the half of its
// values whose index bit s equals its lane bit s and adds the partner's
// copy of that half -- so lane l ends with the total of value l % NV after
// log2(NV) exchanges of NV/2, NV/4, ... values instead of NV full reductions.
template <typename T, int NV>
__device__ __forceinline__ T transpose_sum(T (&w)[NV]) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; (1 << s) < NV; ++s) {
    const bool hi = (lane >> s) & 1;
#pragma unroll
    for (int i = 0; i < (NV >> (s + 1)); ++i) {
      const T keep = hi ? w[2 * i + 1] : w[2 * i];
      const T send = hi ? w[2 * i] : w[2 * i + 1];
      w[i] = keep + __shfl_xor(send, 1 << s, kWave);
    }
  }
  T r = w[0];
#pragma unroll
  for (int s = 0; (NV << s) < kWave; ++s) r += __shfl_xor(r, NV << s, kWave);
  return r;
}

// FP32 variant of transpose_sum with the exchanges across lane bits 3..0 as
// DPP row_mirror / row_half_mirror / quad_perm moves fused into the adds
// (each pairs every lane with one on the other side of that bit; no LDS
// traffic), those across bits 5 and 4 as shuffles.  Bits are taken from the
// top, so lane l ends with the total of value tvalue<NV>(l), whose bit k is
// lane bit 5 - k.
template <int NV>
__device__ __forceinline__ int tvalue(int lane) {
  int v = 0;
#pragma unroll
  for (int k = 0; (1 << k) < NV; ++k) v |= ((lane >> (5 - k)) & 1) << k;
  return v;
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int B>
__device__ __forceinline__ float dpp_partner(float v) {
  // partner across lane bit B (B < 4): 15 - l, 7 - l (within 16 / 8 lanes),
  // l ^ 2, l ^ 1
  if constexpr (B == 3) return dpp_move<0x140>(v);
  else if constexpr (B == 2) return dpp_move<0x141>(v);
  else if constexpr (B == 1) return dpp_move<0x4E>(v);
  else return dpp_move<0xB1>(v);
}
template <int B>
__device__ __forceinline__ float pair_step(float lo_val, float hi_val, bool hi) {
  // lanes with bit B clear keep the total of lo_val over the pair, the
  // others that of hi_val
  if constexpr (B >= 4) {
    const float keep = hi ? hi_val : lo_val, send = hi ? lo_val : hi_val;
    return keep + __shfl_xor(send, 1 << B, kWave);
  } else {
    const float t0 = lo_val + dpp_partner<B>(lo_val), t1 = hi_val + dpp_partner<B>(hi_val);
    return hi ? t1 : t0;
  }
}
template <int NV, int S>
__device__ __forceinline__ void transpose_steps(float (&w)[NV], int lane) {
  if constexpr ((1 << S) < NV) {
    constexpr int B = 5 - S;
    const bool hi = (lane >> B) & 1;
#pragma unroll
    for (int i = 0; i < (NV >> (S + 1)); ++i) w[i] = pair_step<B>(w[2 * i], w[2 * i + 1], hi);
    transpose_steps<NV, S + 1>(w, lane);
  }
}
template <int NV>
__device__ __forceinline__ float transpose_sum_f32(float (&w)[NV]) {
  static_assert(NV >= 4 && NV <= 64 && (NV & (NV - 1)) == 0, "4 <= NV <= 64, a power of two");
  transpose_steps<NV, 0>(w, lane_id());
  // lane bits below the transposed ones (bits < 6 - log2 NV): butterflies
  float r = w[0];
  if constexpr (NV <= 32) r += dpp_partner<0>(r);
  if constexpr (NV <= 16) r += dpp_partner<1>(r);
  if constexpr (NV <= 8) r += dpp_partner<2>(r);
  if constexpr (NV <= 4) r += dpp_partner<3>(r);
  return r;
}

// Transposed sum with its lane layout: value_of(lane) is the value a lane
// ends with, owner(lane) picks one lane per value, lane_of(v) a lane that
// holds value v.
template <typename T, int NV>
struct TSum {
  static __device__ __forceinline__ T sum(T (&w)[NV]) { return transpose_sum<T, NV>(w); }
  static __device__ __forceinline__ int value_of(int lane) { return lane % NV; }
  static __device__ __forceinline__ bool owner(int lane) { return lane < NV; }
  static __device__ __forceinline__ int lane_of(int v) { return v; }
};
template <int NV>
struct TSum<float, NV> {
  static __device__ __forceinline__ float sum(float (&w)[NV]) { return transpose_sum_f32<NV>(w); }
  static __device__ __forceinline__ int value_of(int lane) { return tvalue<NV>(lane); }
  static __device__ __forceinline__ bool owner(int lane) { return (lane & (kWave / NV - 1)) == 0; }
  static __device__ __forceinline__ int lane_of(int v) {
    int l = 0;
#pragma unroll
    for (int k = 0; (1 << k) < NV; ++k) l |= ((v >> k) & 1) << (5 - k);
    return l;
  }
};

// Adjoint of one neighbour block (nb slots) of row g, lanes over columns k
// (as the forward, so each neighbour's table row is one coalesced read):
//   nd[j, c] = sum_k dy[k, c] P_k(x_j) + [c = 0] sum_k P'_k(x_j) sum_c' dy[k, c'] em[j, c']
// per-lane partial sums of QG neighbours x 4 components are reduced across
// the wave with transpose_sum, which leaves neighbour (l / 4) component
// (l % 4) on lane l for one coalesced store.  Sorted-padding tail as the
// forward: the first tail slot carries weight (nb - jt), later slots zero.
template <typename T, int KQ, bool PK>
__device__ void embed_bwd_block(T* __restrict__ nd, const T (&dy)[KQ][4], const T* __restrict__ table,
                                const Info<T>& f, const T* __restrict__ em, int nb, int M) {
  if (nb <= 0) return;
  constexpr int QG = 8;  // neighbours per reduction group
  const int lane = lane_id();
  const T last = em[(size_t)(nb - 1) * 4];
  int jt = nb;  // first neighbour of the sorted zero-padding tail (tabulate.cc:188-201)
  for (int c0 = 0; c0 < nb; c0 += kWave) {
    const int j = c0 + lane;
    bool tail = false;
    if (j < nb) {
      const T* r = em + (size_t)j * 4;
      tail = r[0] == last && r[1] == (T)0 && r[2] == (T)0 && r[3] == (T)0;
    }
    const unsigned long long m = __ballot(tail);
    if (m) {
      jt = c0 + __ffsll((long long)m) - 1;
      break;
    }
  }
  const int jend = jt < nb ? jt + 1 : nb;
  const RowReader<T, KQ, PK> rows(table, M, lane);
  for (int c0 = 0; c0 < jend; c0 += kWave) {
    Nbr<T, 4> r;
    load_nbr(r, f, em + (size_t)c0 * 4, 4, em + (size_t)c0 * 4, lane, jend - c0, last, false);
    const int cnt = min(kWave, jend - c0);
    for (int q0 = 0; q0 < cnt; q0 += QG) {
      T w[QG * 4];
      // rows of U neighbours requested before use (clamped, zero-weighted
      // beyond cnt), as in the forward
      constexpr int U = KQ <= 2 ? 4 : (KQ <= 4 ? 2 : 1);
#pragma unroll
      for (int u0 = 0; u0 < QG; u0 += U) {
        T a[U][KQ][6];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int q = min(q0 + u0 + u, cnt - 1);
          rows.load(bcast(r.idx, q), a[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int q = min(q0 + u0 + u, cnt - 1);
          const T x = bcast(r.x, q), delta = bcast(r.delta, q);
          T ll[4];  // the neighbour's em row, read as uniform (scalar) loads
#pragma unroll
          for (int c = 0; c < 4; ++c) ll[c] = em[(size_t)(c0 + q) * 4 + c];
          // columns k >= M: zero coefficients times dy = 0
          T v[KQ], d[KQ];
#pragma unroll
          for (int kq = 0; kq < KQ; ++kq) quintic(a[u][kq], x, delta, v[kq], d[kq]);
          T p[4];
#pragma unroll
          for (int kq = 0; kq < KQ; ++kq) {
            const T proj = ll[0] * dy[kq][0] + ll[1] * dy[kq][1] + ll[2] * dy[kq][2] + ll[3] * dy[kq][3];
            if (kq == 0) {
              p[0] = v[kq] * dy[kq][0] + d[kq] * proj;
#pragma unroll
              for (int c = 1; c < 4; ++c) p[c] = v[kq] * dy[kq][c];
            } else {
              p[0] += v[kq] * dy[kq][0] + d[kq] * proj;
#pragma unroll
              for (int c = 1; c < 4; ++c) p[c] += v[kq] * dy[kq][c];
            }
          }
#pragma unroll
          for (int c = 0; c < 4; ++c) w[(u0 + u) * 4 + c] = p[c];
        }
      }
      using TS = TSum<T, QG * 4>;
      const T tot = TS::sum(w);
      const int val = TS::value_of(lane), qq = val >> 2;
      const int j = c0 + q0 + qq;
      if (TS::owner(lane) && q0 + qq < cnt) {
        const T scale = (j == jt) ? (T)(nb - jt) : (T)1;
        nd[(size_t)j * 4 + (val & 3)] = tot * scale;
      }
    }
  }
  for (int j = jend + lane; j < nb; j += kWave)
#pragma unroll
    for (int c = 0; c < 4; ++c) nd[(size_t)j * 4 + c] = 0;
}

// dy = dE/dxyz of row g in the lanes-over-columns layout (lane k holds
// column k + 64 kq):  dE/dT[c, k] = sum_a G[k, a] T[c, a] + [k < A] sum_m
// G[m, k] T[c, m] with G = dE/dD (row pos[g]) and T = xyz / nnei, then
// dy = dE/dT / nnei.  Columns k >= M come out zero.
template <typename T, typename TG, int KQ>
__device__ __forceinline__ void bwd_dy(T (&dy)[KQ][4], const T* __restrict__ xyz_g, const TG* __restrict__ G, int M,
                                       int A, T inv_nnei) {
  const int lane = lane_id();
  T Tk[KQ][4];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
#pragma unroll
    for (int c = 0; c < 4; ++c) Tk[kq][c] = k < M ? xyz_g[(size_t)c * M + k] * inv_nnei : (T)0;
  }
  T ta[16][4];  // T[c, a] for a < A, broadcast
#pragma unroll
  for (int a = 0; a < 16; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) ta[a][c] = a < A ? bcast(Tk[0][c], a) : (T)0;
  // dT[c, k] = sum_a G[k, a] T[c, a];  per-lane terms of S[c, a] = sum_m G[m, a] T[c, m]
  T dT[KQ][4];
  T sp[64];
#pragma unroll
  for (int v = 0; v < 64; ++v) sp[v] = 0;
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
#pragma unroll
    for (int c = 0; c < 4; ++c) dT[kq][c] = 0;
    const int k = lane + kq * kWave;
    // every load issued before the first use (indices clamped into the
    // row, the values masked), so the 16 reads are in flight together
    const bool kin = k < M;
    const size_t kr = (size_t)(kin ? k : 0) * A;
    TG gv[16];
#pragma unroll
    for (int a = 0; a < 16; ++a) gv[a] = G[kr + (a < A ? a : 0)];
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const T gk = kin && a < A ? (T)gv[a] : (T)0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        dT[kq][c] += gk * ta[a][c];
        sp[c * 16 + a] += gk * Tk[kq][c];
      }
    }
  }
  // value c*16 + a of the transposed sum is S[c, a]; lane a < A adds S[c, a]
  // to dT[c, a]
  using TS = TSum<T, 64>;
  const T S = TS::sum(sp);  // Tk already carries 1 / nnei
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const T sv = __shfl(S, TS::lane_of(c * 16 + (lane & 15)), kWave);
    if (lane < A) dT[0][c] += sv;
  }
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
    for (int c = 0; c < 4; ++c) dy[kq][c] = dT[kq][c] * inv_nnei;
}

template <typename T, typename TG, int KQ, bool PK>
__global__ void __launch_bounds__(256) k_embed_bwd(T* __restrict__ net_deriv, const T* __restrict__ xyz,
                                                  const TG* __restrict__ gD, const int* __restrict__ pos,
                                                  const T* __restrict__ tables, const long long* __restrict__ tab_off,
                                                  const Info<T>* __restrict__ infos, const int* __restrict__ net_of,
                                                  const T* __restrict__ em, const int* __restrict__ row_type,
                                                  const int* __restrict__ sec, int ntypes, long long nrows, int nnei,
                                                  int M, int A, int ldD, T inv_nnei) {
  const long long g = wave_index();
  const int lane = lane_id();
  if (g >= nrows) return;
  const int ti = row_type[g];
  if (ti < 0) {
    for (int s = lane; s < nnei * 4; s += kWave) net_deriv[(size_t)g * nnei * 4 + s] = 0;
    return;
  }
  T dy[KQ][4];
  bwd_dy<T, TG, KQ>(dy, xyz + (size_t)g * 4 * M, gD + (size_t)pos[g] * ldD, M, A, inv_nnei);
  for (int tj = 0; tj < ntypes; ++tj) {
    const int net = net_of[ti * ntypes + tj];
    const int b = sec[tj], nb = sec[tj + 1] - sec[tj];
    embed_bwd_block<T, KQ, PK>(net_deriv + ((size_t)g * nnei + b) * 4, dy, tables + tab_off[net], infos[net],
                           em + ((size_t)g * nnei + b) * 4, nb, M);
  }
}

// ---------------------------------------------------------------------------
// Adjoint of the mixed engine in groups of 16 lanes.
//
// The lanes-over-columns layout above covers M columns with 64 KQ lane slots
// (M = 100: 128 slots, 22% of every operation idle).  Here a wave splits
// into 4 groups of 16 lanes: group g takes the neighbours q = g (mod 4) of a
// block and lane l of a group holds the C = ceil(M / 16) columns
// k = l + 16 c (M = 100: 112 slots).  A neighbour's record moves from the lane
// that located it with shuffles; a group reads its table row as 16-lane,
// 256-byte spans of packed records (columns k >= M through out-of-range
// buffer offsets, which read zeros), and each neighbour's 4 sums are reduced
// over its 16 lanes with a transposing butterfly (5 exchanges).  The adjoint
// is bound by arithmetic, so the idle slots it saves are time saved; the
// forward, bound by the table reads (whose row index here would first
// travel through a shuffle), keeps the lanes-over-columns layout.

// Packed records of the lane's C columns of any table row (per-lane rows):
// rows hold pk_stride(M) >= 16 C records (zero past column M), so the C
// reads share one vector offset and differ by immediate offsets of 256 c.
template <int C>
struct GRows {
  __amdgpu_buffer_rsrc_t rs;
  int stride16, base;
  __device__ __forceinline__ GRows(const float* t, int M, int gl) {
    rs = pk_rsrc(t);
    stride16 = pk_stride(M) * 16;
    base = gl * 16;
  }
  __device__ __forceinline__ void load(int idx, float (&a)[C][6]) const {
    const int off = idx * stride16 + base;
#pragma unroll
    for (int c = 0; c < C; ++c) load_coef_pk(rs, 0, off + 256 * c, a[c]);
  }
};

// Sum over the 16 lanes of a group of the 4 values p[0..3]: lane l of the
// group ends with the total of value l >> 2 (exchanges across lane bits 3
// and 2 keep half of the values each, then bits 1 and 0 add).
__device__ __forceinline__ float group_sum4(const float (&p)[4], int gl) {
  const bool h3 = gl & 8, h2 = gl & 4;
  const float k0 = h3 ? p[2] : p[0], k1 = h3 ? p[3] : p[1];
  const float s0 = h3 ? p[0] : p[2], s1 = h3 ? p[1] : p[3];
  const float t0 = k0 + __shfl_xor(s0, 8, kWave), t1 = k1 + __shfl_xor(s1, 8, kWave);
  float u = (h2 ? t1 : t0) + __shfl_xor(h2 ? t0 : t1, 4, kWave);
  u += dpp_partner<1>(u);
  u += dpp_partner<0>(u);
  return u;
}

template <int C>
__device__ void bwd_block_g(float* __restrict__ nd, const float (&dy)[C][4], const float* __restrict__ table,
                            const Info<float>& f, const float* __restrict__ em, int nb, int M, int grp, int gl) {
  if (nb <= 0) return;
  const int lane = lane_id();
  const float last = em[(size_t)(nb - 1) * 4];
  int jt = nb;  // first neighbour of the sorted zero-padding tail (tabulate.cc:188-201)
  for (int c0 = 0; c0 < nb; c0 += kWave) {
    const int j = c0 + lane;
    bool tail = false;
    if (j < nb) {
      const float* r = em + (size_t)j * 4;
      tail = r[0] == last && r[1] == 0.f && r[2] == 0.f && r[3] == 0.f;
    }
    const unsigned long long m = __ballot(tail);
    if (m) {
      jt = c0 + __ffsll((long long)m) - 1;
      break;
    }
  }
  const int jend = jt < nb ? jt + 1 : nb;
  const GRows<C> rows(table, M, gl);
  for (int c0 = 0; c0 < jend; c0 += kWave) {
    Nbr<float, 4> r;
    load_nbr(r, f, em + (size_t)c0 * 4, 4, em + (size_t)c0 * 4, lane, jend - c0, last, false);
    const int cnt = min(kWave, jend - c0);
    for (int q0 = 0; q0 < cnt; q0 += 8) {
      float a[2][C][6], x[2], delta[2], ll[2][4];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int qc = min(q0 + 4 * u + grp, cnt - 1);
        x[u] = __shfl(r.x, qc, kWave);
        delta[u] = __shfl(r.delta, qc, kWave);
#pragma unroll
        for (int m = 0; m < 4; ++m) ll[u][m] = __shfl(r.ll[m], qc, kWave);
        rows.load(__shfl(r.idx, qc, kWave), a[u]);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        // columns k >= M: zero coefficients times dy = 0.  p[m] = sum_k
        // P_k dy[k, m] and e[m] = sum_k P'_k dy[k, m] over the lane's
        // columns; the slope term of component 0 is ll . e.
        // outside the table P_k + P'_k delta (linear extension): its share
        // of p is delta * e, added once per neighbour instead of per column
        float p[4], e[4];
#pragma unroll
        for (int c = 0; c < C; ++c) {
          float v, d;
          quintic(a[u][c], x[u], 0.f, v, d);
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            p[m] = c == 0 ? v * dy[c][m] : p[m] + v * dy[c][m];
            e[m] = c == 0 ? d * dy[c][m] : e[m] + d * dy[c][m];
          }
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) p[m] += delta[u] * e[m];
        p[0] += ll[u][0] * e[0] + ll[u][1] * e[1] + ll[u][2] * e[2] + ll[u][3] * e[3];
        const float tot = group_sum4(p, gl);
        const int q = q0 + 4 * u + grp, j = c0 + q;
        if (q < cnt && (gl & 3) == 0) {
          const float scale = (j == jt) ? (float)(nb - jt) : 1.f;
          nd[(size_t)j * 4 + (gl >> 2)] = tot * scale;
        }
      }
    }
  }
  for (int j = jend + lane; j < nb; j += kWave)
#pragma unroll
    for (int c = 0; c < 4; ++c) nd[(size_t)j * 4 + c] = 0.f;
}

template <int C>
__global__ void __launch_bounds__(256) k_embed_bwd_g(float* __restrict__ net_deriv, const float* __restrict__ xyz,
                                                     const float* __restrict__ gD, const int* __restrict__ pos,
                                                     const float* __restrict__ tables,
                                                     const long long* __restrict__ tab_off,
                                                     const Info<float>* __restrict__ infos,
                                                     const int* __restrict__ net_of, const float* __restrict__ em,
                                                     const int* __restrict__ row_type, const int* __restrict__ sec,
                                                     int ntypes, long long nrows, int nnei, int M, int A, int ldD,
                                                     float inv_nnei) {
  constexpr int KQO = (C + 3) / 4;  // lanes-over-columns groups of dy
  const long long g = wave_index();
  const int lane = lane_id(), grp = lane >> 4, gl = lane & 15;
  if (g >= nrows) return;
  const int ti = row_type[g];
  if (ti < 0) {
    for (int s = lane; s < nnei * 4; s += kWave) net_deriv[(size_t)g * nnei * 4 + s] = 0.f;
    return;
  }
  float dyo[KQO][4];
  bwd_dy<float, float, KQO>(dyo, xyz + (size_t)g * 4 * M, gD + (size_t)pos[g] * ldD, M, A, inv_nnei);
  // column k = gl + 16 c sits on lane k % 64 of group k / 64
  float dy[C][4];
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int m = 0; m < 4; ++m) dy[c][m] = __shfl(dyo[c / 4][m], gl + 16 * (c % 4), kWave);
  for (int tj = 0; tj < ntypes; ++tj) {
    const int net = net_of[ti * ntypes + tj];
    const int b = sec[tj], nb = sec[tj + 1] - sec[tj];
    bwd_block_g<C>(net_deriv + ((size_t)g * nnei + b) * 4, dy, tables + tab_off[net], infos[net],
                   em + ((size_t)g * nnei + b) * 4, nb, M, grp, gl);
  }
}

// ---------------------------------------------------------------------------
// Matrix-core kernels of the mixed engine: the table embedding in table space.
//
// With t' = t / s and delta' = delta / s (s the stride of the row's table
// segment) the quintic of neighbour j at table row r is
//   G_k(x_j) = sum_p a'_p[r, k] T_j[p],  a'_p = a_p s^p,
//   T_j[p] = t'^p + delta' p t'^(p-1),
// so the contraction becomes a product over (row, p):
//   xyz[c, k] = sum_{r, p} Q[r, p, c] a'_p[r, k],
//   Q[r, p, c] = sum of em[j, c] T_j[p] over the neighbours j at row r.
// A row block's neighbours are sorted by distance, so their table rows do
// not increase along the list.  A wave takes 4 atoms of one type (N = 4
// atoms x 4 components = 16), walks their neighbour lists together through
// windows of 16 table rows (top row first), accumulates the window's Q in
// LDS and multiplies it with the table on the matrix cores
// (v_mfma_f32_16x16x32_f16, K = 4 rows x 8 p).  Both operands are split
// into FP16 hi + lo parts and hi*hi + hi*lo + lo*hi are accumulated in FP32
// (about 22 bits of each operand).
// The adjoint walks the same windows with the unscaled coefficients (the
// em_x derivative would multiply the errors of a'_p by 1 / s), each power p
// scaled by a power of two sigma_p that brings the largest |a_p| of the net
// near the top of the FP16 range:  H[r, p, (a, c)] = sum_k sigma_p a_p[r, k]
// dy_a[k, c] for the row pairs the window's neighbours use, then per
// neighbour (t, delta unscaled)
//   nd[j, c] = sum_p (t^p + delta p t^(p-1)) / sigma_p  H[r_j, p, c],
//   nd[j, 0] += sum_p p t^(p-1) / sigma_p  sum_c em[j, c] H[r_j, p, c].
// The sorted-padding tail of a block (first padded slot jt, weight nb - jt,
// tabulate.cc:188-201) is evaluated per atom with the FP32 table.
typedef _Float16 mhalf8 __attribute__((ext_vector_type(8)));
typedef float mfloat4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const mhalf8 gmhalf8;

constexpr int kTA = kEmbedTileAtoms;  // atoms per wave
constexpr int kWinRows = 16;          // table rows per window
constexpr int kRS = 36;               // floats per (atom, row): [c][8 p] + 4 pad
constexpr int kAS = kWinRows * kRS;   // floats per atom
constexpr int kWinF = kTA * kAS;      // floats of one wave's window
constexpr int kEW = 4;                // waves per workgroup
constexpr float kHalfMax = 32768.f;   // operands at or above this magnitude are flagged

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

__device__ __forceinline__ void split8(const float4& v0, const float4& v1, mhalf8& hi, mhalf8& lo) {
  const float v[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const _Float16 h = (_Float16)v[i];
    hi[i] = h;
    lo[i] = (_Float16)(v[i] - (float)h);
  }
}

__device__ __forceinline__ bool is_tail(const float4& r, float last) {
  return r.x == last && r.y == 0.f && r.z == 0.f && r.w == 0.f;
}

// First slot of the sorted zero-padding tail of a block of nb slots (nb if
// there is none).
__device__ __forceinline__ int find_tail(const float* __restrict__ e, int nb, float last) {
  const int lane = lane_id();
  for (int c0 = 0; c0 < nb; c0 += kWave) {
    const int j = c0 + lane;
    bool tail = false;
    if (j < nb) tail = is_tail(*reinterpret_cast<const float4*>(e + (size_t)j * 4), last);
    const unsigned long long m = __ballot(tail);
    if (m) return c0 + __ffsll((long long)m) - 1;
  }
  return nb;
}

// Powers of one neighbour: T[p] = t'^p + delta' p t'^(p-1) and, for the
// adjoint, Td[p] = p t'^(p-1) / s.
__device__ __forceinline__ void powers(float tp, float dp, float (&T)[6]) {
  const float t2 = tp * tp, t3 = t2 * tp, t4 = t3 * tp;
  T[0] = 1.f;
  T[1] = tp + dp;
  T[2] = t2 + 2.f * dp * tp;
  T[3] = t3 + 3.f * dp * t2;
  T[4] = t4 + 4.f * dp * t3;
  T[5] = t4 * tp + 5.f * dp * t4;
}

// Walk state of the tile's atoms in one row block.
struct Walk {
  const float* e[kTA];
  float last[kTA];
  int cur[kTA];
  bool done[kTA];
};

__device__ __forceinline__ void walk_init(Walk& w, const float* em, const long long (&g)[kTA], int na, int nnei,
                                          int b, int nb) {
#pragma unroll
  for (int i = 0; i < kTA; ++i) {
    w.done[i] = i >= na;
    w.cur[i] = 0;
    w.e[i] = w.done[i] ? em : em + ((size_t)g[i] * nnei + b) * 4;
    w.last[i] = w.done[i] ? 0.f : uni(w.e[i][(size_t)(nb - 1) * 4]);
  }
}

// Row of the neighbour at each atom's cursor; atoms at the end of their list
// or at the padding tail are done.  Returns the top row (-1: all done).
__device__ __forceinline__ int walk_top(Walk& w, const Info<float>& f, int nb) {
  int rtop = -1;
#pragma unroll
  for (int i = 0; i < kTA; ++i) {
    if (w.done[i]) continue;
    if (w.cur[i] >= nb) {
      w.done[i] = true;
      continue;
    }
    const float4 r = *reinterpret_cast<const float4*>(w.e[i] + (size_t)w.cur[i] * 4);
    if (uni((int)is_tail(r, w.last[i]))) {
      w.done[i] = true;
      continue;
    }
    float x = r.x, d;
    int row;
    locate(f, x, row, d);
    rtop = max(rtop, uni(row));
  }
  return rtop;
}

// Neighbour (lane) of atom i from position `from`: record, row, and whether
// it lies in the window rows [R0, R0 + kWinRows) (not padding).
struct NbrW {
  float4 ll;
  float t, d;
  int row;
  bool in;
};
__device__ __forceinline__ NbrW load_w(const Walk& w, int i, int from, int nb, const Info<float>& f, int R0) {
  NbrW n;
  const int j = from + lane_id();
  n.ll = make_float4(0.f, 0.f, 0.f, 0.f);
  const bool ok = j < nb;
  if (ok) n.ll = *reinterpret_cast<const float4*>(w.e[i] + (size_t)j * 4);
  n.t = n.ll.x;
  locate(f, n.t, n.row, n.d);
  n.in = ok && !is_tail(n.ll, w.last[i]) && n.row >= R0 && n.row < R0 + kWinRows;
  return n;
}

template <int MT>
__global__ void __launch_bounds__(256) k_embed_fwd_mfma(const EmbedArgs a) {
  __shared__ __attribute__((aligned(16))) float s_win[kEW][kWinF];
  const int lane = lane_id();
  const int wave = uni((int)(threadIdx.x / kWave));
  const int tile = blockIdx.x * kEW + wave;
  if (tile >= a.netiles) return;
  const int4 tl = a.etiles[tile];
  const int ti = uni(tl.x), p0 = uni(tl.y), na = uni(tl.z);
  float* win = s_win[wave];
  const float* em = (const float*)a.em;
  const Info<float>* infos = (const Info<float>*)a.infos;
  const int M = a.M, nnei = a.nnei;
  long long g[kTA];
#pragma unroll
  for (int i = 0; i < kTA; ++i) g[i] = i < na ? uni(a.perm[p0 + i]) : 0;
  const int fn = lane & 15, fa = fn >> 2, fc = fn & 3, fk = lane >> 4;  // fragment coordinates
  mfloat4 acc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) acc[mt] = mfloat4{0.f, 0.f, 0.f, 0.f};
  bool range_ok = true;
  if (ti >= 0) {
    for (int tj = 0; tj < a.ntypes; ++tj) {
      const int net = a.net_of[ti * a.ntypes + tj];
      const int b = a.sec[tj], nb = a.sec[tj + 1] - b;
      if (nb <= 0) continue;
      const Info<float> f = infos[net];
      const float is0 = 1.f / f.s0, is1 = 1.f / f.s1;
      const _Float16* thi = a.tab16 + a.tab16_off[4 * net];
      const _Float16* tlo = a.tab16 + a.tab16_off[4 * net + 1];
      Walk w;
      walk_init(w, em, g, na, nnei, b, nb);
      while (true) {
        const int R1 = walk_top(w, f, nb);
        if (R1 < 0) break;
        const int R0 = max(R1 - (kWinRows - 1), 0);
        for (int q = lane; q < kWinF / 4; q += kWave) reinterpret_cast<float4*>(win)[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        unsigned kbits = 0;
#pragma unroll
        for (int i = 0; i < kTA; ++i) {
          if (w.done[i]) continue;
          while (true) {
            const NbrW n = load_w(w, i, w.cur[i], nb, f, R0);
            const unsigned long long m = __ballot(n.in);
            const int cnt = ~m == 0ull ? kWave : (int)__builtin_ctzll(~m);  // leading in-window slots
            if (lane < cnt) {
              const float is = n.row < f.n0 ? is0 : is1;
              float T[6];
              powers(n.t * is, n.d * is, T);
              float* dst = win + i * kAS + (n.row - R0) * kRS;
              const float ll[4] = {n.ll.x, n.ll.y, n.ll.z, n.ll.w};
#pragma unroll
              for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int p = 0; p < 6; ++p) {
                  const float q = T[p] * ll[c];
                  range_ok = range_ok && fabsf(q) < kHalfMax;
                  atomicAdd(dst + c * 8 + p, q);
                }
              kbits |= 1u << ((n.row - R0) >> 2);
            }
            w.cur[i] += cnt;
            if (cnt < kWave) break;
          }
        }
        unsigned kb = 0;
#pragma unroll
        for (int ks = 0; ks < kWinRows / 4; ++ks)
          if (__ballot((kbits >> ks) & 1u)) kb |= 1u << ks;
#pragma unroll
        for (int ks = 0; ks < kWinRows / 4; ++ks) {
          if (!((kb >> ks) & 1u)) continue;
          const float* src = win + fa * kAS + (4 * ks + fk) * kRS + fc * 8;
          mhalf8 bh, bl;
          split8(*reinterpret_cast<const float4*>(src), *reinterpret_cast<const float4*>(src + 4), bh, bl);
          const size_t ro = ((size_t)(R0 + 4 * ks + fk) * (MT * 16) + fn) * 8;
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) {
            const mhalf8 ah = *(gmhalf8*)(thi + ro + mt * 128);
            const mhalf8 al = *(gmhalf8*)(tlo + ro + mt * 128);
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[mt], 0, 0, 0);
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc[mt], 0, 0, 0);
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc[mt], 0, 0, 0);
          }
        }
      }
    }
  }
  if (!range_ok) atomicOr(a.flag, kFlagHalfRange);
  // stage xyz as [atom][c][16 MT] in LDS: lane (fa, fc) holds columns 16 mt + 4 fk .. +3
  constexpr int LD = MT * 16;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
    *reinterpret_cast<float4*>(win + fn * LD + mt * 16 + fk * 4) = make_float4(acc[mt][0], acc[mt][1], acc[mt][2], acc[mt][3]);
  // per atom, lanes over columns: padding-tail term (FP32 table), xyz, descriptor
  float* xyz = (float*)a.xyz;
  _Float16* D = (_Float16*)a.D;
  const float* tab32 = (const float*)a.tables;
  const float inv = (float)a.inv_nnei;
  const int A = a.A;
  constexpr int KQ = (MT * 16 + kWave - 1) / kWave;
  for (int i = 0; i < na; ++i) {
    const long long gi = uni(a.perm[p0 + i]);
    float X[KQ][4];
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) {
      const int k = min(lane + kq * kWave, LD - 1);
#pragma unroll
      for (int c = 0; c < 4; ++c) X[kq][c] = win[(i * 4 + c) * LD + k];
    }
    if (ti >= 0) {
      for (int tj = 0; tj < a.ntypes; ++tj) {
        const int b = a.sec[tj], nb = a.sec[tj + 1] - b;
        if (nb <= 0) continue;
        const int net = a.net_of[ti * a.ntypes + tj];
        const float* e = em + ((size_t)gi * nnei + b) * 4;
        const float last = uni(e[(size_t)(nb - 1) * 4]);
        const int jt = uni(find_tail(e, nb, last));
        if (jt >= nb) continue;
        const float4 r = *reinterpret_cast<const float4*>(e + (size_t)jt * 4);
        float x = r.x, d;
        int idx;
        locate(infos[net], x, idx, d);
        const float wgt = (float)(nb - jt);
        const float* row = tab32 + a.tab_off[net] + (size_t)idx * M * 6;
#pragma unroll
        for (int kq = 0; kq < KQ; ++kq) {
          float c6[6], v, dv;
          load6(row + 6 * min(lane + kq * kWave, M - 1), c6);
          quintic(c6, x, d, v, dv);
          v *= wgt;
          X[kq][0] += v * r.x;
          X[kq][1] += v * r.y;
          X[kq][2] += v * r.z;
          X[kq][3] += v * r.w;
        }
      }
    }
    float ta[16][4];
#pragma unroll
    for (int q = 0; q < 16; ++q)
#pragma unroll
      for (int c = 0; c < 4; ++c) ta[q][c] = q < A ? bcast(X[0][c], q) * inv : 0.f;
    _Float16* out = D + (size_t)(p0 + i) * a.ldD;
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) {
      const int k = lane + kq * kWave;
      if (k >= M) continue;
#pragma unroll
      for (int c = 0; c < 4; ++c) xyz[((size_t)gi * 4 + c) * M + k] = X[kq][c];
      float v[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        v[q] = 0.f;
#pragma unroll
        for (int c = 0; c < 4; ++c) v[q] += (X[kq][c] * inv) * ta[q][c];
      }
      if (A == 16) {
        mhalf8 h0, h1;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          h0[q] = (_Float16)v[q];
          h1[q] = (_Float16)v[q + 8];
        }
        *reinterpret_cast<mhalf8*>(out + (size_t)k * 16) = h0;
        *reinterpret_cast<mhalf8*>(out + (size_t)k * 16 + 8) = h1;
      } else {
#pragma unroll
        for (int q = 0; q < 16; ++q)
          if (q < A) out[(size_t)k * A + q] = (_Float16)v[q];
      }
    }
  }
}

template <int KS>
__global__ void __launch_bounds__(256) k_embed_bwd_mfma(const EmbedArgs a) {
  __shared__ __attribute__((aligned(16))) float s_win[kEW][kWinF];
  const int lane = lane_id();
  const int wave = uni((int)(threadIdx.x / kWave));
  const int tile = blockIdx.x * kEW + wave;
  if (tile >= a.netiles) return;
  const int4 tl = a.etiles[tile];
  const int ti = uni(tl.x), p0 = uni(tl.y), na = uni(tl.z);
  float* win = s_win[wave];
  const float* em = (const float*)a.em;
  const Info<float>* infos = (const Info<float>*)a.infos;
  const int M = a.M, nnei = a.nnei;
  float* ndv = (float*)a.net_deriv;
  long long g[kTA];
#pragma unroll
  for (int i = 0; i < kTA; ++i) g[i] = i < na ? uni(a.perm[p0 + i]) : 0;
  if (ti < 0) {
    for (int i = 0; i < na; ++i) {
      const long long gi = uni(a.perm[p0 + i]);
      for (int s = lane; s < nnei * 4; s += kWave) ndv[(size_t)gi * nnei * 4 + s] = 0.f;
    }
    return;
  }
  constexpr int KP = KS * 32;                     // padded columns
  constexpr int KQ = (KP + kWave - 1) / kWave;    // columns per lane in the dy layout
  const float* tab32 = (const float*)a.tables;
  const float inv = (float)a.inv_nnei;
  // dy of each atom (lanes over columns), the padding-tail slots, and dy
  // staged as [atom][c][KP]
  float dmax = 0.f;
  for (int i = 0; i < kTA; ++i) {
    float dy[KQ][4];
    if (i < na) {
      const long long gi = uni(a.perm[p0 + i]);
      bwd_dy<float, float, KQ>(dy, (const float*)a.xyz + (size_t)gi * 4 * M,
                               (const float*)a.gD + (size_t)(p0 + i) * a.ldD, M, a.A, inv);
      for (int tj = 0; tj < a.ntypes; ++tj) {
        const int b = a.sec[tj], nb = a.sec[tj + 1] - b;
        if (nb <= 0) continue;
        const int net = a.net_of[ti * a.ntypes + tj];
        const float* e = em + ((size_t)gi * nnei + b) * 4;
        const float last = uni(e[(size_t)(nb - 1) * 4]);
        const int jt = uni(find_tail(e, nb, last));
        if (jt >= nb) continue;
        float* nd = ndv + ((size_t)gi * nnei + b) * 4;
        const float4 r = *reinterpret_cast<const float4*>(e + (size_t)jt * 4);
        float x = r.x, d;
        int idx;
        locate(infos[net], x, idx, d);
        const float* row = tab32 + a.tab_off[net] + (size_t)idx * M * 6;
        float pp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kq = 0; kq < KQ; ++kq) {
          float c6[6], v, dv;
          load6(row + 6 * min(lane + kq * kWave, M - 1), c6);
          quintic(c6, x, d, v, dv);
          const bool kin = lane + kq * kWave < M;
          v = kin ? v : 0.f;
          dv = kin ? dv : 0.f;
          const float proj = r.x * dy[kq][0] + r.y * dy[kq][1] + r.z * dy[kq][2] + r.w * dy[kq][3];
          pp[0] += v * dy[kq][0] + dv * proj;
          pp[1] += v * dy[kq][1];
          pp[2] += v * dy[kq][2];
          pp[3] += v * dy[kq][3];
        }
        const float wgt = (float)(nb - jt);
#pragma unroll
        for (int c = 0; c < 4; ++c) pp[c] = wave_sum(pp[c]) * wgt;
        if (lane == 0) *reinterpret_cast<float4*>(nd + (size_t)jt * 4) = make_float4(pp[0], pp[1], pp[2], pp[3]);
        for (int j = jt + 1 + lane; j < nb; j += kWave)
          *reinterpret_cast<float4*>(nd + (size_t)j * 4) = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    } else {
#pragma unroll
      for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
        for (int c = 0; c < 4; ++c) dy[kq][c] = 0.f;
    }
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) {
      const int k = lane + kq * kWave;
      if (k < KP)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const float v = k < M ? dy[kq][c] : 0.f;
          dmax = fmaxf(dmax, fabsf(v));
          win[(i * 4 + c) * KP + k] = v;
        }
    }
  }
  // power-of-two scale that brings the largest |dy| to [2^13, 2^14)
  for (int off = 32; off > 0; off >>= 1) dmax = fmaxf(dmax, __shfl_xor(dmax, off, kWave));
  int ex = 0;
  frexpf(uni(dmax), &ex);
  const float sc = dmax > 0.f ? ldexpf(1.f, 14 - ex) : 1.f, isc = 1.f / sc;
  const int fn = lane & 15, fa = fn >> 2, fc = fn & 3, fk = lane >> 4;
  mhalf8 bh[KS], bl[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const float* src = win + fn * KP + ks * 32 + fk * 8;
    float4 v0 = *reinterpret_cast<const float4*>(src), v1 = *reinterpret_cast<const float4*>(src + 4);
    v0.x *= sc; v0.y *= sc; v0.z *= sc; v0.w *= sc;
    v1.x *= sc; v1.y *= sc; v1.z *= sc; v1.w *= sc;
    split8(v0, v1, bh[ks], bl[ks]);
  }
  for (int tj = 0; tj < a.ntypes; ++tj) {
    const int net = a.net_of[ti * a.ntypes + tj];
    const int b = a.sec[tj], nb = a.sec[tj + 1] - b;
    if (nb <= 0) continue;
    const Info<float> f = infos[net];
    float isg[6];  // 1 / sigma_p
#pragma unroll
    for (int q = 0; q < 6; ++q) isg[q] = a.tab16_pscale[8 * net + q];
    const _Float16* thi = a.tab16 + a.tab16_off[4 * net + 2];
    const _Float16* tlo = a.tab16 + a.tab16_off[4 * net + 3];
    Walk w;
    walk_init(w, em, g, na, nnei, b, nb);
    while (true) {
      const int R1 = walk_top(w, f, nb);
      if (R1 < 0) break;
      const int R0 = max(R1 - (kWinRows - 1), 0);
      // pass 1: in-window extent of each atom and the row pairs used
      int cnt[kTA];
      unsigned pbits = 0;
#pragma unroll
      for (int i = 0; i < kTA; ++i) {
        cnt[i] = 0;
        if (w.done[i]) continue;
        while (true) {
          const NbrW n = load_w(w, i, w.cur[i] + cnt[i], nb, f, R0);
          const unsigned long long m = __ballot(n.in);
          const int c = ~m == 0ull ? kWave : (int)__builtin_ctzll(~m);
          if (lane < c) pbits |= 1u << ((n.row - R0) >> 1);
          cnt[i] += c;
          if (c < kWave) break;
        }
      }
      unsigned pb = 0;
#pragma unroll
      for (int q = 0; q < kWinRows / 2; ++q)
        if (__ballot((pbits >> q) & 1u)) pb |= 1u << q;
      // H of the used row pairs: A row fn = (row R0 + 2 mt + fn / 8, p = fn % 8)
#pragma unroll
      for (int mt = 0; mt < kWinRows / 2; ++mt) {
        if (!((pb >> mt) & 1u)) continue;
        mfloat4 h = mfloat4{0.f, 0.f, 0.f, 0.f};
        const size_t ao = ((size_t)(R0 + 2 * mt + (fn >> 3)) * 8 + (fn & 7)) * KP + fk * 8;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const mhalf8 ah = *(gmhalf8*)(thi + ao + ks * 32);
          const mhalf8 al = *(gmhalf8*)(tlo + ao + ks * 32);
          h = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[ks], h, 0, 0, 0);
          h = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[ks], h, 0, 0, 0);
          h = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[ks], h, 0, 0, 0);
        }
        // lane holds rows 4 fk + e: row R0 + 2 mt + fk / 2, p = 4 (fk % 2) + e
        *reinterpret_cast<float4*>(win + fa * kAS + (2 * mt + (fk >> 1)) * kRS + fc * 8 + 4 * (fk & 1)) =
            make_float4(h[0] * isc, h[1] * isc, h[2] * isc, h[3] * isc);
      }
      // pass 2: net_deriv of the window's neighbours
#pragma unroll
      for (int i = 0; i < kTA; ++i) {
        if (w.done[i] || cnt[i] == 0) continue;
        float* nd = ndv + ((size_t)g[i] * nnei + b) * 4;
        for (int c0 = 0; c0 < cnt[i]; c0 += kWave) {
          const NbrW n = load_w(w, i, w.cur[i] + c0, nb, f, R0);
          if (c0 + lane < cnt[i]) {
            const float tp = n.t;
            float T[6];
            powers(tp, n.d, T);
            const float t2 = tp * tp;
            float Td[6] = {0.f, 1.f, 2.f * tp, 3.f * t2, 4.f * t2 * tp, 5.f * t2 * t2};
#pragma unroll
            for (int q = 0; q < 6; ++q) {
              T[q] *= isg[q];
              Td[q] *= isg[q];
            }
            const float* hs = win + i * kAS + (n.row - R0) * kRS;
            const float ll[4] = {n.ll.x, n.ll.y, n.ll.z, n.ll.w};
            float o[4] = {0.f, 0.f, 0.f, 0.f};
            float pr[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              const float4 h0 = *reinterpret_cast<const float4*>(hs + c * 8);
              const float2 h1 = *reinterpret_cast<const float2*>(hs + c * 8 + 4);
              const float hv[6] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y};
#pragma unroll
              for (int p = 0; p < 6; ++p) {
                o[c] += T[p] * hv[p];
                pr[p] += ll[c] * hv[p];
              }
            }
#pragma unroll
            for (int p = 1; p < 6; ++p) o[0] += Td[p] * pr[p];
            *reinterpret_cast<float4*>(nd + (size_t)(w.cur[i] + c0 + lane) * 4) = make_float4(o[0], o[1], o[2], o[3]);
          }
        }
        w.cur[i] += cnt[i];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Adjoint of the mixed engine from the table rows of each atom's neighbour
// runs, on the matrix cores (one wavefront per centre row).
//
// With T_j[p] = t^p + p delta t^(p-1) (t the neighbour's offset in its table
// row r_j, delta its linear extension beyond the table) the table value is
// P_k(x_j) = sum_p a_p[r_j, k] T_j[p] (tabulate.cc:234-341), so
//   nd[j, c]  = sum_k P_k(x_j) dy[c, k]     = sum_p T_j[p] H[r_j, p, c],
//   nd[j, 0] += sum_k P'_k(x_j) sum_c em[j, c] dy[c, k]
//             = sum_p p t^(p-1) sum_c em[j, c] H[r_j, p, c],
//   H[r, p, c] = sum_k a_p[r, k] dy[c, k].
// A block's neighbours are sorted by distance, so equal table rows come in
// runs (2-3 neighbours in water, about 9 in fcc copper with davg 0 / dstd 1):
// H is formed once per run, 24 products per (row, column) instead of 17 per
// (neighbour, column), from one read of the row's coefficients.
// H of two runs is one chain of v_mfma_f32_16x16x16_f16 over the columns.
// Its A operand holds 8 rows per table row -- sigma_0 a_0 and sigma_1 a_1 as
// FP16 hi + lo parts, sigma_p a_p (p >= 2) in FP16, as the packed FP32
// records hold them -- and its B operand the columns dy_hi (0..3) and dy_lo
// (4..7) of dy scaled by a power of two per atom, so each product
// accumulates (a_hi + a_lo)(dy_hi + dy_lo) in FP32.  Runs are taken per
// chunk of 64 neighbours (a run that crosses a chunk boundary is formed in
// both); H reaches the neighbour's lane through LDS.
// Run tables (build_run_tables), per net: [row][kc][kq][a][4 halves] holds
// column k = 16 kc + 4 kq + i of A-row a, 256 bytes per (row, kc), so the
// 32 lanes of one table row read 256 contiguous bytes per product.
typedef _Float16 mhalf4 __attribute__((ext_vector_type(4)));

constexpr int kRunWaves = 4;  // wavefronts per workgroup
constexpr int kRunH = 24;     // floats of H per run: [p][c]

// Four halves of a run table through a buffer resource (kDeadOffset: zeros).
__device__ __forceinline__ mhalf4 load_h4(__amdgpu_buffer_rsrc_t rs, int off) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
  return __builtin_bit_cast(mhalf4, v);
}

// Lane i of each row of 16 takes lane i + 4's value (DPP row_shl:4, on the
// vector ALU instead of an LDS permute); lanes 12..15 of a row read zero.
__device__ __forceinline__ float dpp_row_shl4(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x104, 0xF, 0xF, true));
}

// Largest of a non-negative value over the wavefront, in every lane: row
// maxima by DPP row shifts (lanes past a row's end read zero), then the four
// rows' lane 0 (max is exact in any order).
__device__ __forceinline__ float wave_max_nonneg(float v) {
#pragma unroll
  for (int n = 1; n < 16; n <<= 1) {
    const int ctrl = 0x100 + n;
    float u;
    switch (ctrl) {
      case 0x101: u = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x101, 0xF, 0xF, true)); break;
      case 0x102: u = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x102, 0xF, 0xF, true)); break;
      case 0x104: u = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x104, 0xF, 0xF, true)); break;
      default: u = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x108, 0xF, 0xF, true)); break;
    }
    v = fmaxf(v, u);
  }
  const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
  const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
  const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
  const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
  return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

// Orders the LDS accesses of one wavefront's lanes (wave-private buffers).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int KC>
__global__ void __launch_bounds__(256) k_embed_bwd_runs(const EmbedArgs a) {
  constexpr int KP = KC * 16;                   // padded columns
  constexpr int KQ = (KP + kWave - 1) / kWave;  // columns per lane in the dy layout
  constexpr int GS = 17;                        // LDS row stride of dE/dD (A <= 16; odd, so lanes k miss no bank)
  constexpr int TOFF = KP * GS;                 // T = xyz / nnei staged after dE/dD, [c][KP]
  constexpr int BUF = TOFF + 4 * KP > kWave * kRunH ? TOFF + 4 * KP : kWave * kRunH;
  // per wavefront: dE/dD and T, then dy, then H of the runs of a chunk
  __shared__ __attribute__((aligned(16))) float s_buf[kRunWaves][BUF];
  __shared__ int s_row[kRunWaves][kWave];
  const int lane = lane_id();
  const int wave = uni((int)(threadIdx.x / kWave));
  const long long g = (long long)blockIdx.x * kRunWaves + wave;
  if (g >= a.nrows) return;
  const int nnei = a.nnei, M = a.M;
  float* ndg = (float*)a.net_deriv + (size_t)g * nnei * 4;
  const int ti = uni(a.row_type[g]);
  if (ti < 0) {
    for (int s = lane; s < nnei * 4; s += kWave) ndg[s] = 0.f;
    return;
  }
  const int A = a.A, ldD = a.ldD > 0 ? a.ldD : M * A;
  const float inv = (float)a.inv_nnei;
  float* buf = s_buf[wave];
  // dy = dE/dxyz from G = dE/dD (se_a.py:844-851 backward), T = xyz / nnei:
  //   dT[c, k] = sum_a G[k, a] T[c, a] + [k < A] S[c, k],  S[c, a] = sum_m T[c, m] G[m, a],
  //   dy = dT / nnei.
  // G ([M][A] at row pos[g]) and T are read coalesced into LDS; S is a
  // product on the matrix cores (v_mfma_f32_16x16x4f32, four columns m per
  // step), dT a per-lane sum (lanes over columns k = lane + 64 kq).
  {
    const float* G = (const float*)a.gD + (size_t)uni(a.pos[g]) * ldD;
    const float* xyz_g = (const float*)a.xyz + (size_t)g * 4 * M;
    const int MA = M * A;
    if (A == 16 && (ldD & 3) == 0) {
      // four consecutive values of one row k per lane: LDS index e + e / 16
      for (int e0 = 0; e0 < MA; e0 += 4 * 4 * kWave) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = e0 + 4 * (u * kWave + lane);
          v[u] = *reinterpret_cast<const float4*>(G + min(e, MA - 4));
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = e0 + 4 * (u * kWave + lane);
          if (e < MA) {
            float* d = buf + e + (e >> 4);
            d[0] = v[u].x;
            d[1] = v[u].y;
            d[2] = v[u].z;
            d[3] = v[u].w;
          }
        }
      }
    } else {
      for (int e0 = 0; e0 < MA; e0 += 8 * kWave) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = G[min(e0 + u * kWave + lane, MA - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = e0 + u * kWave + lane;
          if (e < MA) {
            const int k = e / A;
            buf[k * GS + (e - k * A)] = v[u];
          }
        }
      }
    }
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) {
      const int k = lane + kq * kWave;
      float t[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) t[c] = xyz_g[(size_t)c * M + min(k, M - 1)];
      if (k < KP)
#pragma unroll
        for (int c = 0; c < 4; ++c) buf[TOFF + c * KP + k] = k < M ? t[c] * inv : 0.f;
    }
  }
  wave_lds_sync();
  float dyo[KQ][4];
  {
    float ta[16][4];  // T[c, a], wave-uniform
#pragma unroll
    for (int a4 = 0; a4 < 16; ++a4)
#pragma unroll
      for (int c = 0; c < 4; ++c) ta[a4][c] = a4 < A ? buf[TOFF + c * KP + a4] : 0.f;
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) {
      const int k = lane + kq * kWave;
      const bool kin = k < M;
      const int kr = (kin ? k : 0) * GS;
      float gk[16];
#pragma unroll
      for (int a4 = 0; a4 < 16; ++a4) gk[a4] = kin && a4 < A ? buf[kr + a4] : 0.f;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        float d = 0.f;
#pragma unroll
        for (int a4 = 0; a4 < 16; ++a4) d += gk[a4] * ta[a4][c];
        dyo[kq][c] = d;
      }
    }
    // S: A operand lane (i, q) = T[i, 4 s + q] (rows i >= 4 zero), B operand
    // lane (j, q) = G[4 s + q, j]; lane j < 16 ends with S[0..3, j]
    const int i16 = lane & 15, q4 = lane >> 4;
    mfloat4 S = mfloat4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s0 = 0; s0 < 4 * KC; s0 += 4) {
      float av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int m = 4 * (s0 + u) + q4;
        av[u] = i16 < 4 ? buf[TOFF + i16 * KP + m] : 0.f;
        bv[u] = m < M && i16 < A ? buf[m * GS + i16] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) S = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], bv[u], S, 0, 0, 0);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (lane < A) dyo[0][c] += S[c];
#pragma unroll
      for (int kq = 0; kq < KQ; ++kq) dyo[kq][c] *= inv;
    }
  }
  wave_lds_sync();  // dE/dD and T are dead; the buffer holds dy next
  // power-of-two scale that brings the largest |dy| to [2^13, 2^14)
  float dmax = 0.f;
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
    for (int c = 0; c < 4; ++c) dmax = fmaxf(dmax, fabsf(dyo[kq][c]));
  dmax = uni(wave_max_nonneg(dmax));
  int ex = 0;
  frexpf(dmax, &ex);
  const float sc = dmax > 0.f ? ldexpf(1.f, 14 - ex) : 1.f, isc = 1.f / sc;
  // scaled dy split once per element into FP16 hi and lo parts, staged as
  // halves [hi / lo][c][KP] (zero for k >= M); B fragments: lane
  // (n = lane % 16, q = lane / 16) holds columns k = 16 kc + 4 q + i of
  // dy_hi[n] (n < 4), dy_lo[n - 4] (n < 8), else zero
  _Float16* hbuf = reinterpret_cast<_Float16*>(buf);
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
    if (k < KP)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float v = dyo[kq][c] * sc;
        const _Float16 hi = (_Float16)v;
        hbuf[c * KP + k] = hi;
        hbuf[(4 + c) * KP + k] = (_Float16)(v - (float)hi);
      }
  }
  wave_lds_sync();
  const int n = lane & 15, q = lane >> 4;
  mhalf4 bf[KC];
  {
    const _Float16* src = hbuf + (n & 7) * KP + q * 4;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      const mhalf4 v = *reinterpret_cast<const mhalf4*>(src + kc * 16);
      bf[kc] = n < 8 ? v : mhalf4{(_Float16)0.f, (_Float16)0.f, (_Float16)0.f, (_Float16)0.f};
    }
  }
  wave_lds_sync();
  const Info<float>* infos = (const Info<float>*)a.infos;
  const float* emg = (const float*)a.em + (size_t)g * nnei * 4;
  int* srow = s_row[wave];
  for (int tj = 0; tj < a.ntypes; ++tj) {
    const int b = a.sec[tj], nb = a.sec[tj + 1] - b;
    if (nb <= 0) continue;
    const int net = uni(a.net_of[ti * a.ntypes + tj]);
    const Info<float> f = infos[net];
    // wave-uniform table base, so the buffer resource sits in scalar registers
    const long long toff = a.rtab_off[net];
    const long long toff_u = ((long long)uni((int)(toff >> 32)) << 32) | (unsigned int)uni((int)(toff & 0xffffffff));
    const __amdgpu_buffer_rsrc_t rs = pk_rsrc(a.rtab + toff_u);
    const int lane_b = q * 64 + (lane & 7) * 8;  // byte offset of the lane in a (row, kc) block
    float psc[6];  // 1 / (sigma_p * dy scale)
#pragma unroll
    for (int p = 0; p < 6; ++p) psc[p] = a.rtab_pscale[8 * net + p] * isc;
    const float* e = emg + (size_t)b * 4;
    float* nd = ndg + (size_t)b * 4;
    // first slot of the sorted zero-padding tail (tabulate.cc:188-201): it
    // carries weight nb - jt, later slots are zero
    const float last = uni(e[(size_t)(nb - 1) * 4]);
    const int jt = uni(find_tail(e, nb, last));
    const int jend = jt < nb ? jt + 1 : nb;
    for (int c0 = 0; c0 < jend; c0 += kWave) {
      const int j = c0 + lane;
      const bool valid = j < jend;
      const float4 ll = valid ? *reinterpret_cast<const float4*>(e + (size_t)j * 4) : make_float4(0.f, 0.f, 0.f, 0.f);
      float t = ll.x, delta;
      int row;
      locate(f, t, row, delta);
      // the previous lane's row (DPP wave_shr:1; lane 0 does not use it)
      const int prow = __builtin_amdgcn_mov_dpp(row, 0x138, 0xF, 0xF, true);
      const bool start = valid && (lane == 0 || row != prow);
      const unsigned long long sm = __ballot(start);
      const int R = __popcll(sm);
      const int ridx = __popcll(sm & (~0ull >> (63 - lane))) - 1;  // this lane's run
      if (start) srow[ridx] = row;
      wave_lds_sync();
      // H of runs (r0, r0 + 1) and (r0 + 2, r0 + 3): lanes with n < 8 read
      // the first table row of a pair, n >= 8 the second (clamped runs past
      // R repeat the last one; their H is not stored)
      const bool second = lane & 8;
      for (int r0 = 0; r0 < R; r0 += 4) {
        // table rows of the two pairs, selected per lane without an indexed
        // array (which would live in scratch memory); lanes of a run past R,
        // or whose four columns all lie past M, read zeros through an
        // out-of-range offset (no memory traffic)
        const int ra0 = uni(srow[min(r0, R - 1)]), rb0 = uni(srow[min(r0 + 1, R - 1)]);
        const int ra1 = uni(srow[min(r0 + 2, R - 1)]), rb1 = uni(srow[min(r0 + 3, R - 1)]);
        const int off0 = (second ? rb0 : ra0) * (KC * 256) + lane_b;
        const int off1 = (second ? rb1 : ra1) * (KC * 256) + lane_b;
        const bool live0 = (second ? r0 + 1 : r0) < R, live1 = (second ? r0 + 3 : r0 + 2) < R;
        mhalf4 av[2][KC];
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
          const bool col = 16 * kc + 4 * q < M;
          av[0][kc] = load_h4(rs, live0 && col ? off0 + kc * 256 : kDeadOffset);
          av[1][kc] = load_h4(rs, live1 && col ? off1 + kc * 256 : kDeadOffset);
        }
        mfloat4 h[2] = {mfloat4{0.f, 0.f, 0.f, 0.f}, mfloat4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
#pragma unroll
          for (int pp = 0; pp < 2; ++pp) h[pp] = __builtin_amdgcn_mfma_f32_16x16x16f16(av[pp][kc], bf[kc], h[pp], 0, 0, 0);
        // lane (n, q) holds A-rows 4 q + v of column n; add the dy_lo column
        // n + 4 to the dy_hi column n (lanes n < 4 use the sum)
#pragma unroll
        for (int pp = 0; pp < 2; ++pp) {
          float hv[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) hv[v] = h[pp][v] + dpp_row_shl4(h[pp][v]);
          const int run = r0 + 2 * pp + (q >> 1);
          if (n < 4 && run < R) {
            float* hs = buf + run * kRunH;
            if ((q & 1) == 0) {  // A-rows 0..3: a_0 hi, lo, a_1 hi, lo
              hs[n] = (hv[0] + hv[1]) * psc[0];
              hs[4 + n] = (hv[2] + hv[3]) * psc[1];
            } else {  // A-rows 4..7: a_2 .. a_5
#pragma unroll
              for (int v = 0; v < 4; ++v) hs[4 * (2 + v) + n] = hv[v] * psc[2 + v];
            }
          }
        }
      }
      wave_lds_sync();
      if (valid) {
        const float* hs = buf + ridx * kRunH;
        const float t2 = t * t, t3 = t2 * t, t4 = t3 * t;
        const float Td[6] = {0.f, 1.f, 2.f * t, 3.f * t2, 4.f * t3, 5.f * t4};
        const float Tv[6] = {1.f, t + delta, t2 + delta * Td[2], t3 + delta * Td[3], t4 + delta * Td[4],
                             t4 * t + delta * Td[5]};
        float o[4] = {0.f, 0.f, 0.f, 0.f}, sl = 0.f;
#pragma unroll
        for (int p = 0; p < 6; ++p) {
          const float4 H = *reinterpret_cast<const float4*>(hs + 4 * p);
          o[0] += Tv[p] * H.x;
          o[1] += Tv[p] * H.y;
          o[2] += Tv[p] * H.z;
          o[3] += Tv[p] * H.w;
          if (p > 0) sl += Td[p] * (ll.x * H.x + ll.y * H.y + ll.z * H.z + ll.w * H.w);
        }
        const float scale = j == jt ? (float)(nb - jt) : 1.f;
        *reinterpret_cast<float4*>(nd + (size_t)j * 4) =
            make_float4((o[0] + sl) * scale, o[1] * scale, o[2] * scale, o[3] * scale);
      }
      wave_lds_sync();  // H and the run rows are rewritten by the next chunk
    }
    for (int jj = jend + lane; jj < nb; jj += kWave)
      *reinterpret_cast<float4*>(nd + (size_t)jj * 4) = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// ---------------------------------------------------------------------------
// Forward embedding of the mixed engine from each atom's table-row runs, on
// the matrix cores (one wavefront per centre row).
//
// With the row's stride s, t' = t / s, delta' = delta / s and a'_p = a_p s^p,
// the weighted contraction of a block (tabulate.cc:162-232, the sorted
// padding tail weighted nb - jt) is a sum over the runs of equal rows:
//   xyz[c, k] = sum_j w_j em[j, c] P_k(x_j) = sum_runs sum_p Q[run, p, c] a'_p[r, k],
//   Q[run, p, c] = sum_{j in run} w_j em[j, c] T'_j[p],  T'_j[p] = t'^p + p delta' t'^(p-1).
// Per chunk of 64 neighbours Q of each run is summed in LDS, scaled by a power
// of two (largest |Q| in [2^13, 2^14)) and split into FP16 hi + lo parts; two
// runs make one v_mfma_f32_16x16x16_f16 per block of 16 columns: A rows
// (c, hi) and (c, lo), K = 2 runs x 8 A-rows of the forward run table
// (sigma a'_0 hi, lo, sigma a'_1 hi, lo, sigma a'_2 .. sigma a'_5 in FP16,
// sigma a power of two per net), B columns k.  Each chunk's products are
// unscaled into FP32 sums; T and the descriptor follow as in k_embed_fwd.
// Forward run tables (build_fwd_run_tables), per net: [row][kb][quad][j][4
// halves] holds A-rows 4 quad + i of column k = 16 kb + j, 256 bytes per
// (row, kb).
template <int KC>
__global__ void __launch_bounds__(256) k_embed_fwd_runs(const EmbedArgs a) {
  constexpr int KP = KC * 16;                   // padded columns
  constexpr int KQ = (KP + kWave - 1) / kWave;  // columns per lane in the output layout
  // Q of a chunk as [c][8][run] (p -> slot 0, 1, 4, 5, 6, 7): the lanes of
  // different runs touch consecutive words (no bank conflicts)
  constexpr int QN = 32 * kWave;
  __shared__ __attribute__((aligned(16))) float s_q[kRunWaves][QN];
  __shared__ __attribute__((aligned(16))) float s_x[kRunWaves][4 * KP];  // xyz sums [c][KP]
  __shared__ int s_row[kRunWaves][kWave];
  const int lane = lane_id();
  const int wave = uni((int)(threadIdx.x / kWave));
  const long long g = (long long)blockIdx.x * kRunWaves + wave;
  if (g >= a.nrows) return;
  const int nnei = a.nnei, M = a.M, A = a.A;
  const int ti = uni(a.row_type[g]);
  float* sq = s_q[wave];
  float* sx = s_x[wave];
  int* srow = s_row[wave];
  const int j16 = lane & 15, q = lane >> 4;
  for (int i = lane; i < 4 * KP; i += kWave) sx[i] = 0.f;
  if (ti >= 0) {
    const Info<float>* infos = (const Info<float>*)a.infos;
    const float* emg = (const float*)a.em + (size_t)g * nnei * 4;
    // A-operand lane roles: rows i = j16 < 8 hold (c = i % 4, hi / lo = i / 4);
    // K rows 4 q .. 4 q + 3 = run 2 m + q / 2, A-rows 4 (q % 2) .. + 3
    const int cc = j16 & 3;
    const bool lo_row = j16 & 4, zero_row = j16 >= 8, odd = q & 1;
    for (int tj = 0; tj < a.ntypes; ++tj) {
      const int b = a.sec[tj], nb = a.sec[tj + 1] - b;
      if (nb <= 0) continue;
      const int net = uni(a.net_of[ti * a.ntypes + tj]);
      const Info<float> f = infos[net];
      const long long toff = a.ftab_off[net];
      const long long toff_u = ((long long)uni((int)(toff >> 32)) << 32) | (unsigned int)uni((int)(toff & 0xffffffff));
      const __amdgpu_buffer_rsrc_t rs = pk_rsrc(a.ftab + toff_u);
      const float inv_sigma = a.ftab_scale[net];
      const float is0 = 1.f / f.s0, is1 = 1.f / f.s1;
      const float* e = emg + (size_t)b * 4;
      const float last = uni(e[(size_t)(nb - 1) * 4]);
      const int jt = uni(find_tail(e, nb, last));
      const int jend = jt < nb ? jt + 1 : nb;
      for (int c0 = 0; c0 < jend; c0 += kWave) {
        const int j = c0 + lane;
        const bool valid = j < jend;
        const float4 ll = valid ? *reinterpret_cast<const float4*>(e + (size_t)j * 4) : make_float4(0.f, 0.f, 0.f, 0.f);
        float t = ll.x, delta;
        int row;
        locate(f, t, row, delta);
        const int prow = __shfl_up(row, 1, kWave);
        const bool start = valid && (lane == 0 || row != prow);
        const unsigned long long sm = __ballot(start);
        const int R = __popcll(sm);
        const int ridx = __popcll(sm & (~0ull >> (63 - lane))) - 1;
        if (start) srow[ridx] = row;
        // this lane's terms w em[c] T'[p]; a run is a contiguous range of
        // lanes, summed by a segmented inclusive scan (step d adds the value
        // d lanes down when that lane is in the same run) whose last lane
        // holds the run's Q -- deterministic, no atomics
        float v[4][6];
        {
          const float is = row < f.n0 ? is0 : is1;
          float T[6];
          powers(t * is, delta * is, T);
          const float w = valid ? (j == jt ? (float)(nb - jt) : 1.f) : 0.f;
          const float l4[4] = {ll.x * w, ll.y * w, ll.z * w, ll.w * w};
#pragma unroll
          for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int p = 0; p < 6; ++p) v[c][p] = l4[c] * T[p];
        }
        const int pos = lane - (63 - __builtin_clzll(sm & (~0ull >> (63 - lane))));  // place in the run
        int lmax = valid ? pos + 1 : 0;
        for (int off = 32; off > 0; off >>= 1) lmax = max(lmax, __shfl_xor(lmax, off, kWave));
        lmax = uni(lmax);
        for (int d = 1; d < lmax; d <<= 1) {
          const bool take = pos >= d;
#pragma unroll
          for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int p = 0; p < 6; ++p) {
              const float u = __shfl_up(v[c][p], d, kWave);
              if (take) v[c][p] += u;
            }
        }
        const bool is_end = valid && (lane == kWave - 1 || j + 1 == jend || ((sm >> (lane + 1)) & 1ull));
        float qm = 0.f;
        if (is_end) {
          float* dst = sq + ridx;
#pragma unroll
          for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int p = 0; p < 6; ++p) {
              dst[(c * 8 + (p < 2 ? p : p + 2)) * kWave] = v[c][p];
              qm = fmaxf(qm, fabsf(v[c][p]));
            }
        }
        wave_lds_sync();
        for (int off = 32; off > 0; off >>= 1) qm = fmaxf(qm, __shfl_xor(qm, off, kWave));
        qm = uni(qm);
        int ex = 0;
        frexpf(qm, &ex);
        const float qsc = qm > 0.f ? ldexpf(1.f, 14 - ex) : 1.f;
        mfloat4 acc[KC];
#pragma unroll
        for (int kb = 0; kb < KC; ++kb) acc[kb] = mfloat4{0.f, 0.f, 0.f, 0.f};
        for (int m2 = 0; m2 < R; m2 += 2) {
          const int run = m2 + (q >> 1);
          const bool rl = run < R;
          // slots 4 odd + i of Q[c][.][run]: p = 0, 0, 1, 1 (even quad) or 2 .. 5
          const float* qs = sq + (cc * 8 + (odd ? 4 : 0)) * kWave + min(run, R - 1);
          const float x0 = qs[0], x1 = qs[kWave], x2 = qs[2 * kWave], x3 = qs[3 * kWave];
          const float v[4] = {x0, odd ? x1 : x0, odd ? x2 : x1, odd ? x3 : x1};
          mhalf4 av;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float s = v[i] * qsc;
            const _Float16 hi = (_Float16)s;
            const _Float16 lo = (_Float16)(s - (float)hi);
            av[i] = zero_row || !rl ? (_Float16)0.f : (lo_row ? lo : hi);
          }
          const int r0 = uni(srow[m2]), r1 = uni(srow[min(m2 + 1, R - 1)]);
          const int base = ((q >> 1) ? r1 : r0) * (KC * 256) + odd * 128 + j16 * 8;
          mhalf4 bv[KC];
#pragma unroll
          for (int kb = 0; kb < KC; ++kb) bv[kb] = load_h4(rs, rl && 16 * kb + j16 < M ? base + kb * 256 : kDeadOffset);
#pragma unroll
          for (int kb = 0; kb < KC; ++kb) acc[kb] = __builtin_amdgcn_mfma_f32_16x16x16f16(av, bv[kb], acc[kb], 0, 0, 0);
        }
        // rows (c, hi) sit in lanes 0..15, (c, lo) in lanes 16..31; the
        // chunk's unscaled sums are added to xyz in LDS by lanes 0..15
        const float us = inv_sigma / qsc;
#pragma unroll
        for (int kb = 0; kb < KC; ++kb)
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const float sum = (acc[kb][c] + __shfl_xor(acc[kb][c], 16, kWave)) * us;
            if (lane < 16) sx[c * KP + 16 * kb + lane] += sum;
          }
        wave_lds_sync();  // Q and the run rows are rewritten by the next chunk
      }
    }
  }
  // xyz from LDS into the lanes-over-columns layout
  wave_lds_sync();
  float X[KQ][4];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
#pragma unroll
    for (int c = 0; c < 4; ++c) X[kq][c] = k < M ? sx[c * KP + k] : 0.f;
  }
  float* xyz = (float*)a.xyz;
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
    if (k < M)
#pragma unroll
      for (int c = 0; c < 4; ++c) xyz[((size_t)g * 4 + c) * M + k] = X[kq][c];
  }
  // descriptor D[pos[g], k, a] = sum_c T[c, k] T[c, a], T = xyz / nnei (a < A in lanes 0..A-1)
  const float inv = (float)a.inv_nnei;
  const int ldD = a.ldD > 0 ? a.ldD : M * A;
  _Float16* out = (_Float16*)a.D + (size_t)uni(a.pos[g]) * ldD;
  float ta[16][4];
#pragma unroll
  for (int q2 = 0; q2 < 16; ++q2)
#pragma unroll
    for (int c = 0; c < 4; ++c) ta[q2][c] = q2 < A ? bcast(X[0][c], q2) * inv : 0.f;
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
    if (k >= M) continue;
    float v[16];
#pragma unroll
    for (int q2 = 0; q2 < 16; ++q2) {
      v[q2] = 0.f;
#pragma unroll
      for (int c = 0; c < 4; ++c) v[q2] += (X[kq][c] * inv) * ta[q2][c];
    }
    if (A == 16 && (ldD & 7) == 0) {
      mhalf8 h0, h1;
#pragma unroll
      for (int q2 = 0; q2 < 8; ++q2) {
        h0[q2] = (_Float16)v[q2];
        h1[q2] = (_Float16)v[q2 + 8];
      }
      *reinterpret_cast<mhalf8*>(out + (size_t)k * 16) = h0;
      *reinterpret_cast<mhalf8*>(out + (size_t)k * 16 + 8) = h1;
    } else {
#pragma unroll
      for (int q2 = 0; q2 < 16; ++q2)
        if (q2 < A) out[(size_t)k * A + q2] = (_Float16)v[q2];
    }
  }
}

// ---------------------------------------------------------------------------
// Uncompressed embedding of the FP64 engine (se_a.py:785-839 without
// compression).  The embedding net runs as row-wise layer products
// (fit64.hip) over rows laid out net by net within a chunk of descriptor
// positions: the slots of block tj of atom g are the rows
// rowbase[g * ntypes + tj] + s.  These kernels move values between the
// neighbour slots and those rows, with one wavefront per atom (lanes own
// columns k = lane + 64 kq):
//   gather:    x0[row] = em[g, j, 0]
//   contract:  xyz[g, c, k] = sum_j em[g, j, c] G[row(j), k], then T and D
//              as the table kernels form them
//   adjoint:   dy = dE/dxyz from dE/dD;  dE/dG[row, k] = sum_c em[g, j, c] dy[c, k]
//              (kept for the residual of the last layer) and its dE/dz
//              (times idt and tanh'), and net_deriv[g, j, c] = sum_k G[row, k] dy[c, k]
//   slope:     net_deriv[g, j, 0] += dE/dx0[row]   (the net's input is em[g, j, 0])
// Padding slots are evaluated like the others, as the reference does.

__global__ void __launch_bounds__(256) k_net_gather(const double* __restrict__ em, const int* __restrict__ perm,
                                                    const int* __restrict__ type, const int* __restrict__ rowbase,
                                                    const int* __restrict__ sec, int ntypes, int nnei, int q0, int q1,
                                                    double* __restrict__ x0) {
  const long long q = q0 + wave_index();
  if (q >= q1) return;
  const int lane = lane_id(), g = perm[q];
  if (type[g] < 0) return;
  for (int tj = 0; tj < ntypes; ++tj) {
    const int rb = rowbase[(size_t)g * ntypes + tj], b = sec[tj], nb = sec[tj + 1] - b;
    for (int s = lane; s < nb; s += kWave) x0[(size_t)rb + s] = em[((size_t)g * nnei + b + s) * 4];
  }
}

template <int KQ>
__global__ void __launch_bounds__(256) k_net_contract(const double* __restrict__ G, const double* __restrict__ em,
                                                      const int* __restrict__ perm, const int* __restrict__ pos,
                                                      const int* __restrict__ type, const int* __restrict__ rowbase,
                                                      const int* __restrict__ sec, int ntypes, int nnei, int M, int A,
                                                      int ldD, double inv_nnei, int q0, int q1,
                                                      double* __restrict__ xyz, double* __restrict__ D) {
  const long long q = q0 + wave_index();
  if (q >= q1) return;
  const int lane = lane_id(), g = perm[q], ti = type[g];
  double acc[KQ][4];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[kq][c] = 0;
  if (ti >= 0) {
    for (int tj = 0; tj < ntypes; ++tj) {
      const int rb = rowbase[(size_t)g * ntypes + tj], b = sec[tj], nb = sec[tj + 1] - b;
      const double* e = em + ((size_t)g * nnei + b) * 4;
      for (int s = 0; s < nb; ++s) {
        const double e0 = e[4 * s], e1 = e[4 * s + 1], e2 = e[4 * s + 2], e3 = e[4 * s + 3];
        const double* gr = G + ((size_t)rb + s) * M;
#pragma unroll
        for (int kq = 0; kq < KQ; ++kq) {
          const int k = lane + kq * kWave;
          const double v = k < M ? gr[k] : 0.0;
          acc[kq][0] += e0 * v;
          acc[kq][1] += e1 * v;
          acc[kq][2] += e2 * v;
          acc[kq][3] += e3 * v;
        }
      }
    }
  }
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
    if (k < M)
#pragma unroll
      for (int c = 0; c < 4; ++c) xyz[((size_t)g * 4 + c) * M + k] = acc[kq][c];
  }
  double* out = D + (size_t)pos[g] * ldD;
  for (int a = 0; a < A; ++a) {
    double ta[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) ta[c] = bcast(acc[0][c], a) * inv_nnei;
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) {
      const int k = lane + kq * kWave;
      if (k < M) {
        double v = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) v += (acc[kq][c] * inv_nnei) * ta[c];
        out[(size_t)k * A + a] = v;
      }
    }
  }
}

template <int KQ>
__global__ void __launch_bounds__(256) k_net_adjoint(
    const double* __restrict__ G, const double* __restrict__ ytop, const double* const* __restrict__ idt_top,
    int act_top, const double* __restrict__ xyz, const double* __restrict__ gD, const double* __restrict__ em,
    const int* __restrict__ perm, const int* __restrict__ pos, const int* __restrict__ type,
    const int* __restrict__ rowbase, const int* __restrict__ sec, const int* __restrict__ net_of, int ntypes,
    int nnei, int M, int A, int ldD, double inv_nnei, int q0, int q1, double* __restrict__ gG,
    double* __restrict__ dz, double* __restrict__ net_deriv) {
  const long long q = q0 + wave_index();
  if (q >= q1) return;
  const int lane = lane_id(), g = perm[q], ti = type[g];
  if (ti < 0) {
    for (int s = lane; s < nnei * 4; s += kWave) net_deriv[(size_t)g * nnei * 4 + s] = 0;
    return;
  }
  double dy[KQ][4];
  bwd_dy<double, double, KQ>(dy, xyz + (size_t)g * 4 * M, gD + (size_t)pos[g] * ldD, M, A, inv_nnei);
  for (int tj = 0; tj < ntypes; ++tj) {
    const int rb = rowbase[(size_t)g * ntypes + tj], b = sec[tj], nb = sec[tj + 1] - b;
    const double* __restrict__ idt = idt_top[net_of[ti * ntypes + tj]];
    const double* e = em + ((size_t)g * nnei + b) * 4;
    double* nd = net_deriv + ((size_t)g * nnei + b) * 4;
    for (int s = 0; s < nb; ++s) {
      const double e0 = e[4 * s], e1 = e[4 * s + 1], e2 = e[4 * s + 2], e3 = e[4 * s + 3];
      const size_t row = (size_t)rb + s;
      double p[4] = {0, 0, 0, 0};
#pragma unroll
      for (int kq = 0; kq < KQ; ++kq) {
        const int k = lane + kq * kWave;
        if (k < M) {
          const double gv = G[row * M + k];
          const double gg = e0 * dy[kq][0] + e1 * dy[kq][1] + e2 * dy[kq][2] + e3 * dy[kq][3];
          gG[row * M + k] = gg;
          double z = idt ? gg * idt[k] : gg;
          if (act_top) {
            const double yv = ytop[row * M + k];
            z *= (1. - yv * yv);
          }
          dz[row * M + k] = z;
#pragma unroll
          for (int c = 0; c < 4; ++c) p[c] += gv * dy[kq][c];
        }
      }
      const double tot = TSum<double, 4>::sum(p);
      if (lane < 4) nd[4 * s + lane] = tot;
    }
  }
}

__global__ void __launch_bounds__(256) k_net_slope(const double* __restrict__ ds, const int* __restrict__ perm,
                                                   const int* __restrict__ type, const int* __restrict__ rowbase,
                                                   const int* __restrict__ sec, int ntypes, int nnei, int q0, int q1,
                                                   double* __restrict__ net_deriv) {
  const long long q = q0 + wave_index();
  if (q >= q1) return;
  const int lane = lane_id(), g = perm[q];
  if (type[g] < 0) return;
  for (int tj = 0; tj < ntypes; ++tj) {
    const int rb = rowbase[(size_t)g * ntypes + tj], b = sec[tj], nb = sec[tj + 1] - b;
    for (int s = lane; s < nb; s += kWave) net_deriv[((size_t)g * nnei + b + s) * 4] += ds[(size_t)rb + s];
  }
}

}  // namespace

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

// Packed FP32 tables, A <= 16, M <= 128: the group-of-16 adjoint (C columns
// per lane).
bool embed_groups_apply(const EmbedArgs& a) {
  return a.packed && a.A <= 16 && a.M <= 128 && a.M > 0 && pk_stride(a.M) >= (a.M + 15) / 16 * 16;
}
#define DPMI_C_DISPATCH(M, ...)                                          \
  do {                                                                   \
    switch (((M) + 15) / 16) {                                           \
      case 1: { constexpr int C = 1; __VA_ARGS__; } break;               \
      case 2: { constexpr int C = 2; __VA_ARGS__; } break;               \
      case 3: { constexpr int C = 3; __VA_ARGS__; } break;               \
      case 4: { constexpr int C = 4; __VA_ARGS__; } break;               \
      case 5: { constexpr int C = 5; __VA_ARGS__; } break;               \
      case 6: { constexpr int C = 6; __VA_ARGS__; } break;               \
      case 7: { constexpr int C = 7; __VA_ARGS__; } break;               \
      default: { constexpr int C = 8; __VA_ARGS__; } break;              \
    }                                                                    \
  } while (0)

template <typename T, typename TD>
void launch_embed_fwd(const EmbedArgs& a, hipStream_t st) {
  if (a.nrows <= 0) return;
  const int tpb = kWave * (a.waves > 0 ? a.waves : 4);
  const dim3 grid(ceil_div(a.nrows * kWave, tpb));
  const T* tables = (const T*)(a.packed ? a.packed : a.tables);
  const long long* off = a.packed ? a.packed_off : a.tab_off;
#define DPMI_FWD_LAUNCH(PKV)                                                                                        \
  DPMI_KQ_DISPATCH(a.M, hipLaunchKernelGGL((k_embed_fwd<T, TD, KQ, PKV>), grid, dim3(tpb), 0, st, (T*)a.xyz, (TD*)a.D, \
                                           a.pos, tables, off, (const Info<T>*)a.infos, a.net_of, (const T*)a.em,     \
                                           a.row_type, a.sec, a.ntypes, a.nrows, a.nnei, a.M, a.A,                     \
                                           a.ldD > 0 ? a.ldD : a.M * a.A, (T)a.inv_nnei))
  if constexpr (sizeof(T) == 4) {
    if (a.packed)
      DPMI_FWD_LAUNCH(true);
    else
      DPMI_FWD_LAUNCH(false);
  } else {
    DPMI_FWD_LAUNCH(false);
  }
#undef DPMI_FWD_LAUNCH
  DPMI_LAUNCH_CHECK();
}

template <typename T, typename TG>
void launch_embed_bwd(const EmbedArgs& a, hipStream_t st) {
  if (a.nrows <= 0) return;
  const int tpb = kWave * (a.waves_bwd > 0 ? a.waves_bwd : 1);
  const dim3 grid(ceil_div(a.nrows * kWave, tpb));
  const T* tables = (const T*)(a.packed ? a.packed : a.tables);
  const long long* off = a.packed ? a.packed_off : a.tab_off;
  if constexpr (sizeof(T) == 4 && sizeof(TG) == 4) {
    if (embed_groups_apply(a)) {
      DPMI_C_DISPATCH(a.M, hipLaunchKernelGGL((k_embed_bwd_g<C>), grid, dim3(tpb), 0, st, (float*)a.net_deriv,
                                              (const float*)a.xyz, (const float*)a.gD, a.pos, (const float*)tables,
                                              off, (const Info<float>*)a.infos, a.net_of, (const float*)a.em,
                                              a.row_type, a.sec, a.ntypes, a.nrows, a.nnei, a.M, a.A,
                                              a.ldD > 0 ? a.ldD : a.M * a.A, (float)a.inv_nnei));
      DPMI_LAUNCH_CHECK();
      return;
    }
  }
#define DPMI_BWD_LAUNCH(PKV)                                                                                        \
  DPMI_KQ_DISPATCH(a.M, hipLaunchKernelGGL((k_embed_bwd<T, TG, KQ, PKV>), grid, dim3(tpb), 0, st, (T*)a.net_deriv,    \
                                           (const T*)a.xyz, (const TG*)a.gD, a.pos, tables, off,                      \
                                           (const Info<T>*)a.infos, a.net_of, (const T*)a.em, a.row_type, a.sec,      \
                                           a.ntypes, a.nrows, a.nnei, a.M, a.A, a.ldD > 0 ? a.ldD : a.M * a.A,        \
                                           (T)a.inv_nnei))
  if constexpr (sizeof(T) == 4) {
    if (a.packed)
      DPMI_BWD_LAUNCH(true);
    else
      DPMI_BWD_LAUNCH(false);
  } else {
    DPMI_BWD_LAUNCH(false);
  }
#undef DPMI_BWD_LAUNCH
  DPMI_LAUNCH_CHECK();
}

bool embed_mfma_supported(int M) { return M > 0 && M <= 128 && M % 4 == 0; }

void build_packed_tables(const double* table, const long long* tab_off, const int* nrows, int nnets, int M,
                         std::vector<float>& out, std::vector<long long>& off) {
  const int S = pk_stride(M);
  off.assign(nnets, 0);
  long long tot = 0;
  for (int n = 0; n < nnets; ++n) {
    off[n] = tot;
    tot += (long long)nrows[n] * S * 4;
  }
  out.assign((size_t)tot, 0.f);  // padding records stay zero
  for (int n = 0; n < nnets; ++n)
    for (long long q = 0; q < (long long)nrows[n] * M; ++q) {
      const double* a = table + tab_off[n] + q * 6;
      float* r = out.data() + off[n] + (q / M * S + q % M) * 4;
      r[0] = (float)a[0];
      r[1] = (float)a[1];
      const tab_half2 h23 = {(_Float16)a[2], (_Float16)a[3]}, h45 = {(_Float16)a[4], (_Float16)a[5]};
      r[2] = __builtin_bit_cast(float, h23);
      r[3] = __builtin_bit_cast(float, h45);
    }
}

void build_split_tables(const double* table, const long long* tab_off, const int* nrows, const double* infos,
                        int nnets, int M, std::vector<_Float16>& out, std::vector<long long>& off,
                        std::vector<float>& pscale) {
  const int LDf = (M + 15) / 16 * 16, KP = (M + 31) / 32 * 32;
  off.assign(4 * (size_t)nnets, 0);
  long long tot = 0;
  for (int n = 0; n < nnets; ++n) {
    const long long R = nrows[n] + kWinRows;
    off[4 * n] = tot;
    tot += R * LDf * 8;
    off[4 * n + 1] = tot;
    tot += R * LDf * 8;
    off[4 * n + 2] = tot;
    tot += R * 8 * KP;
    off[4 * n + 3] = tot;
    tot += R * 8 * KP;
  }
  out.assign((size_t)tot, (_Float16)0.f);
  pscale.assign(8 * (size_t)nnets, 1.f);
  auto put = [&](long long o_hi, long long o_lo, size_t at, double v) {
    const _Float16 hi = (_Float16)v;
    out[o_hi + at] = hi;
    out[o_lo + at] = (_Float16)(v - (double)hi);
  };
  for (int n = 0; n < nnets; ++n) {
    float fi[6];
    for (int q = 0; q < 6; ++q) fi[q] = (float)infos[n * 6 + q];
    const Info<float> f = make_info(fi);  // the segment rule of the kernels
    if (f.ntail >= nrows[n]) throw std::invalid_argument("table has fewer rows than its info addresses");
    const double* T = table + tab_off[n];
    // adjoint scale per power: 2^e with 2^13 <= sigma_p max|a_p| < 2^14
    double sig[6];
    for (int p = 0; p < 6; ++p) {
      double mx = 0.0;
      for (size_t q = 0; q < (size_t)nrows[n] * M; ++q) mx = std::max(mx, std::fabs(T[q * 6 + p]));
      int ex = 0;
      if (mx > 0.0) std::frexp(mx, &ex);
      sig[p] = mx > 0.0 ? std::ldexp(1.0, 14 - ex) : 1.0;
      pscale[8 * n + p] = (float)(1.0 / sig[p]);
    }
    for (int r = 0; r < nrows[n]; ++r) {
      const double s = r < f.n0 ? (double)f.s0 : (double)f.s1;
      for (int k = 0; k < M; ++k) {
        double sp = 1.0;
        for (int p = 0; p < 6; ++p) {
          const double a = T[((size_t)r * M + k) * 6 + p];
          put(off[4 * n], off[4 * n + 1], ((size_t)r * LDf + k) * 8 + p, a * sp);
          put(off[4 * n + 2], off[4 * n + 3], ((size_t)r * 8 + p) * KP + k, a * sig[p]);
          sp *= s;
        }
      }
    }
  }
}

void launch_embed_fwd_mfma(const EmbedArgs& a, hipStream_t st) {
  if (a.netiles <= 0) return;
  if (!embed_mfma_supported(a.M)) throw std::invalid_argument("matrix-core embedding needs M <= 128, M % 4 == 0");
  const dim3 grid(ceil_div(a.netiles, kEW)), block(kEW * kWave);
  switch ((a.M + 15) / 16) {
    case 1: hipLaunchKernelGGL(k_embed_fwd_mfma<1>, grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL(k_embed_fwd_mfma<2>, grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL(k_embed_fwd_mfma<3>, grid, block, 0, st, a); break;
    case 4: hipLaunchKernelGGL(k_embed_fwd_mfma<4>, grid, block, 0, st, a); break;
    case 5: hipLaunchKernelGGL(k_embed_fwd_mfma<5>, grid, block, 0, st, a); break;
    case 6: hipLaunchKernelGGL(k_embed_fwd_mfma<6>, grid, block, 0, st, a); break;
    case 7: hipLaunchKernelGGL(k_embed_fwd_mfma<7>, grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL(k_embed_fwd_mfma<8>, grid, block, 0, st, a); break;
  }
  DPMI_LAUNCH_CHECK();
}

void launch_embed_bwd_mfma(const EmbedArgs& a, hipStream_t st) {
  if (a.netiles <= 0) return;
  if (!embed_mfma_supported(a.M)) throw std::invalid_argument("matrix-core embedding needs M <= 128, M % 4 == 0");
  const dim3 grid(ceil_div(a.netiles, kEW)), block(kEW * kWave);
  switch ((a.M + 31) / 32) {
    case 1: hipLaunchKernelGGL(k_embed_bwd_mfma<1>, grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL(k_embed_bwd_mfma<2>, grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL(k_embed_bwd_mfma<3>, grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL(k_embed_bwd_mfma<4>, grid, block, 0, st, a); break;
  }
  DPMI_LAUNCH_CHECK();
}

void check_table_rows(const double* infos, const int* nrows, int nnets) {
  for (int n = 0; n < nnets; ++n) {
    float fi[6];
    for (int q = 0; q < 6; ++q) fi[q] = (float)infos[n * 6 + q];
    const Info<double> fd = make_info(infos + n * 6);
    const Info<float> ff = make_info(fi);
    if (fd.n0 < 0 || ff.n0 < 0 || fd.ntail >= nrows[n] || ff.ntail >= nrows[n])
      throw std::invalid_argument("table_info of net " + std::to_string(n) + " addresses rows past its " +
                                  std::to_string(nrows[n]) + " table rows");
  }
}

bool embed_runs_supported(int M, int A) { return M > 0 && M <= 128 && A > 0 && A <= 16; }

void build_run_tables(const double* table, const long long* tab_off, const int* nrows, int nnets, int M,
                      std::vector<_Float16>& out, std::vector<long long>& off, std::vector<float>& pscale) {
  const int KC = (M + 15) / 16;
  off.assign(nnets, 0);
  long long tot = 0;
  for (int n = 0; n < nnets; ++n) {
    off[n] = tot;
    tot += (long long)nrows[n] * KC * 128;
  }
  out.assign((size_t)tot, (_Float16)0.f);  // columns k >= M stay zero
  pscale.assign(8 * (size_t)nnets, 1.f);
  for (int n = 0; n < nnets; ++n) {
    const double* T = table + tab_off[n];
    // scale per power: 2^e with 2^13 <= sigma_p max|a_p| < 2^14
    double sig[6];
    for (int p = 0; p < 6; ++p) {
      double mx = 0.0;
      for (size_t q = 0; q < (size_t)nrows[n] * M; ++q) mx = std::max(mx, std::fabs(T[q * 6 + p]));
      int ex = 0;
      if (mx > 0.0) std::frexp(mx, &ex);
      sig[p] = mx > 0.0 ? std::ldexp(1.0, 14 - ex) : 1.0;
      pscale[8 * n + p] = (float)(1.0 / sig[p]);
    }
    for (int r = 0; r < nrows[n]; ++r)
      for (int k = 0; k < M; ++k) {
        // element (row r, column k, A-row a) at base + 4 a
        const size_t base = (size_t)off[n] + (((size_t)r * KC + k / 16) * 4 + (k % 16) / 4) * 32 + k % 4;
        const double* c = T + ((size_t)r * M + k) * 6;
        for (int p = 0; p < 2; ++p) {
          const double v = c[p] * sig[p];
          const _Float16 hi = (_Float16)v;
          out[base + 4 * (2 * p)] = hi;
          out[base + 4 * (2 * p + 1)] = (_Float16)(v - (double)hi);
        }
        for (int p = 2; p < 6; ++p) out[base + 4 * (p + 2)] = (_Float16)(c[p] * sig[p]);
      }
  }
}

void build_fwd_run_tables(const double* table, const long long* tab_off, const int* nrows, const double* infos,
                          int nnets, int M, std::vector<_Float16>& out, std::vector<long long>& off,
                          std::vector<float>& scale) {
  const int KC = (M + 15) / 16;
  off.assign(nnets, 0);
  long long tot = 0;
  for (int n = 0; n < nnets; ++n) {
    off[n] = tot;
    tot += (long long)nrows[n] * KC * 128;
  }
  out.assign((size_t)tot, (_Float16)0.f);  // columns k >= M stay zero
  scale.assign(nnets, 1.f);
  for (int n = 0; n < nnets; ++n) {
    float fi[6];
    for (int q = 0; q < 6; ++q) fi[q] = (float)infos[n * 6 + q];
    const Info<float> f = make_info(fi);  // the kernels' segment rule
    const double* T = table + tab_off[n];
    auto stride = [&](int r) { return r < f.n0 ? (double)f.s0 : (double)f.s1; };
    // one power of two per net: 2^13 <= sigma max|a_p s^p| < 2^14
    double mx = 0.0;
    for (int r = 0; r < nrows[n]; ++r) {
      const double s = stride(r);
      for (int k = 0; k < M; ++k) {
        double sp = 1.0;
        for (int p = 0; p < 6; ++p, sp *= s) mx = std::max(mx, std::fabs(T[((size_t)r * M + k) * 6 + p] * sp));
      }
    }
    int ex = 0;
    if (mx > 0.0) std::frexp(mx, &ex);
    const double sig = mx > 0.0 ? std::ldexp(1.0, 14 - ex) : 1.0;
    scale[n] = (float)(1.0 / sig);
    for (int r = 0; r < nrows[n]; ++r) {
      const double s = stride(r);
      for (int k = 0; k < M; ++k) {
        // A-row u of column k at ((r KC + k / 16) 2 + u / 4) 64 + (k % 16) 4 + u % 4
        const size_t base = ((size_t)r * KC + k / 16) * 128 + (k % 16) * 4;
        auto put = [&](int u, _Float16 v) { out[(size_t)off[n] + base + (u / 4) * 64 + u % 4] = v; };
        const double* c = T + ((size_t)r * M + k) * 6;
        double sp = sig;
        for (int p = 0; p < 6; ++p, sp *= s) {
          const double v = c[p] * sp;
          const _Float16 hi = (_Float16)v;
          if (p < 2) {
            put(2 * p, hi);
            put(2 * p + 1, (_Float16)(v - (double)hi));
          } else {
            put(p + 2, hi);
          }
        }
      }
    }
  }
}

void launch_embed_fwd_runs(const EmbedArgs& a, hipStream_t st) {
  if (a.nrows <= 0) return;
  if (!embed_runs_supported(a.M, a.A) || !a.ftab)
    throw std::invalid_argument("row-run forward needs forward run tables, M <= 128 and A <= 16");
  const dim3 grid(ceil_div(a.nrows, kRunWaves)), block(kRunWaves * kWave);
  switch ((a.M + 15) / 16) {
    case 1: hipLaunchKernelGGL(k_embed_fwd_runs<1>, grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL(k_embed_fwd_runs<2>, grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL(k_embed_fwd_runs<3>, grid, block, 0, st, a); break;
    case 4: hipLaunchKernelGGL(k_embed_fwd_runs<4>, grid, block, 0, st, a); break;
    case 5: hipLaunchKernelGGL(k_embed_fwd_runs<5>, grid, block, 0, st, a); break;
    case 6: hipLaunchKernelGGL(k_embed_fwd_runs<6>, grid, block, 0, st, a); break;
    case 7: hipLaunchKernelGGL(k_embed_fwd_runs<7>, grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL(k_embed_fwd_runs<8>, grid, block, 0, st, a); break;
  }
  DPMI_LAUNCH_CHECK();
}

void launch_embed_bwd_runs(const EmbedArgs& a, hipStream_t st) {
  if (a.nrows <= 0) return;
  if (!embed_runs_supported(a.M, a.A) || !a.rtab)
    throw std::invalid_argument("row-run adjoint needs run tables, M <= 128 and A <= 16");
  const dim3 grid(ceil_div(a.nrows, kRunWaves)), block(kRunWaves * kWave);
  switch ((a.M + 15) / 16) {
    case 1: hipLaunchKernelGGL(k_embed_bwd_runs<1>, grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL(k_embed_bwd_runs<2>, grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL(k_embed_bwd_runs<3>, grid, block, 0, st, a); break;
    case 4: hipLaunchKernelGGL(k_embed_bwd_runs<4>, grid, block, 0, st, a); break;
    case 5: hipLaunchKernelGGL(k_embed_bwd_runs<5>, grid, block, 0, st, a); break;
    case 6: hipLaunchKernelGGL(k_embed_bwd_runs<6>, grid, block, 0, st, a); break;
    case 7: hipLaunchKernelGGL(k_embed_bwd_runs<7>, grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL(k_embed_bwd_runs<8>, grid, block, 0, st, a); break;
  }
  DPMI_LAUNCH_CHECK();
}

template void launch_embed_fwd<double, double>(const EmbedArgs&, hipStream_t);
template void launch_embed_bwd<double, double>(const EmbedArgs&, hipStream_t);
template void launch_embed_fwd<float, _Float16>(const EmbedArgs&, hipStream_t);
template void launch_embed_bwd<float, _Float16>(const EmbedArgs&, hipStream_t);
template void launch_embed_bwd<float, float>(const EmbedArgs&, hipStream_t);


// ---------------------------------------------------------- uncompressed --
void launch_net_gather(const NetArgs& a, int q0, int q1, double* x0, hipStream_t st) {
  if (q1 <= q0) return;
  hipLaunchKernelGGL(k_net_gather, dim3(ceil_div(q1 - q0, 4)), dim3(256), 0, st, a.em, a.perm, a.type, a.rowbase,
                     a.sec, a.ntypes, a.nnei, q0, q1, x0);
  DPMI_LAUNCH_CHECK();
}

void launch_net_contract(const NetArgs& a, const double* G, int q0, int q1, hipStream_t st) {
  if (q1 <= q0) return;
  const dim3 grid(ceil_div(q1 - q0, 4)), block(256);
  const int ldD = a.ldD ? a.ldD : a.M * a.A;
  DPMI_KQ_DISPATCH(a.M, hipLaunchKernelGGL((k_net_contract<KQ>), grid, block, 0, st, G, a.em, a.perm, a.pos, a.type,
                                           a.rowbase, a.sec, a.ntypes, a.nnei, a.M, a.A, ldD, a.inv_nnei, q0, q1,
                                           a.xyz, a.D));
  DPMI_LAUNCH_CHECK();
}

void launch_net_adjoint(const NetArgs& a, const double* G, const double* ytop, const double* const* idt_top,
                        int act_top, int q0, int q1, double* gG, double* dz, hipStream_t st) {
  if (q1 <= q0) return;
  const dim3 grid(ceil_div(q1 - q0, 4)), block(256);
  const int ldD = a.ldD ? a.ldD : a.M * a.A;
  DPMI_KQ_DISPATCH(a.M, hipLaunchKernelGGL((k_net_adjoint<KQ>), grid, block, 0, st, G, ytop, idt_top, act_top,
                                           a.xyz, a.gD, a.em, a.perm, a.pos, a.type, a.rowbase, a.sec, a.net_of,
                                           a.ntypes, a.nnei, a.M, a.A, ldD, a.inv_nnei, q0, q1, gG, dz,
                                           a.net_deriv));
  DPMI_LAUNCH_CHECK();
}

void launch_net_slope(const NetArgs& a, const double* ds, int q0, int q1, hipStream_t st) {
  if (q1 <= q0) return;
  hipLaunchKernelGGL(k_net_slope, dim3(ceil_div(q1 - q0, 4)), dim3(256), 0, st, ds, a.perm, a.type, a.rowbase, a.sec,
                     a.ntypes, a.nnei, q0, q1, a.net_deriv);
  DPMI_LAUNCH_CHECK();
}

}  // namespace dpmi
