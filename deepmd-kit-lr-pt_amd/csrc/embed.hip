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

#include "dpmi_common.h"
#include "kernels.h"
#include "tabulate_core.h"

namespace dpmi {

namespace {

using namespace tab;

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

template <typename T, typename TD, int KQ>
__global__ void __launch_bounds__(256) k_embed_fwd(T* __restrict__ xyz, TD* __restrict__ D, const int* __restrict__ pos,
                                                   const T* __restrict__ tables, const long long* __restrict__ tab_off,
                                                   const Info<T>* __restrict__ infos, const int* __restrict__ net_of,
                                                   const T* __restrict__ em, const int* __restrict__ row_type,
                                                   const int* __restrict__ sec, int ntypes, long long nrows, int nnei,
                                                   int M, int A, int ldD, T inv_nnei) {
  const long long g = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
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
      fwd_block<T, 4, KQ>(acc, tables + tab_off[net], infos[net], e, 4, e, nb, M, true);
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

// Adjoint of one neighbour block (nb slots) of row g, lanes over columns k
// (as the forward, so each neighbour's table row is one coalesced read):
//   nd[j, c] = sum_k dy[k, c] P_k(x_j) + [c = 0] sum_k P'_k(x_j) sum_c' dy[k, c'] em[j, c']
// per-lane partial sums of QG neighbours x 4 components are reduced across
// the wave with transpose_sum, which leaves neighbour (l / 4) component
// (l % 4) on lane l for one coalesced store.  Sorted-padding tail as the
// forward: the first tail slot carries weight (nb - jt), later slots zero.
template <typename T, int KQ>
__device__ void embed_bwd_block(T* __restrict__ nd, const T (&dy)[KQ][4], const T* __restrict__ table,
                                const Info<T>& f, const T* __restrict__ em, int nb, int M) {
  if (nb <= 0) return;
  constexpr int QG = sizeof(T) == 8 ? 8 : 16;  // neighbours per reduction group
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
          const T* row = table + (size_t)bcast(r.idx, q) * M * 6;
#pragma unroll
          for (int kq = 0; kq < KQ; ++kq) load6(row + 6 * min(lane + kq * kWave, M - 1), a[u][kq]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int q = min(q0 + u0 + u, cnt - 1);
          const T x = bcast(r.x, q), delta = bcast(r.delta, q);
          T ll[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) ll[c] = bcast(r.ll[c], q);
          T p[4] = {0, 0, 0, 0};
#pragma unroll
          for (int kq = 0; kq < KQ; ++kq) {
            const bool kin = lane + kq * kWave < M;
            T v, d;
            quintic(a[u][kq], x, delta, v, d);
            v = kin ? v : (T)0;
            d = kin ? d : (T)0;
            const T proj = ll[0] * dy[kq][0] + ll[1] * dy[kq][1] + ll[2] * dy[kq][2] + ll[3] * dy[kq][3];
            p[0] += v * dy[kq][0] + d * proj;
#pragma unroll
            for (int c = 1; c < 4; ++c) p[c] += v * dy[kq][c];
          }
#pragma unroll
          for (int c = 0; c < 4; ++c) w[(u0 + u) * 4 + c] = p[c];
        }
      }
      const T tot = transpose_sum<T, QG * 4>(w);
      const int qq = (lane >> 2) & (QG - 1);
      const int j = c0 + q0 + qq;
      if (lane < QG * 4 && q0 + qq < cnt) {
        const T scale = (j == jt) ? (T)(nb - jt) : (T)1;
        nd[(size_t)j * 4 + (lane & 3)] = tot * scale;
      }
    }
  }
  for (int j = jend + lane; j < nb; j += kWave)
#pragma unroll
    for (int c = 0; c < 4; ++c) nd[(size_t)j * 4 + c] = 0;
}

template <typename T, typename TG, int KQ>
__global__ void __launch_bounds__(256) k_embed_bwd(T* __restrict__ net_deriv, const T* __restrict__ xyz,
                                                  const TG* __restrict__ gD, const int* __restrict__ pos,
                                                  const T* __restrict__ tables, const long long* __restrict__ tab_off,
                                                  const Info<T>* __restrict__ infos, const int* __restrict__ net_of,
                                                  const T* __restrict__ em, const int* __restrict__ row_type,
                                                  const int* __restrict__ sec, int ntypes, long long nrows, int nnei,
                                                  int M, int A, int ldD, T inv_nnei) {
  const long long g = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = lane_id();
  if (g >= nrows) return;
  const int ti = row_type[g];
  if (ti < 0) {
    for (int s = lane; s < nnei * 4; s += kWave) net_deriv[(size_t)g * nnei * 4 + s] = 0;
    return;
  }
  // T[c, k] for the lane's columns
  T Tk[KQ][4];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
#pragma unroll
    for (int c = 0; c < 4; ++c) Tk[kq][c] = k < M ? xyz[((size_t)g * 4 + c) * M + k] * inv_nnei : (T)0;
  }
  T ta[16][4];  // T[c, a] for a < A, broadcast
#pragma unroll
  for (int a = 0; a < 16; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) ta[a][c] = a < A ? bcast(Tk[0][c], a) : (T)0;
  const TG* G = gD + (size_t)pos[g] * ldD;
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
    if (k < M) {
#pragma unroll
      for (int a = 0; a < 16; ++a) {
        if (a < A) {
          const T gk = (T)G[(size_t)k * A + a];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            dT[kq][c] += gk * ta[a][c];
            sp[c * 16 + a] += gk * Tk[kq][c];
          }
        }
      }
    }
  }
  // S[c, a] lands on lane c*16 + a; lane a < A adds S[c, a] to dT[c, a]
  const T S = transpose_sum<T, 64>(sp);  // Tk already carries 1 / nnei
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const T sv = __shfl(S, c * 16 + (lane & 15), kWave);
    if (lane < A) dT[0][c] += sv;
  }
  T dy[KQ][4];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
    for (int c = 0; c < 4; ++c) dy[kq][c] = dT[kq][c] * inv_nnei;
  for (int tj = 0; tj < ntypes; ++tj) {
    const int net = net_of[ti * ntypes + tj];
    const int b = sec[tj], nb = sec[tj + 1] - sec[tj];
    embed_bwd_block<T, KQ>(net_deriv + ((size_t)g * nnei + b) * 4, dy, tables + tab_off[net], infos[net],
                           em + ((size_t)g * nnei + b) * 4, nb, M);
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

template <typename T, typename TD>
void launch_embed_fwd(const EmbedArgs& a, hipStream_t st) {
  if (a.nrows <= 0) return;
  const int tpb = kWave * (a.waves > 0 ? a.waves : 4);
  const dim3 grid(ceil_div(a.nrows * kWave, tpb));
  DPMI_KQ_DISPATCH(a.M, hipLaunchKernelGGL((k_embed_fwd<T, TD, KQ>), grid, dim3(tpb), 0, st, (T*)a.xyz, (TD*)a.D,
                                           a.pos, (const T*)a.tables, a.tab_off, (const Info<T>*)a.infos, a.net_of,
                                           (const T*)a.em, a.row_type, a.sec, a.ntypes, a.nrows, a.nnei, a.M, a.A,
                                           a.ldD > 0 ? a.ldD : a.M * a.A, (T)a.inv_nnei));
  DPMI_LAUNCH_CHECK();
}

template <typename T, typename TG>
void launch_embed_bwd(const EmbedArgs& a, hipStream_t st) {
  if (a.nrows <= 0) return;
  const int tpb = kWave * (a.waves_bwd > 0 ? a.waves_bwd : 1);
  DPMI_KQ_DISPATCH(a.M, hipLaunchKernelGGL((k_embed_bwd<T, TG, KQ>), dim3(ceil_div(a.nrows * kWave, tpb)), dim3(tpb), 0, st,
                                           (T*)a.net_deriv, (const T*)a.xyz, (const TG*)a.gD, a.pos,
                                           (const T*)a.tables, a.tab_off, (const Info<T>*)a.infos, a.net_of,
                                           (const T*)a.em, a.row_type, a.sec, a.ntypes, a.nrows, a.nnei, a.M, a.A,
                                           a.ldD > 0 ? a.ldD : a.M * a.A, (T)a.inv_nnei));
  DPMI_LAUNCH_CHECK();
}

template void launch_embed_fwd<double, double>(const EmbedArgs&, hipStream_t);
template void launch_embed_bwd<double, double>(const EmbedArgs&, hipStream_t);
template void launch_embed_fwd<float, _Float16>(const EmbedArgs&, hipStream_t);
template void launch_embed_bwd<float, _Float16>(const EmbedArgs&, hipStream_t);
template void launch_embed_bwd<float, float>(const EmbedArgs&, hipStream_t);

}  // namespace dpmi
