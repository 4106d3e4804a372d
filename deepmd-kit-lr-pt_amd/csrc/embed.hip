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
// The adjoint loop uses a transposed lane layout: 16 neighbours x 4 column
// groups per wave, each lane walking every fourth column of its neighbour's
// table row, so the reduction over columns is one 2-step butterfly per 16
// neighbours instead of a 6-step butterfly per neighbour.
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

// Adjoint of one neighbour block (nb slots) of row g.  dyk: [M][4] in LDS.
template <typename T>
__device__ void embed_bwd_block(T* __restrict__ nd, const T* dyk, const T* __restrict__ table, const Info<T>& f,
                                const T* __restrict__ em, int nb, int M) {
  if (nb <= 0) return;
  const int lane = lane_id();
  const int jl = lane & 15, kg = lane >> 4;
  const T last = em[(size_t)(nb - 1) * 4];
  // first neighbour of the sorted zero-padding tail (tabulate.cc:188-201)
  int jt = nb;
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
  for (int j0 = 0; j0 < jend; j0 += 16) {
    const int j = j0 + jl;
    const bool valid = j < jend;
    T ll[4] = {0, 0, 0, 0};
    if (valid)
#pragma unroll
      for (int c = 0; c < 4; ++c) ll[c] = em[(size_t)j * 4 + c];
    T x = ll[0], delta = 0;
    int idx = 0;
    if (valid) locate(f, x, idx, delta);
    const T* row = table + (size_t)idx * M * 6;
    T px = 0, pm[4] = {0, 0, 0, 0};
    for (int k = kg; k < M; k += 4) {
      T a[6];
      load6(row + 6 * k, a);
      T v, d;
      quintic(a, x, delta, v, d);
      const T* dy = dyk + 4 * k;
      const T proj = ll[0] * dy[0] + ll[1] * dy[1] + ll[2] * dy[2] + ll[3] * dy[3];
      px += d * proj;
#pragma unroll
      for (int c = 0; c < 4; ++c) pm[c] += v * dy[c];
    }
    px += __shfl_xor(px, 16, kWave);
    px += __shfl_xor(px, 32, kWave);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      pm[c] += __shfl_xor(pm[c], 16, kWave);
      pm[c] += __shfl_xor(pm[c], 32, kWave);
    }
    if (kg == 0 && valid) {
      const T scale = (j == jt) ? (T)(nb - jt) : (T)1;
      nd[(size_t)j * 4 + 0] = (pm[0] + px) * scale;
#pragma unroll
      for (int c = 1; c < 4; ++c) nd[(size_t)j * 4 + c] = pm[c] * scale;
    }
  }
  for (int j = jend + lane; j < nb; j += kWave)
#pragma unroll
    for (int c = 0; c < 4; ++c) nd[(size_t)j * 4 + c] = 0;
}

template <typename T, typename TG, int KQ>
__global__ void __launch_bounds__(64) k_embed_bwd(T* __restrict__ net_deriv, const T* __restrict__ xyz,
                                                  const TG* __restrict__ gD, const int* __restrict__ pos,
                                                  const T* __restrict__ tables, const long long* __restrict__ tab_off,
                                                  const Info<T>* __restrict__ infos, const int* __restrict__ net_of,
                                                  const T* __restrict__ em, const int* __restrict__ row_type,
                                                  const int* __restrict__ sec, int ntypes, long long nrows, int nnei,
                                                  int M, int A, int ldD, T inv_nnei) {
  __shared__ T dyk[KQ * kWave * 4];
  const long long g = blockIdx.x;
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
  const TG* G = gD + (size_t)pos[g] * ldD;
  // sum_a G[k, a] T[c, a]
  T dT[KQ][4];
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq)
#pragma unroll
    for (int c = 0; c < 4; ++c) dT[kq][c] = 0;
  for (int a = 0; a < A; ++a) {
    T ta[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) ta[c] = bcast(Tk[0][c], a);
#pragma unroll
    for (int kq = 0; kq < KQ; ++kq) {
      const int k = lane + kq * kWave;
      if (k < M) {
        const T gk = (T)G[(size_t)k * A + a];
#pragma unroll
        for (int c = 0; c < 4; ++c) dT[kq][c] += gk * ta[c];
      }
    }
  }
  // S[c, a] = sum_m G[m, a] T[c, m] on lane c*A + a (A <= 16)
  T S = 0;
  {
    const int c = lane / A, a = lane - (lane / A) * A;
    if (c < 4) {
      const T* xr = xyz + ((size_t)g * 4 + c) * M;
      for (int m = 0; m < M; ++m) S += (T)G[(size_t)m * A + a] * xr[m];
      S *= inv_nnei;
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const T s = __shfl(S, c * A + (lane < A ? lane : 0), kWave);
    if (lane < A) dT[0][c] += s;
  }
#pragma unroll
  for (int kq = 0; kq < KQ; ++kq) {
    const int k = lane + kq * kWave;
#pragma unroll
    for (int c = 0; c < 4; ++c) dyk[k * 4 + c] = dT[kq][c] * inv_nnei;
  }
  __syncthreads();
  for (int tj = 0; tj < ntypes; ++tj) {
    const int net = net_of[ti * ntypes + tj];
    const int b = sec[tj], nb = sec[tj + 1] - sec[tj];
    embed_bwd_block<T>(net_deriv + ((size_t)g * nnei + b) * 4, dyk, tables + tab_off[net], infos[net],
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
  const dim3 grid(ceil_div(a.nrows * kWave, 256));
  DPMI_KQ_DISPATCH(a.M, hipLaunchKernelGGL((k_embed_fwd<T, TD, KQ>), grid, dim3(256), 0, st, (T*)a.xyz, (TD*)a.D,
                                           a.pos, (const T*)a.tables, a.tab_off, (const Info<T>*)a.infos, a.net_of,
                                           (const T*)a.em, a.row_type, a.sec, a.ntypes, a.nrows, a.nnei, a.M, a.A,
                                           a.ldD > 0 ? a.ldD : a.M * a.A, (T)a.inv_nnei));
  DPMI_LAUNCH_CHECK();
}

template <typename T, typename TG>
void launch_embed_bwd(const EmbedArgs& a, hipStream_t st) {
  if (a.nrows <= 0) return;
  DPMI_KQ_DISPATCH(a.M, hipLaunchKernelGGL((k_embed_bwd<T, TG, KQ>), dim3(a.nrows), dim3(kWave), 0, st,
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
