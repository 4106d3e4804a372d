// Environment matrix of se_e2_a (gfx950): one thread per (centre, slot).
//
// For slot s of centre i with neighbour j = nlist[i, s]:
//   r = x_j - x_i,  value = s(|r|) * [1/|r|, r_x/|r|^2, r_y/|r|^2, r_z/|r|^2]
//   deriv[a][d] = d(value_a)/d(x_i)_d            (reference env_mat.cc:124-222)
//   em = (value - avg[t_i]) / std[t_i], em_deriv = deriv / std[t_i]
//   padding slots: em = -avg/std, deriv = 0, rij = 0; centre type < 0: zeros
//   (prod_env_mat.cc:107-123).
// Consecutive threads own consecutive slots, so the 4-, 12- and 3-wide rows
// are written as contiguous, coalesced spans.
#include "dpmi_common.h"
#include "env_core.h"
#include "kernels.h"

namespace dpmi {

namespace {

// TO: storage type of em / em_deriv / rij (double for the FP64 engine,
// float for the mixed engine: the image displacement, the switch and every
// intermediate are evaluated in FP64 and only the stored values are rounded).
// DERIV false: em only (the mixed engine's force kernel recomputes the
// derivative from the coordinates).
template <typename TO, bool DERIV, typename TC = double>
__global__ void k_env_mat_a(const TC* __restrict__ coord, const int* __restrict__ type,
                            const int* __restrict__ nlist, const int* __restrict__ row_frame,
                            const int* __restrict__ row_ext_off, const int* __restrict__ row_center,
                            int strided_nloc, long long nrows, int nnei, int nall_stride, int ntypes,
                            const double* __restrict__ avg,
                            const double* __restrict__ stdv, double rmin, double rmax, TO* __restrict__ em,
                            TO* __restrict__ em_deriv, TO* __restrict__ rij) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nrows * nnei) return;
  const long long g = tid / nnei;
  const int s = (int)(tid - g * nnei);
  // row -> (frame extended offset, centre index in the frame)
  size_t eo;
  int i;
  if (row_ext_off) {
    eo = (size_t)row_ext_off[row_frame[g]];
    i = row_center[g];
  } else {
    eo = (size_t)(g / strided_nloc) * nall_stride;
    i = (int)(g % strided_nloc);
  }
  // a centre type outside [0, ntypes) has no avg/std row: its row is zero
  // (the neighbour-list stage reports such types)
  int ti = type[eo + i];
  if (ti >= ntypes) ti = -1;
  const int j = nlist[g * nnei + s];
  double val[4] = {0., 0., 0., 0.};
  double der[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) der[k] = 0.;
  double rr[3] = {0., 0., 0.};
  if (ti >= 0 && j >= 0) {
    for (int d = 0; d < 3; ++d) rr[d] = (double)coord[3 * (eo + j) + d] - (double)coord[3 * (eo + i) + d];
    envc::slot_terms<sizeof(TO) == 4, DERIV>(rr, rmin, rmax, val, der);
  }
  const size_t base = (size_t)g * nnei + s;
  if (ti >= 0) {
    const double* av = avg + (size_t)ti * nnei * 4 + s * 4;
    const double* sd = stdv + (size_t)ti * nnei * 4 + s * 4;
    envc::em_row<TO>(val, av, sd, em + base * 4);
    if constexpr (DERIV) {
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        if constexpr (sizeof(TO) == 4) {
          const double isd = 1. / sd[a];
#pragma unroll
          for (int d = 0; d < 3; ++d) em_deriv[base * 12 + a * 3 + d] = (TO)(der[a * 3 + d] * isd);
        } else {
#pragma unroll
          for (int d = 0; d < 3; ++d) em_deriv[base * 12 + a * 3 + d] = (TO)(der[a * 3 + d] / sd[a]);
        }
      }
    }
  } else {
#pragma unroll
    for (int a = 0; a < 4; ++a) em[base * 4 + a] = (TO)0;
    if constexpr (DERIV)
#pragma unroll
      for (int k = 0; k < 12; ++k) em_deriv[base * 12 + k] = (TO)0;
  }
  if constexpr (DERIV)
#pragma unroll
    for (int d = 0; d < 3; ++d) rij[base * 3 + d] = (TO)rr[d];
}

}  // namespace

void launch_env_mat_a(const EnvArgs& a, hipStream_t st) {
  const long long n = a.nrows * (long long)a.nnei;
  if (n <= 0) return;
  const int tpb = a.tpb > 0 ? a.tpb : 256;
  if (a.em_deriv)
    hipLaunchKernelGGL((k_env_mat_a<double, true>), dim3(ceil_div(n, tpb)), dim3(tpb), 0, st, a.coord, a.type, a.nlist, a.row_frame,
                       a.row_ext_off, a.row_center, a.strided_nloc, a.nrows, a.nnei, a.nall_stride, a.ntypes, a.avg, a.std_, a.rmin, a.rmax,
                       a.em, a.em_deriv, a.rij);
  else
    hipLaunchKernelGGL((k_env_mat_a<double, false>), dim3(ceil_div(n, tpb)), dim3(tpb), 0, st, a.coord, a.type, a.nlist, a.row_frame,
                       a.row_ext_off, a.row_center, a.strided_nloc, a.nrows, a.nnei, a.nall_stride, a.ntypes, a.avg, a.std_, a.rmin, a.rmax,
                       a.em, a.em_deriv, a.rij);
  DPMI_LAUNCH_CHECK();
}

void launch_env_mat_a_f32(const EnvArgs& a, float* em, float* em_deriv, float* rij, hipStream_t st) {
  const long long n = a.nrows * (long long)a.nnei;
  if (n <= 0) return;
  const int tpb = a.tpb > 0 ? a.tpb : 256;
  if (a.coord_f) {  // FP32 coordinates (the float instantiation of prod_env_mat_a)
    if (em_deriv)
      hipLaunchKernelGGL((k_env_mat_a<float, true, float>), dim3(ceil_div(n, tpb)), dim3(tpb), 0, st, a.coord_f,
                         a.type, a.nlist, a.row_frame, a.row_ext_off, a.row_center, a.strided_nloc, a.nrows, a.nnei,
                         a.nall_stride, a.ntypes, a.avg, a.std_, a.rmin, a.rmax, em, em_deriv, rij);
    else
      hipLaunchKernelGGL((k_env_mat_a<float, false, float>), dim3(ceil_div(n, tpb)), dim3(tpb), 0, st, a.coord_f,
                         a.type, a.nlist, a.row_frame, a.row_ext_off, a.row_center, a.strided_nloc, a.nrows, a.nnei,
                         a.nall_stride, a.ntypes, a.avg, a.std_, a.rmin, a.rmax, em, em_deriv, rij);
    DPMI_LAUNCH_CHECK();
    return;
  }
  if (em_deriv)
    hipLaunchKernelGGL((k_env_mat_a<float, true>), dim3(ceil_div(n, tpb)), dim3(tpb), 0, st, a.coord, a.type, a.nlist,
                       a.row_frame, a.row_ext_off, a.row_center, a.strided_nloc, a.nrows, a.nnei, a.nall_stride, a.ntypes, a.avg,
                       a.std_, a.rmin, a.rmax, em, em_deriv, rij);
  else
    hipLaunchKernelGGL((k_env_mat_a<float, false>), dim3(ceil_div(n, tpb)), dim3(tpb), 0, st, a.coord, a.type,
                       a.nlist, a.row_frame, a.row_ext_off, a.row_center, a.strided_nloc, a.nrows, a.nnei,
                       a.nall_stride, a.ntypes, a.avg, a.std_, a.rmin, a.rmax, em, em_deriv, rij);
  DPMI_LAUNCH_CHECK();
}

}  // namespace dpmi
