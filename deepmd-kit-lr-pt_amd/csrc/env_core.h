// Per-slot terms of the se_e2_a environment matrix, shared by the
// environment kernel and the force kernel that recomputes the derivative.
#pragma once

#include <hip/hip_runtime.h>

namespace dpmi {
namespace envc {

template <typename T>
__device__ __forceinline__ void smooth_switch(T& v, T& dv, T r, T rmin, T rmax) {
  // switcher.h:61-80
  if (r < rmin) {
    v = 1;
    dv = 0;
  } else if (r < rmax) {
    const T u = (r - rmin) / (rmax - rmin);
    const T du = (T)1 / (rmax - rmin);
    v = u * u * u * ((T)-6 * u * u + (T)15 * u - (T)10) + (T)1;
    dv = ((T)3 * u * u * ((T)-6 * u * u + (T)15 * u - (T)10) + u * u * u * ((T)-12 * u + (T)15)) * du;
  } else {
    v = 0;
    dv = 0;
  }
}

// For the displacement rr = x_j - x_i: val = s(|r|) [1/|r|, r/|r|^2] and, when
// DERIV, der[a][d] = d(val_a)/d(x_i)_d (env_mat.cc:124-222).  FP32: the
// values are stored rounded to FP32, so reciprocal products replace the
// quotients (they differ by about one FP64 ulp).
template <bool FP32, bool DERIV>
__device__ __forceinline__ void slot_terms(const double (&rr)[3], double rmin, double rmax, double (&val)[4],
                                           double (&der)[12]) {
  const double nr2 = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
  const double inr = 1. / sqrt(nr2);
  const double nr = nr2 * inr;
  const double inr2 = inr * inr;
  const double inr4 = inr2 * inr2;
  const double inr3 = inr4 * nr;
  double sw, dsw;
  smooth_switch(sw, dsw, nr, rmin, rmax);
  if constexpr (FP32) {
    val[0] = inr;
    val[1] = rr[0] * inr2;
    val[2] = rr[1] * inr2;
    val[3] = rr[2] * inr2;
  } else {
    val[0] = 1. / nr;
    val[1] = rr[0] / nr2;
    val[2] = rr[1] / nr2;
    val[3] = rr[2] / nr2;
  }
  if constexpr (DERIV) {
#pragma unroll
    for (int d = 0; d < 3; ++d) der[d] = rr[d] * inr3 * sw - val[0] * dsw * rr[d] * inr;
#pragma unroll
    for (int a = 1; a < 4; ++a) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double t = 2. * rr[a - 1] * rr[d] * inr4;
        if (d == a - 1) t -= inr2;
        der[a * 3 + d] = t * sw - val[a] * dsw * rr[d] * inr;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) val[a] *= sw;
}

// Stored environment-matrix row of a slot: (val - avg) / std (FP32
// storage: times the FP64 reciprocal of std, rounded once).
template <typename TO>
__device__ __forceinline__ void em_row(const double (&val)[4], const double* __restrict__ av,
                                       const double* __restrict__ sd, TO* __restrict__ out) {
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if constexpr (sizeof(TO) == 4)
      out[a] = (TO)((val[a] - av[a]) * (1. / sd[a]));
    else
      out[a] = (TO)((val[a] - av[a]) / sd[a]);
  }
}

// Environment-matrix row of slot s of centre i (type ti, coordinates ci)
// with neighbour j (< 0: padding), as the environment kernel stores it
// without derivatives (prod_env_mat.cc:107-123): padding slots give
// -avg/std, a centre of negative type zeros.
template <typename TO>
__device__ __forceinline__ void slot_em(const double (&ci)[3], const double* __restrict__ cj, int j, int ti, int s,
                                        int nnei, const double* __restrict__ avg, const double* __restrict__ stdv,
                                        double rmin, double rmax, TO* __restrict__ out) {
  if (ti < 0) {
#pragma unroll
    for (int a = 0; a < 4; ++a) out[a] = (TO)0;
    return;
  }
  double val[4] = {0., 0., 0., 0.}, der[12];
  if (j >= 0) {
    double rr[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) rr[d] = cj[d] - ci[d];
    slot_terms<sizeof(TO) == 4, false>(rr, rmin, rmax, val, der);
  }
  const size_t o = (size_t)ti * nnei * 4 + (size_t)s * 4;
  em_row<TO>(val, avg + o, stdv + o, out);
}

}  // namespace envc
}  // namespace dpmi
