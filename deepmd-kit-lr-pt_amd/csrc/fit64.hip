// FP64 fitting layers of the plain engine on the matrix cores (gfx950,
// v_mfma_f64_16x16x4f64), after deepmd/pt/model/network/mlp.py:188-219 and
// deepmd/pt/model/task/fitting.py (energy = last layer + bias_atom_e):
//   forward:   z = x W, y = act(z + b), x_next = y * idt + residual(x)
//   backward:  g_in = (g * idt * act'(y)) W^T + residual(g)
//   output:    e = x w_out + b_out + bias_atom_e
// One launch per layer and direction covers every atom type.  Every wave
// is independent (no LDS, no barriers): it owns WM row blocks x WN column
// blocks of 16 of one type and walks K in chunks of 16.  The weights are
// stored pre-swizzled into MFMA B fragments (fit64_swizzle): fragment
// (chunk c, column block cb, step s) is 64 contiguous doubles, lane j + 16 q
// holding W[16 c + 4 q + s][16 cb + j], so a wave reads each fragment with
// one coalesced 512-byte load and the waves of a CU share them through L1.
// Lane i + 16 q reads the 4 consecutive A values A[row i][16 c + 4 q ..+3]
// of a chunk (two 16-byte loads) and feeds them to the 4 steps: the
// reduction runs over k in the order (c, s, q) instead of k ascending.
// Chunks are double buffered in registers (the next chunk's loads are in
// flight during the current chunk's products).  Bias, tanh, timestep and
// residual terms are applied to the accumulators before the store, and so is
// the next product's operand: the forward of the top hidden layer also
// stores its dE/dz (from w_out), and each backward layer stores dE/dz of the
// layer below (dE/dx * idt * act'(y)), so every backward product reads a
// plain operand.
//
// Operand layout of v_mfma_f64_16x16x4f64 (checked on the device): A[i][k]
// in lane i + 16 k, B[k][j] in lane j + 16 k, D[i][j] in lane j + 16 (i % 4),
// element i / 4.
#include "dpmi_common.h"
#include "kernels.h"

namespace dpmi {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

// A values of one chunk for WM row blocks, and B fragments for WN blocks x 4 steps
template <int WM, int WN>
struct Chunk {
  double a[WM][4];
  double b[WN][4];
};

template <bool BWD, int WM, int WN, bool KFULL>
__global__ __launch_bounds__(256) void k_fit64(const Fit64Layer p) {
  const int K = p.K, N = p.N;
  const int ncb = (N + 15) / 16, ncg = (ncb + WN - 1) / WN, nch = (K + 15) / 16;
  const int lane = threadIdx.x & 63;
  long long tile = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  int t = 0;
  for (; t < p.ntypes; ++t) {
    const long long nt = (p.rows[t] + 16 * WM - 1) / (16 * WM) * ncg;
    if (tile < nt) break;
    tile -= nt;
  }
  if (t >= p.ntypes) return;
  const long long nrows = p.rows[t], base = p.p0[t];
  const long long r0 = tile / ncg * (16 * WM);
  const int cb0 = (int)(tile % ncg) * WN;
  const int i = lane & 15, q = lane >> 4;

  // this lane's A rows (clamped into the type's rows; extra rows are dropped
  // at the store) and the fragment base of each column block (clamped to the
  // last block)
  const double* arow[WM];
#pragma unroll
  for (int m = 0; m < WM; ++m) {
    const long long r = r0 + 16 * m + i;
    arow[m] = p.a + (base + (r < nrows ? r : nrows - 1)) * (long long)p.lda;
  }
  const double* bcol[WN];
#pragma unroll
  for (int n = 0; n < WN; ++n) {
    const int cb = cb0 + n < ncb ? cb0 + n : ncb - 1;
    bcol[n] = p.w[t] + (size_t)cb * 256 + lane;
  }
  const size_t bchunk = (size_t)ncb * 256;

  auto load = [&](Chunk<WM, WN>& ch, int c) {
    const int k0 = 16 * c + 4 * q;
#pragma unroll
    for (int m = 0; m < WM; ++m) {
      if (KFULL) {
        const d2 lo = *reinterpret_cast<const d2*>(arow[m] + k0);
        const d2 hi = *reinterpret_cast<const d2*>(arow[m] + k0 + 2);
        ch.a[m][0] = lo.x;
        ch.a[m][1] = lo.y;
        ch.a[m][2] = hi.x;
        ch.a[m][3] = hi.y;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double v = arow[m][k0 + e < K ? k0 + e : K - 1];
          ch.a[m][e] = k0 + e < K ? v : 0.0;
        }
      }
    }
#pragma unroll
    for (int n = 0; n < WN; ++n)
#pragma unroll
      for (int s = 0; s < 4; ++s) ch.b[n][s] = bcol[n][(size_t)c * bchunk + s * 64];
  };

  d4 acc[WM][WN];
#pragma unroll
  for (int m = 0; m < WM; ++m)
#pragma unroll
    for (int n = 0; n < WN; ++n) acc[m][n] = d4{0.0, 0.0, 0.0, 0.0};
  auto mma = [&](const Chunk<WM, WN>& ch) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int m = 0; m < WM; ++m)
#pragma unroll
        for (int n = 0; n < WN; ++n)
          acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(ch.a[m][s], ch.b[n][s], acc[m][n], 0, 0, 0);
  };

  // two register buffers and no branch inside the loop (a branch lets the
  // compiler sink a prefetch to its use); scheduling barriers keep each
  // prefetch ahead of the products before it.  The last pair re-reads the
  // last chunk (in bounds) when nch is even; an odd tail runs after the loop.
  Chunk<WM, WN> x, y;
  load(x, 0);
  int c = 0;
  for (; c + 1 < nch; c += 2) {
    load(y, c + 1);
    __builtin_amdgcn_sched_barrier(0);
    mma(x);
    __builtin_amdgcn_sched_barrier(0);
    load(x, c + 2 < nch ? c + 2 : nch - 1);
    __builtin_amdgcn_sched_barrier(0);
    mma(y);
    __builtin_amdgcn_sched_barrier(0);
  }
  if (c < nch) mma(x);

  // epilogue: element v of lane l is row 4 v + l / 16, column l % 16
  const double* __restrict__ idt = p.idt[t];
  const double* __restrict__ wout = p.wout[t];
#pragma unroll
  for (int m = 0; m < WM; ++m)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const long long row = r0 + 16 * m + 4 * v + q;
      if (row >= nrows) continue;
      const long long gr = base + row;
#pragma unroll
      for (int n = 0; n < WN; ++n) {
        const int col = (cb0 + n) * 16 + i;
        if (col >= N) continue;
        const double z = acc[m][n][v];
        const size_t o = (size_t)gr * p.ldo + col;
        if (!BWD) {
          double yv = z + p.bias[t][col];
          if (p.act) yv = tanh(yv);
          if (p.out0) p.out0[o] = yv;
          double xo = idt ? yv * idt[col] : yv;
          const double* x0 = p.a + gr * (long long)p.lda;
          if (p.res_mode == 1)
            xo += x0[col];
          else if (p.res_mode == 2)
            xo += x0[col < K ? col : col - K];
          p.out1[o] = xo;
          if (p.out2) {  // top hidden layer: dE/dz = w_out * idt * act'
            double gz = wout[col];
            if (idt) gz *= idt[col];
            if (p.act) gz *= (1. - yv * yv);
            p.out2[o] = gz;
          }
        } else {
          double gi = z;
          if (p.res_mode) {
            const double* g = wout ? wout : p.g + gr * (long long)p.ldg;
            gi += g[col];
            if (p.res_mode == 2) gi += g[col + N];
          }
          if (!p.out1) {
            p.out0[o] = gi;
          } else {  // dE/dz of the layer below
            if (p.out0) p.out0[o] = gi;
            const double* idn = p.idt_next[t];
            double gz = idn ? gi * idn[col] : gi;
            if (p.act_next) {
              const double yv = p.y[gr * (long long)p.ldy + col];
              gz *= (1. - yv * yv);
            }
            p.out1[o] = gz;
          }
        }
      }
    }
}

// e[perm[p0 + r]] = x[r] . w + (b_out + bias): one wave per row
__global__ __launch_bounds__(256) void k_fit64_out(const Fit64Out p) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int bid = blockIdx.x, t = 0;
  for (; t < p.ntypes; ++t) {
    const int nb = (int)((p.rows[t] + 3) / 4);
    if (bid < nb) break;
    bid -= nb;
  }
  if (t >= p.ntypes) return;
  const long long r = (long long)bid * 4 + wave;
  if (r >= p.rows[t]) return;
  const double* __restrict__ x = p.x + (p.p0[t] + r) * (long long)p.K;
  const double* __restrict__ w = p.w[t];
  double s = 0.0;
  for (int j = lane; j < p.K; j += 64) s = fma(x[j], w[j], s);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) p.atom_e[p.perm[p.p0[t] + r]] = s + p.b_out[t] + p.bias[t];
}

}  // namespace

namespace {
template <bool BWD, int WM, int WN>
void fit64_launch(const Fit64Layer& p, hipStream_t st) {
  const int ncb = (p.N + 15) / 16, ncg = (ncb + WN - 1) / WN;
  long long tiles = 0;
  for (int t = 0; t < p.ntypes; ++t) tiles += (p.rows[t] + 16 * WM - 1) / (16 * WM) * ncg;
  if (tiles <= 0) return;
  const long long blocks = (tiles + 3) / 4;
  if (blocks > (1LL << 31) - 1) throw HipError("fit64: too many tiles");
  if (p.K % 16 == 0 && p.lda % 2 == 0)
    hipLaunchKernelGGL((k_fit64<BWD, WM, WN, true>), dim3((unsigned)blocks), dim3(256), 0, st, p);
  else
    hipLaunchKernelGGL((k_fit64<BWD, WM, WN, false>), dim3((unsigned)blocks), dim3(256), 0, st, p);
  DPMI_LAUNCH_CHECK();
}
}  // namespace

// Two row blocks x 3 column blocks per wave (a 240-wide layer splits into 5
// column groups, so 12,288 rows make 1,920 waves: close to a whole number of
// waves per SIMD).
void launch_fit64(const Fit64Layer& p, bool backward, hipStream_t st) {
  if (p.ntypes <= 0 || p.ntypes > kFit64MaxTypes || p.K <= 0 || p.N <= 0)
    throw HipError("fit64: layer shape out of range");
  if (backward)
    fit64_launch<true, 2, 3>(p, st);
  else
    fit64_launch<false, 2, 3>(p, st);
}

std::vector<double> fit64_swizzle(const double* w, int K, int N) {
  const int nc = (K + 15) / 16, ncb = (N + 15) / 16;
  std::vector<double> out((size_t)nc * ncb * 256, 0.0);
  for (int c = 0; c < nc; ++c)
    for (int cb = 0; cb < ncb; ++cb)
      for (int s = 0; s < 4; ++s)
        for (int l = 0; l < 64; ++l) {
          const int k = 16 * c + 4 * (l >> 4) + s, n = 16 * cb + (l & 15);
          if (k < K && n < N) out[(((size_t)c * ncb + cb) * 4 + s) * 64 + l] = w[(size_t)k * N + n];
        }
  return out;
}

void launch_fit64_out(const Fit64Out& p, hipStream_t st) {
  if (p.ntypes <= 0 || p.ntypes > kFit64MaxTypes || p.K <= 0) throw HipError("fit64: output layer out of range");
  long long blocks = 0;
  for (int t = 0; t < p.ntypes; ++t) blocks += (p.rows[t] + 3) / 4;
  if (blocks <= 0) return;
  hipLaunchKernelGGL(k_fit64_out, dim3((unsigned)blocks), dim3(256), 0, st, p);
  DPMI_LAUNCH_CHECK();
}

}  // namespace dpmi
