// Fitting net of the mixed-precision engine on FP16 MFMA (gfx950).
//
// One GEMM kernel with fused prologue/epilogue covers every fitting stage
// (deepmd/pt/model/task/fitting.py:1079-1104, network/mlp.py:188-219):
//   forward layer:  z = x W (FP16 x FP16 -> FP32),  y = tanh(z + b),
//                   x' = y * idt + x (equal widths) -- y kept in FP32 for the
//                   backward, x' stored in FP16 as the next GEMM input; the
//                   last hidden layer instead accumulates e += x' . w_out
//                   from the FP32 x' (its output feeds nothing else)
//   backward layer: gz = g * idt * (1 - y^2) formed while loading the A tile
//                   (g from memory, or w_out broadcast for the last hidden
//                   layer), g_in = gz W^T + g (residual), stored in FP32.
// Tiles: 64 x 64 outputs per 256-thread workgroup, 4 waves of 32 x 32, each
// 2 x 2 x 2 v_mfma_f32_16x16x32_f16 per 32-wide K step (weights split into
// FP16 hi + lo parts, both multiplied); A and B tiles are staged through LDS
// with padded rows.  Widths are padded to multiples of 64 with
// zero weights (padded activations stay exactly zero).
#include <hip/hip_fp16.h>

#include "dpmi_common.h"
#include "kernels.h"

namespace dpmi {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32, LDS_PAD = 8;

enum { kAHalf = 0, kAGradMat = 1, kAGradBcast = 2 };
enum { kEFwd = 0, kEBwd = 1 };

struct GemmArgs {
  const _Float16* A;  // kAHalf: [rows][lda]
  int lda;
  const float* g;     // kAGradMat: [rows][ldg]; kAGradBcast: [K]
  int ldg;
  const float* y;     // [rows][ldy] (kAGrad*)
  int ldy;
  const float* a_idt; // [K] or null (kAGrad*)
  int a_act;          // kAGrad*: 1 = tanh (factor 1 - y^2), 0 = linear
  const _Float16* Bt; // [N][ldb] (hi part)
  const _Float16* Btl; // [N][ldb] (lo part)
  int ldb;
  int rows, N, K;
  // forward epilogue
  const float* bias;  // [N]
  const float* idt;   // [N] or null
  int act;
  const _Float16* xres;  // [rows][ldr] residual input (half) or null
  int ldr;
  float* y_out;       // [rows][ldyo]
  int ldyo;
  _Float16* x_out;    // [rows][ldxo]
  int ldxo;
  const float* out_w; // [N] or null: e[row] += sum_n x_out * out_w
  float* e_acc;       // [rows]
  // backward epilogue
  const float* gres;  // [rows][ldgr] residual gradient, or (gres_bcast) [N]
  int ldgr;
  int gres_bcast;
  float* g_out;       // [rows][ldgo]
  int ldgo;
};

template <int AMODE, int EMODE>
__global__ void __launch_bounds__(256) k_fit_gemm(GemmArgs p) {
  __shared__ _Float16 As[BM][BK + LDS_PAD];
  __shared__ _Float16 Bs[BN][BK + LDS_PAD];
  __shared__ _Float16 Bsl[BN][BK + LDS_PAD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int row0 = blockIdx.y * BM, col0 = blockIdx.x * BN;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  const int lr = tid >> 2, lk = (tid & 3) * 8;  // tile loader: row, k chunk
  for (int k0 = 0; k0 < p.K; k0 += BK) {
    // ---- A tile
    {
      const int r = row0 + lr;
      half8 v;
      if (r < p.rows) {
        if constexpr (AMODE == kAHalf) {
          v = *reinterpret_cast<const half8*>(p.A + (size_t)r * p.lda + k0 + lk);
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int k = k0 + lk + e;
            float g = (AMODE == kAGradMat) ? p.g[(size_t)r * p.ldg + k] : p.g[k];
            if (p.a_idt) g *= p.a_idt[k];
            if (p.a_act) {
              const float yy = p.y[(size_t)r * p.ldy + k];
              g *= 1.f - yy * yy;
            }
            v[e] = (_Float16)g;
          }
        }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = (_Float16)0.f;
      }
      *reinterpret_cast<half8*>(&As[lr][lk]) = v;
    }
    // ---- B tile (rows of Bt are output columns)
    {
      const int n = col0 + lr;
      const half8 v = *reinterpret_cast<const half8*>(p.Bt + (size_t)n * p.ldb + k0 + lk);
      const half8 vl = *reinterpret_cast<const half8*>(p.Btl + (size_t)n * p.ldb + k0 + lk);
      *reinterpret_cast<half8*>(&Bs[lr][lk]) = v;
      *reinterpret_cast<half8*>(&Bsl[lr][lk]) = vl;
    }
    __syncthreads();
    half8 af[2], bf[2], bl[2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
      af[i] = *reinterpret_cast<const half8*>(&As[wr * 32 + i * 16 + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      bf[j] = *reinterpret_cast<const half8*>(&Bs[wc * 32 + j * 16 + (lane & 15)][(lane >> 4) * 8]);
      bl[j] = *reinterpret_cast<const half8*>(&Bsl[wc * 32 + j * 16 + (lane & 15)][(lane >> 4) * 8]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(af[i], bf[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(af[i], bl[j], acc[i][j], 0, 0, 0);
      }
    __syncthreads();
  }

  // ---- epilogue: lane holds rows (lane>>4)*4 + e, column lane & 15 of each tile
  float epart[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) epart[i][e] = 0.f;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = col0 + wc * 32 + j * 16 + (lane & 15);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = row0 + wr * 32 + i * 16 + (lane >> 4) * 4 + e;
        if (r >= p.rows) continue;
        const float z = acc[i][j][e];
        if constexpr (EMODE == kEFwd) {
          float v = z + p.bias[n];
          if (p.act) v = tanhf(v);
          p.y_out[(size_t)r * p.ldyo + n] = v;
          float o = p.idt ? v * p.idt[n] : v;
          if (p.xres) o += (float)p.xres[(size_t)r * p.ldr + n];
          if (p.out_w)
            epart[i][e] += o * p.out_w[n];  // the last hidden output feeds only the energy
          else
            p.x_out[(size_t)r * p.ldxo + n] = (_Float16)o;
        } else {
          float o = z;
          if (p.gres) o += p.gres_bcast ? p.gres[n] : p.gres[(size_t)r * p.ldgr + n];
          p.g_out[(size_t)r * p.ldgo + n] = o;
        }
      }
    }
  }
  if constexpr (EMODE == kEFwd) {
    if (p.out_w) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float s = epart[i][e];
          s += __shfl_xor(s, 1, 64);
          s += __shfl_xor(s, 2, 64);
          s += __shfl_xor(s, 4, 64);
          s += __shfl_xor(s, 8, 64);
          const int r = row0 + wr * 32 + i * 16 + (lane >> 4) * 4 + e;
          if ((lane & 15) == 0 && r < p.rows) atomicAdd(p.e_acc + r, s);
        }
    }
  }
}

__global__ void k_energy_out(const float* __restrict__ e_perm, const int* __restrict__ pos,
                             const int* __restrict__ row_type, const double* __restrict__ shift, long long n,
                             double* __restrict__ atom_e) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int t = row_type[g];
  atom_e[g] = t >= 0 ? (double)e_perm[pos[g]] + shift[t] : 0.0;
}

}  // namespace

void launch_fit16_fwd(const Fit16Layer& L, const _Float16* x, int ldx, int rows, const _Float16* xres, int ldr,
                      float* y, int ldy, _Float16* xo, int ldxo, float* e_acc, hipStream_t st) {
  if (rows <= 0) return;
  GemmArgs p{};
  p.A = x;
  p.lda = ldx;
  p.Bt = L.wt;
  p.Btl = L.wt_lo;
  p.ldb = L.k_pad;
  p.rows = rows;
  p.N = L.n_pad;
  p.K = L.k_pad;
  p.bias = L.b;
  p.idt = L.idt;
  p.act = L.act;
  p.xres = xres;
  p.ldr = ldr;
  p.y_out = y;
  p.ldyo = ldy;
  p.x_out = xo;
  p.ldxo = ldxo;
  p.out_w = e_acc ? L.out_w : nullptr;
  p.e_acc = e_acc;
  const dim3 grid(L.n_pad / BN, ceil_div(rows, BM));
  hipLaunchKernelGGL((k_fit_gemm<kAHalf, kEFwd>), grid, dim3(256), 0, st, p);
  DPMI_LAUNCH_CHECK();
}

void launch_fit16_bwd(const Fit16Layer& L, const float* g, int ldg, bool g_bcast, const float* y, int ldy, int rows,
                      bool residual, float* g_in, int ld_in, hipStream_t st) {
  if (rows <= 0) return;
  GemmArgs p{};
  p.g = g;
  p.ldg = ldg;
  p.y = y;
  p.ldy = ldy;
  p.a_idt = L.idt;
  p.a_act = L.act;
  p.Bt = L.w;  // [n_in_pad][n_out_pad]: rows are the outputs of the backward GEMM
  p.Btl = L.w_lo;
  p.ldb = L.n_pad;
  p.rows = rows;
  p.N = L.k_pad;
  p.K = L.n_pad;
  p.gres = residual ? g : nullptr;
  p.ldgr = ldg;
  p.gres_bcast = g_bcast ? 1 : 0;
  p.g_out = g_in;
  p.ldgo = ld_in;
  const dim3 grid(L.k_pad / BN, ceil_div(rows, BM));
  if (g_bcast)
    hipLaunchKernelGGL((k_fit_gemm<kAGradBcast, kEBwd>), grid, dim3(256), 0, st, p);
  else
    hipLaunchKernelGGL((k_fit_gemm<kAGradMat, kEBwd>), grid, dim3(256), 0, st, p);
  DPMI_LAUNCH_CHECK();
}

void launch_energy_out(const float* e_perm, const int* pos, const int* row_type, const double* shift, long long n,
                       double* atom_e, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_energy_out, dim3(ceil_div(n, 256)), dim3(256), 0, st, e_perm, pos, row_type, shift, n, atom_e);
  DPMI_LAUNCH_CHECK();
}

}  // namespace dpmi
