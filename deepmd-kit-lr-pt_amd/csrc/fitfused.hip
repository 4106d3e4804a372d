// Fitting net of the mixed-precision engine as one fused kernel (gfx950).
//
// deepmd/pt/model/task/fitting.py:1079-1104 and network/mlp.py:188-219 for
// one atom type: hidden layers x_{l+1} = act(x_l W_l + b_l) * idt_l (+ x_l
// when widths are equal), energy e = x_H . w_out + b_out + bias_atom_e, and
// the adjoint dE/dx_0 back to the descriptor.
//
// One workgroup (8 waves) owns a tile of up to 16 RT descriptor rows of one
// type (RT = 4, or 2 when 64-row tiles would leave compute units idle) and
// runs every layer forward, the energy, and every layer backward:
//   * every wave covers all rows of the tile and one eighth of the output
//     columns, so no weight fragment is fetched twice by a workgroup;
//   * the descriptor tile streams from HBM through a double-buffered LDS
//     stage of kKC k-steps (A-fragment order, one barrier per stage); the
//     layer inputs x_l and the backward operands gz_l live in LDS
//     (FP16); the activation derivatives idt*(1 - y^2) (FP16 pairs) and the
//     residual gradients stay in registers (a wave owns the same output
//     positions in every layer);
//   * weights are pre-swizzled on the host into MFMA fragment order, so one
//     B fragment (16 columns x 32 reductions) is one 1 KB coalesced read, and
//     are prefetched kPf k-steps ahead through a register ring;
//   * forward products with the weights split into FP16 hi + lo parts (the
//     energy sees FP32-level weights), backward products with the hi parts
//     (relative gradient error 2^-12, far inside the force tolerance, for
//     half the backward weight traffic);
//   * products on v_mfma_f32_16x16x32_f16 with FP32 accumulation; bias,
//     tanh, timestep, residual and the energy dot product in FP32; dE/dD is
//     written in FP32.
// Every reduction width is a multiple of 32 kPf (the host pads widths to
// multiples of 128).  The host orders tiles so that, with workgroups dealt
// round-robin to the XCDs, each XCD's L2 serves the weights of as few types
// as possible.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <stdexcept>
#include <type_traits>

#include "dpmi_common.h"
#include "kernels.h"

namespace dpmi {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
template <int N>
using halfv = _Float16 __attribute__((ext_vector_type(N)));

constexpr int kWaves = 8;
constexpr int kMaxW = 256;       // widest hidden layer (padded)
constexpr int kXld = kMaxW + 8;  // LDS row stride of the activation tiles (halves)
constexpr int kPf = 4;           // prefetch distance in k-steps
constexpr int kKC = kPf;         // k-steps per descriptor chunk (one barrier each)

__device__ __forceinline__ half8 zero8() {
  half8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = (_Float16)0.f;
  return v;
}

// B fragment (k-step kt, 16-column tile ct) of a swizzled weight matrix with
// NT column tiles: 64 lanes x 8 halves contiguous.
// (Global address space: a generic load would also count against the LDS
// wait counter and drain at every LDS wait.)
typedef __attribute__((address_space(1))) const half8 ghalf8;
__device__ __forceinline__ half8 load_bfrag(const _Float16* __restrict__ w, int kt, int NT, int ct, int lane) {
  return *(ghalf8*)(w + ((size_t)(kt * NT + ct) * 64 + lane) * 8);
}

// C[64 x (ntw tiles of 16 per wave)] += A[64 x K] (B hi [+ B lo when LO])
// over K = 32 KT, A from an LDS tile (row stride ld), B fragments from
// swizzled matrices with NT column tiles; tile j of this wave is column tile
// ct0 + j (used if < NT).  KT is a multiple of kPf.  Loads are
// unconditional (clamped tile and step indices) so that the prefetch ring is
// not serialised by the wait counters.
template <int RT, bool LO, int KTC, int CT = 2>
__device__ __forceinline__ void gemm_lds_a_kt(floatx4 (&acc)[RT][CT], const _Float16* a_lds, int ld,
                                              const _Float16* __restrict__ w, const _Float16* __restrict__ wl,
                                              int KT_, int NT, int ct0, int ntw, int lane) {
  // KTC > 0: the step count is a compile-time constant, the loop is unrolled
  // completely (no loop-carried ring, so the wait counters stay exact) and
  // the last kPf steps issue no refills
  const int KT = KTC > 0 ? KTC : KT_;
  const int lr = lane & 15, lg = lane >> 4;
  // CT column tiles per wave (CT = 1: the second tile's code and loads are
  // compiled out)
  const bool use0 = ntw > 0 && ct0 < NT, use1 = CT > 1 && ntw > 1 && ct0 + 1 < NT;
  const int c0 = min(ct0, NT - 1), c1 = min(ct0 + 1, NT - 1);
  half8 bq[kPf][CT], bl[kPf][CT];
#pragma unroll
  for (int s = 0; s < kPf; ++s) {
    bq[s][0] = load_bfrag(w, s, NT, c0, lane);
    if constexpr (CT > 1) bq[s][CT - 1] = load_bfrag(w, s, NT, c1, lane);
    if constexpr (LO) {
      bl[s][0] = load_bfrag(wl, s, NT, c0, lane);
      if constexpr (CT > 1) bl[s][CT - 1] = load_bfrag(wl, s, NT, c1, lane);
    }
  }
  auto step = [&](int k, int s) {
    half8 a[RT];
#pragma unroll
    for (int i = 0; i < RT; ++i) a[i] = *reinterpret_cast<const half8*>(a_lds + (i * 16 + lr) * ld + k * 32 + lg * 8);
    if (use0)
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], bq[s][0], acc[i][0], 0, 0, 0);
        if constexpr (LO) acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], bl[s][0], acc[i][0], 0, 0, 0);
      }
    if constexpr (CT > 1)
      if (use1)
#pragma unroll
        for (int i = 0; i < RT; ++i) {
          acc[i][CT - 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], bq[s][CT - 1], acc[i][CT - 1], 0, 0, 0);
          if constexpr (LO)
            acc[i][CT - 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], bl[s][CT - 1], acc[i][CT - 1], 0, 0, 0);
        }
    if (KTC > 0 && k + kPf >= KTC) return;  // folds away after unrolling
    const int kn = min(k + kPf, KT - 1);
    bq[s][0] = load_bfrag(w, kn, NT, c0, lane);
    if constexpr (CT > 1) bq[s][CT - 1] = load_bfrag(w, kn, NT, c1, lane);
    if constexpr (LO) {
      bl[s][0] = load_bfrag(wl, kn, NT, c0, lane);
      if constexpr (CT > 1) bl[s][CT - 1] = load_bfrag(wl, kn, NT, c1, lane);
    }
  };
  if constexpr (KTC > 0) {
#pragma unroll
    for (int kt0 = 0; kt0 < KTC; kt0 += kPf)
#pragma unroll
      for (int s = 0; s < kPf; ++s) step(kt0 + s, s);
  } else {
    for (int kt0 = 0; kt0 < KT; kt0 += kPf)
#pragma unroll
      for (int s = 0; s < kPf; ++s) step(kt0 + s, s);
  }
}

template <int RT, bool LO, int CT = 2>
__device__ __forceinline__ void gemm_lds_a(floatx4 (&acc)[RT][CT], const _Float16* a_lds, int ld,
                                           const _Float16* __restrict__ w, const _Float16* __restrict__ wl, int KT,
                                           int NT, int ct0, int ntw, int lane) {
  if (KT == 8)  // 256-wide hidden layers
    gemm_lds_a_kt<RT, LO, 8, CT>(acc, a_lds, ld, w, wl, KT, NT, ct0, ntw, lane);
  else
    gemm_lds_a_kt<RT, LO, 0, CT>(acc, a_lds, ld, w, wl, KT, NT, ct0, ntw, lane);
}

// Lane i of each row of 16 takes lane i + N's value (DPP row_shl; lanes past
// the row end read zero).  Lane 0 of a row then sums like a xor butterfly:
// s + shl1, + shl2, + shl4, + shl8 adds the same pairs in the same order.
template <int N>
__device__ __forceinline__ float dpp_row_shl(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x100 + N, 0xF, 0xF, true));
}

template <int RT, int CT = 2>
__device__ __forceinline__ void zero_acc(floatx4 (&acc)[RT][CT]) {
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < CT; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
}

// One 256-wide product (8 k-steps) of a chain whose fragment ring carries
// over: the ring holds this product's first kPf steps on entry, and its last
// kPf steps fetch the first kPf of the next product (hi, and lo when LON).
// Products and their order are those of gemm_lds_a_kt.
template <int RT, int CT, bool LO, bool LON>
__device__ __forceinline__ void gemm8_chain(floatx4 (&acc)[RT][CT], const _Float16* a_lds, half8 (&rq)[kPf][CT],
                                            half8 (&rl)[kPf][CT], bool use0, bool use1, const _Float16* __restrict__ w,
                                            const _Float16* __restrict__ wl, int NT, int c0, int c1,
                                            const _Float16* __restrict__ wn, const _Float16* __restrict__ wln, int NTn,
                                            int n0, int n1, int lane) {
  constexpr int KS = 8;
  const int lr = lane & 15, lg = lane >> 4;
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int s = k % kPf;
    half8 a[RT];
#pragma unroll
    for (int i = 0; i < RT; ++i) a[i] = *reinterpret_cast<const half8*>(a_lds + (i * 16 + lr) * kXld + k * 32 + lg * 8);
    if (use0)
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], rq[s][0], acc[i][0], 0, 0, 0);
        if constexpr (LO) acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], rl[s][0], acc[i][0], 0, 0, 0);
      }
    if constexpr (CT > 1)
      if (use1)
#pragma unroll
        for (int i = 0; i < RT; ++i) {
          acc[i][CT - 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], rq[s][CT - 1], acc[i][CT - 1], 0, 0, 0);
          if constexpr (LO)
            acc[i][CT - 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], rl[s][CT - 1], acc[i][CT - 1], 0, 0, 0);
        }
    if (k + kPf < KS) {
      rq[s][0] = load_bfrag(w, k + kPf, NT, c0, lane);
      if constexpr (CT > 1) rq[s][CT - 1] = load_bfrag(w, k + kPf, NT, c1, lane);
      if constexpr (LO) {
        rl[s][0] = load_bfrag(wl, k + kPf, NT, c0, lane);
        if constexpr (CT > 1) rl[s][CT - 1] = load_bfrag(wl, k + kPf, NT, c1, lane);
      }
    } else {
      rq[s][0] = load_bfrag(wn, k + kPf - KS, NTn, n0, lane);
      if constexpr (CT > 1) rq[s][CT - 1] = load_bfrag(wn, k + kPf - KS, NTn, n1, lane);
      if constexpr (LON) {
        rl[s][0] = load_bfrag(wln, k + kPf - KS, NTn, n0, lane);
        if constexpr (CT > 1) rl[s][CT - 1] = load_bfrag(wln, k + kPf - KS, NTn, n1, lane);
      }
    }
  }
}

// SPLIT: layer 0's products run in k_fit_layer0 / k_fit_dgrad; this kernel
// reads x_1 and idt_0 act'(z_0) and writes dE/dz_0 instead.  NW waves per
// workgroup, each with CT = 16 / NW of the (at most 16) column tiles of a
// layer: 8 waves x 2 tiles, or 16 waves x 1 tile (twice the waves per SIMD
// to hide the fragment reads, the same fragments fetched once).
template <int H, int RT, bool SPLIT = false, int NW = 8>
__global__ void __launch_bounds__(NW * 64) k_fit_fused(FusedFitArgs p) {
  constexpr int CT = kMaxW / 16 / NW;         // column tiles per wave
  static_assert(CT * NW * 16 == kMaxW, "the waves share the widest layer's column tiles evenly");
  constexpr int kRows = 16 * RT;              // rows per workgroup tile
  constexpr int kFr = kKC * RT / NW;          // descriptor fragments per wave and chunk
  static_assert(kFr * NW == kKC * RT, "a chunk's fragments are shared evenly by the waves");
  __shared__ _Float16 xs[2][kRows][kXld];
  __shared__ half8 afr[2][kKC * RT][kWave];  // descriptor chunks, A-fragment order
  __shared__ float esum[NW][kRows];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int4 ti = p.tiles[blockIdx.x];  // type, first row, row count
  const int t = ti.x, row0 = ti.y, nrows = ti.z;
  if (nrows <= 0) return;  // padding slot of the XCD-aware tile order
  const FusedLayer* Ls = p.layers + (size_t)t * H;
  const int lr = lane & 15, lg = lane >> 4;
  // destination of this thread's atom energy, read now rather than after the
  // forward (the backward's first barrier would wait for that load)
  const int erow = tid < nrows ? p.perm[row0 + tid] : 0;
  const double eshift = p.shift[t];  // uniform: scalar registers
  // 16 waves of one column tile: with every hidden width 256, the products
  // after layer 0's forward run as one chain over a fragment ring that never
  // drains (gemm8_chain); the ring's 32 registers fit beside the rest
  constexpr bool CHAIN = NW == 16 && !SPLIT && H >= 2;
  bool chain = CHAIN;
#pragma unroll
  for (int l = 0; l < H; ++l) chain = chain && Ls[l].N == kMaxW && (l == 0 || Ls[l].K == kMaxW);
  half8 rq[kPf][CT], rl[kPf][CT];  // the chain's fragment ring (hi, lo)

  // activation derivative idt * act'(z) of every hidden layer at this wave's
  // output positions (row tile i, column tile j, rows lg*4 + e), two per register
  half2v dpk[H][RT][CT][2];
  float epart[RT][4];
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) epart[i][e] = 0.f;

  // ------------------------------------------------------------ forward --
#pragma unroll
  for (int l = 0; l < H; ++l) {
    const FusedLayer L = Ls[l];
    const int KT = L.K / 32, NT = L.N / 16, ntw = (NT + NW - 1) / NW;
    const int ct0 = wave * ntw;
    const int src = (l + 1) & 1, dst = l & 1;  // layer l >= 1 reads xs[src]
    if (SPLIT && l == 0) {
      // x_1 rows into LDS (rows past the tile: zeros), the activation
      // derivative at this wave's positions into registers
      const int N8 = L.N / 8;
      for (int q = tid; q < kRows * N8; q += NW * kWave) {
        const int r = q / N8, c8 = q - r * N8;
        const half8 v = r < nrows ? *reinterpret_cast<const half8*>(p.x1 + (size_t)(row0 + r) * p.ldX + c8 * 8)
                                  : zero8();
        *reinterpret_cast<half8*>(&xs[dst][r][c8 * 8]) = v;
      }
#pragma unroll
      for (int j = 0; j < CT; ++j) {
        const int n = min(ct0 + j, NT - 1) * 16 + lr;
#pragma unroll
        for (int i = 0; i < RT; ++i) {
          _Float16 d[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int r = i * 16 + lg * 4 + e;
            d[e] = r < nrows ? p.dv1[(size_t)(row0 + r) * p.ldX + n] : (_Float16)0.f;
          }
          dpk[l][i][j][0] = half2v{d[0], d[1]};
          dpk[l][i][j][1] = half2v{d[2], d[3]};
        }
      }
      __syncthreads();
      continue;
    }
    floatx4 acc[RT][CT];
    zero_acc(acc);
    if (l == 0) {
      // descriptor rows through LDS in chunks of kKC k-steps, stored in
      // A-fragment order (a fragment is 16 rows x 32 k; lane r + 16 g holds
      // row r, k 8g..8g+7, so a wave reads one with one conflict-free
      // 16-byte LDS read per lane); each wave loads kFr fragments of chunk
      // c + 1 into registers while chunk c is multiplied, one barrier per
      // chunk.  Rows past the tile read the tile's first row and are zeroed.
      typedef __attribute__((address_space(1))) const half8 gh8;
      const int NC = KT / kKC;
      const bool use0 = ntw > 0 && ct0 < NT, use1 = CT > 1 && ntw > 1 && ct0 + 1 < NT;
      const int c0 = min(ct0, NT - 1), c1 = min(ct0 + 1, NT - 1);
      const _Float16* dsrc[kFr];
      bool dval[kFr];
#pragma unroll
      for (int m = 0; m < kFr; ++m) {
        const int f = wave + NW * m, r = (f % RT) * 16 + lr;
        dval[m] = r < nrows;
        dsrc[m] = p.D + (size_t)(row0 + (dval[m] ? r : 0)) * p.ldD + (f / RT) * 32 + lg * 8;
      }
      half8 ar[kFr];
      half8(&bq)[kPf][CT] = rq;
      half8(&bl)[kPf][CT] = rl;
      // layer 1's first fragments (fetched by the last kPf steps when chained)
      const FusedLayer L1 = Ls[H > 1 ? 1 : 0];
      const int NT1 = L1.N / 16, ntw1 = (NT1 + NW - 1) / NW;
      const int n10 = min(wave * ntw1, NT1 - 1), n11 = min(wave * ntw1 + 1, NT1 - 1);
#pragma unroll
      for (int m = 0; m < kFr; ++m) ar[m] = __builtin_nontemporal_load((gh8*)dsrc[m]);
#pragma unroll
      for (int s = 0; s < kPf; ++s) {
        bq[s][0] = load_bfrag(L.wf, s, NT, c0, lane);
        bl[s][0] = load_bfrag(L.wf_lo, s, NT, c0, lane);
        if constexpr (CT > 1) {
          bq[s][CT - 1] = load_bfrag(L.wf, s, NT, c1, lane);
          bl[s][CT - 1] = load_bfrag(L.wf_lo, s, NT, c1, lane);
        }
      }
#pragma unroll
      for (int m = 0; m < kFr; ++m) afr[0][wave + NW * m][lane] = dval[m] ? ar[m] : zero8();
      __syncthreads();
      // one chunk; the last one is peeled so that only its refills choose a
      // source (layer 1's first fragments when chained)
      auto chunk = [&](int c, auto last_tag) {
        constexpr bool LAST = decltype(last_tag)::value;
        if constexpr (!LAST)
#pragma unroll
          for (int m = 0; m < kFr; ++m) ar[m] = __builtin_nontemporal_load((gh8*)(dsrc[m] + (c + 1) * kKC * 32));
#pragma unroll
        for (int s = 0; s < kKC; ++s) {
          const int k = c * kKC + s;
          half8 a[RT];
#pragma unroll
          for (int i = 0; i < RT; ++i) a[i] = afr[c & 1][s * RT + i][lane];
          if (use0)
#pragma unroll
            for (int i = 0; i < RT; ++i) {
              acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], bq[s][0], acc[i][0], 0, 0, 0);
              acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], bl[s][0], acc[i][0], 0, 0, 0);
            }
          if constexpr (CT > 1)
            if (use1)
#pragma unroll
              for (int i = 0; i < RT; ++i) {
                acc[i][CT - 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], bq[s][CT - 1], acc[i][CT - 1], 0, 0, 0);
                acc[i][CT - 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], bl[s][CT - 1], acc[i][CT - 1], 0, 0, 0);
              }
          // unconditional loads keep the wait counts exact; in the last chunk
          // step k + kPf is step s of the next product (kKC == kPf), or a
          // repeat of the last step when nothing follows
          const _Float16* wq = L.wf;
          const _Float16* wql = L.wf_lo;
          int kn = k + kPf, NTq = NT, q0 = c0, q1 = c1;
          if constexpr (LAST) {
            kn = chain ? s : KT - 1;
            wq = chain ? L1.wf : L.wf;
            wql = chain ? L1.wf_lo : L.wf_lo;
            NTq = chain ? NT1 : NT;
            q0 = chain ? n10 : c0;
            q1 = chain ? n11 : c1;
          }
          bq[s][0] = load_bfrag(wq, kn, NTq, q0, lane);
          bl[s][0] = load_bfrag(wql, kn, NTq, q0, lane);
          if constexpr (CT > 1) {
            bq[s][CT - 1] = load_bfrag(wq, kn, NTq, q1, lane);
            bl[s][CT - 1] = load_bfrag(wql, kn, NTq, q1, lane);
          }
        }
        // chunk c + 1 goes to the buffer read in chunk c - 1 (every wave has
        // passed the barrier after that chunk)
        if constexpr (!LAST)
#pragma unroll
          for (int m = 0; m < kFr; ++m) afr[(c + 1) & 1][wave + NW * m][lane] = dval[m] ? ar[m] : zero8();
        __syncthreads();
      };
      for (int c = 0; c + 1 < NC; ++c) chunk(c, std::false_type{});
      chunk(NC - 1, std::true_type{});
    } else if (CHAIN && chain) {
      const bool use0 = ntw > 0 && ct0 < NT, use1 = CT > 1 && ntw > 1 && ct0 + 1 < NT;
      const int c0 = min(ct0, NT - 1), c1 = min(ct0 + 1, NT - 1);
      if (l + 1 < H) {  // next: layer l + 1 forward
        const FusedLayer Ln = Ls[l + 1 < H ? l + 1 : l];
        const int NTn = Ln.N / 16, ntwn = (NTn + NW - 1) / NW;
        gemm8_chain<RT, CT, true, true>(acc, &xs[src][0][0], rq, rl, use0, use1, L.wf, L.wf_lo, NT, c0, c1, Ln.wf,
                                        Ln.wf_lo, NTn, min(wave * ntwn, NTn - 1), min(wave * ntwn + 1, NTn - 1), lane);
      } else {  // next: this layer's backward product
        const int OT = L.K / 16, otw = (OT + NW - 1) / NW;
        gemm8_chain<RT, CT, true, false>(acc, &xs[src][0][0], rq, rl, use0, use1, L.wf, L.wf_lo, NT, c0, c1, L.wb,
                                         L.wb, OT, min(wave * otw, OT - 1), min(wave * otw + 1, OT - 1), lane);
      }
    } else {
      gemm_lds_a<RT, true, CT>(acc, &xs[src][0][0], kXld, L.wf, L.wf_lo, KT, NT, ct0, ntw, lane);
    }
    // epilogue: lane holds rows lg*4 + e of each 16x16 tile, column lr
#pragma unroll
    for (int j = 0; j < CT; ++j) {
      if (j >= ntw || ct0 + j >= NT) continue;
      const int n = (ct0 + j) * 16 + lr;
      const float bn = L.b[n];
      const float idn = L.idt ? L.idt[n] : 1.f;
      const float own = (l == H - 1) ? L.out_w[n] : 0.f;
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        float dv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = i * 16 + lg * 4 + e;
          const float z = acc[i][j][e] + bn;
          const float y = L.act ? tanhf(z) : z;
          dv[e] = L.act ? idn * (1.f - y * y) : idn;
          float o = y * idn;
          if (L.res) o += (float)xs[src][r][n];  // equal widths: x_l at the same position
          if (l == H - 1)
            epart[i][e] += o * own;
          else
            xs[dst][r][n] = (_Float16)o;
        }
        dpk[l][i][j][0] = half2v{(_Float16)dv[0], (_Float16)dv[1]};
        dpk[l][i][j][1] = half2v{(_Float16)dv[2], (_Float16)dv[3]};
      }
    }
    __syncthreads();
  }

  // energy: row sums over the 16 column lanes, then over the 8 waves
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float s = epart[i][e];
      s += dpp_row_shl<1>(s);
      s += dpp_row_shl<2>(s);
      s += dpp_row_shl<4>(s);
      s += dpp_row_shl<8>(s);
      if (lr == 0) esum[wave][i * 16 + lg * 4 + e] = s;
    }
  __syncthreads();
  if (tid < nrows) {
    float e = 0.f;
#pragma unroll
    for (int w = 0; w < NW; ++w) e += esum[w][tid];
    p.atom_e[erow] = (double)e + eshift;
  }

  // ----------------------------------------------------------- backward --
  // g: dE/d(output of layer l) at this wave's positions; top: w_out
  float g[RT][CT][4];
  {
    const FusedLayer L = Ls[H - 1];
    const int NT = L.N / 16, ntw = (NT + NW - 1) / NW;
#pragma unroll
    for (int j = 0; j < CT; ++j) {
      const int ct = wave * ntw + j;
      const float w = (j < ntw && ct < NT) ? L.out_w[ct * 16 + lr] : 0.f;
#pragma unroll
      for (int i = 0; i < RT; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) g[i][j][e] = w;
    }
  }
#pragma unroll
  for (int l = H - 1; l >= 0; --l) {
    const FusedLayer L = Ls[l];
    const int NT = L.N / 16, ntw = (NT + NW - 1) / NW;
    const int buf = l & 1;
    // gz = g * idt * act'(z) -> LDS (the A operand, reduction over N)
#pragma unroll
    for (int j = 0; j < CT; ++j) {
      const int ct = wave * ntw + j;
      if (j >= ntw || ct >= NT) continue;
      const int n = ct * 16 + lr;
#pragma unroll
      for (int i = 0; i < RT; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float d = (float)dpk[l][i][j][e >> 1][e & 1];
          const _Float16 gzv = (_Float16)(g[i][j][e] * d);
          const int r = i * 16 + lg * 4 + e;
          if (SPLIT && l == 0) {
            if (r < nrows) p.gz0[(size_t)(row0 + r) * p.ldX + n] = gzv;
          } else {
            xs[buf][r][n] = gzv;
          }
        }
    }
    if (SPLIT && l == 0) return;  // the descriptor gradient is k_fit_dgrad's
    __syncthreads();
    const int KT = L.N / 32;  // reduction steps
    const int OT = L.K / 16;  // output column tiles (input width of the layer)
    if (l > 0) {
      const int otw = (OT + NW - 1) / NW;
      floatx4 acc[RT][CT];
      zero_acc(acc);
      if (CHAIN && chain) {
        const int ct0 = wave * otw;
        const bool use0 = otw > 0 && ct0 < OT, use1 = CT > 1 && otw > 1 && ct0 + 1 < OT;
        const FusedLayer Ln = Ls[l - 1];  // next: layer l - 1's backward product
        const int OTn = Ln.K / 16;
        int n0, n1;
        if (l > 1) {
          const int otwn = (OTn + NW - 1) / NW;
          n0 = min(wave * otwn, OTn - 1);
          n1 = min(wave * otwn + 1, OTn - 1);
        } else {  // layer 0: the first chunk of output tiles
          n0 = min(wave * CT, OTn - 1);
          n1 = min(wave * CT + CT - 1, OTn - 1);
        }
        gemm8_chain<RT, CT, false, false>(acc, &xs[buf][0][0], rq, rl, use0, use1, L.wb, L.wb, OT, min(ct0, OT - 1),
                                          min(ct0 + 1, OT - 1), Ln.wb, Ln.wb, OTn, n0, n1, lane);
      } else {
        gemm_lds_a<RT, false, CT>(acc, &xs[buf][0][0], kXld, L.wb, L.wb, KT, OT, wave * otw, otw, lane);
      }
      // g_l = gz W^T (+ g_{l+1} for residual layers: same positions)
#pragma unroll
      for (int i = 0; i < RT; ++i)
#pragma unroll
        for (int j = 0; j < CT; ++j)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            g[i][j][e] = (j < otw && wave * otw + j < OT ? acc[i][j][e] : 0.f) + (L.res ? g[i][j][e] : 0.f);
    } else if (KT == 8) {
      // descriptor gradient, 256-wide layer 0: K0 output columns in chunks of
      // NW waves x CT tiles, back to back -- the last kPf k-steps of a chunk
      // fetch the first fragments of the next, so the fragment ring does not
      // drain between chunks (same products in the same order as below)
      constexpr int KS = 8;
      const int nchunk = (OT + CT * NW - 1) / (CT * NW);
      auto tile = [&](int c, int j) { return min(c * CT * NW + wave * CT + j, OT - 1); };
      half8(&bq)[kPf][CT] = rq;  // holds chunk 0's first steps when chained
      if (!(CHAIN && chain))
#pragma unroll
        for (int s = 0; s < kPf; ++s)
#pragma unroll
          for (int j = 0; j < CT; ++j) bq[s][j] = load_bfrag(L.wb, s, OT, tile(0, j), lane);
      for (int c = 0; c < nchunk; ++c) {
        // the A fragments are the same for every chunk: read them per chunk
        // (kept across chunks they would take 32 RT registers)
        asm volatile("" ::: "memory");
        const int cn = min(c + 1, nchunk - 1);
        floatx4 acc[RT][CT];
        zero_acc(acc);
#pragma unroll
        for (int k = 0; k < KS; ++k) {
          const int s = k % kPf;
          half8 a[RT];
#pragma unroll
          for (int i = 0; i < RT; ++i)
            a[i] = *reinterpret_cast<const half8*>(&xs[buf][i * 16 + lr][k * 32 + lg * 8]);
#pragma unroll
          for (int j = 0; j < CT; ++j)
#pragma unroll
            for (int i = 0; i < RT; ++i)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], bq[s][j], acc[i][j], 0, 0, 0);
#pragma unroll
          for (int j = 0; j < CT; ++j)
            bq[s][j] = k + kPf < KS ? load_bfrag(L.wb, k + kPf, OT, tile(c, j), lane)
                                    : load_bfrag(L.wb, k + kPf - KS, OT, tile(cn, j), lane);
        }
        const int ot0 = c * CT * NW + wave * CT;
#pragma unroll
        for (int j = 0; j < CT; ++j) {
          if (ot0 + j >= OT) continue;
          const int col = (ot0 + j) * 16 + lr;
#pragma unroll
          for (int i = 0; i < RT; ++i)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int r = i * 16 + lg * 4 + e;
              if (r < nrows) __builtin_nontemporal_store(acc[i][j][e], p.gD + (size_t)(row0 + r) * p.ldG + col);
            }
        }
      }
    } else {
      // descriptor gradient: K0 output columns in chunks of NW waves x CT tiles
      const int nchunk = (OT + CT * NW - 1) / (CT * NW);
      for (int c = 0; c < nchunk; ++c) {
        const int ot0 = c * CT * NW + wave * CT;
        floatx4 acc[RT][CT];
        zero_acc(acc);
        gemm_lds_a<RT, false, CT>(acc, &xs[buf][0][0], kXld, L.wb, L.wb, KT, OT, ot0, CT, lane);
#pragma unroll
        for (int j = 0; j < CT; ++j) {
          if (ot0 + j >= OT) continue;
          const int col = (ot0 + j) * 16 + lr;
#pragma unroll
          for (int i = 0; i < RT; ++i)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int r = i * 16 + lg * 4 + e;
              if (r < nrows) __builtin_nontemporal_store(acc[i][j][e], p.gD + (size_t)(row0 + r) * p.ldG + col);
            }
        }
      }
    }
  }
}

// Workgroup b of a launch over (row tile, column group) pairs: workgroups
// are dealt round-robin to the nx XCDs, and the groups of one tile go to the
// XCD that the tile order gives it, one after another (the tile's rows are
// read from that XCD's L2 after the first group).
__device__ __forceinline__ void tile_group(int b, int nx, int ngroups, int* tix, int* cg) {
  const int xi = b % nx, k = b / nx;
  *cg = k % ngroups;
  *tix = (k / ngroups) * nx + xi;
}

// Layer 0 forward as a product tiled over rows and output columns: a
// workgroup of 4 waves takes one row tile and 4 column tiles (wave w: column
// tile 4 cg + w, every row of the tile), so it streams a quarter of W_0
// instead of all of it.  The descriptor rows stage through LDS as in the
// fused kernel; the epilogue writes x_1 = act(z) idt and idt act'(z) (FP16).
// Per output the products run in the fused kernel's order (k ascending, hi
// then lo), so the results are the same bits.
template <int RT>
__global__ void __launch_bounds__(256) k_fit_layer0(FusedFitArgs p, int ntiles, int ngroups) {
  constexpr int kW = 4;
  constexpr int kFr = kKC * RT / kW;
  static_assert(kFr * kW == kKC * RT, "a chunk's fragments are shared evenly by the waves");
  __shared__ half8 afr[2][kKC * RT][kWave];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  int tix, cg;
  tile_group(blockIdx.x, p.nxcd, ngroups, &tix, &cg);
  if (tix >= ntiles) return;
  const int4 ti = p.tiles[tix];
  const int t = ti.x, row0 = ti.y, nrows = ti.z;
  if (nrows <= 0) return;
  const FusedLayer L = p.layers[(size_t)t * p.H];
  const int KT = L.K / 32, NT = L.N / 16, NC = KT / kKC;
  const int lr = lane & 15, lg = lane >> 4;
  const int ct = cg * kW + wave;
  const bool use = ct < NT;
  const int c0 = min(ct, NT - 1);
  typedef __attribute__((address_space(1))) const half8 gh8;
  const _Float16* dsrc[kFr];
  bool dval[kFr];
#pragma unroll
  for (int m = 0; m < kFr; ++m) {
    const int f = wave + kW * m, r = (f % RT) * 16 + lr;
    dval[m] = r < nrows;
    dsrc[m] = p.D + (size_t)(row0 + (dval[m] ? r : 0)) * p.ldD + (f / RT) * 32 + lg * 8;
  }
  floatx4 acc[RT];
#pragma unroll
  for (int i = 0; i < RT; ++i) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
  // descriptor chunks in a ring of three register sets: chunk c + 1 is
  // requested two chunks before it is stored to LDS (at the end of chunk c)
  half8 ar0[kFr], ar1[kFr], ar2[kFr];
  half8 bq[kPf], bl[kPf];
  auto dchunk = [&](half8(&r)[kFr], int c) {
    const int cc = min(c, NC - 1);
#pragma unroll
    for (int m = 0; m < kFr; ++m) r[m] = *(gh8*)(dsrc[m] + cc * kKC * 32);
  };
  dchunk(ar0, 0);
  dchunk(ar1, 1);
  dchunk(ar2, 2);
#pragma unroll
  for (int s = 0; s < kPf; ++s) {
    bq[s] = load_bfrag(L.wf, s, NT, c0, lane);
    bl[s] = load_bfrag(L.wf_lo, s, NT, c0, lane);
  }
#pragma unroll
  for (int m = 0; m < kFr; ++m) afr[0][wave + kW * m][lane] = dval[m] ? ar0[m] : zero8();
  __syncthreads();
  for (int c = 0; c < NC; ++c) {
    dchunk(ar0, c + 3);
#pragma unroll
    for (int s = 0; s < kKC; ++s) {
      const int k = c * kKC + s;
      if (use)
#pragma unroll
        for (int i = 0; i < RT; ++i) {
          const half8 a = afr[c & 1][s * RT + i][lane];
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, bq[s], acc[i], 0, 0, 0);
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, bl[s], acc[i], 0, 0, 0);
        }
      const int kn = min(k + kPf, KT - 1);
      bq[s] = load_bfrag(L.wf, kn, NT, c0, lane);
      bl[s] = load_bfrag(L.wf_lo, kn, NT, c0, lane);
    }
    if (c + 1 < NC)
#pragma unroll
      for (int m = 0; m < kFr; ++m) afr[(c + 1) & 1][wave + kW * m][lane] = dval[m] ? ar1[m] : zero8();
#pragma unroll
    for (int m = 0; m < kFr; ++m) {
      ar1[m] = ar2[m];
      ar2[m] = ar0[m];
    }
    __syncthreads();
  }
  if (!use) return;
  const int n = ct * 16 + lr;
  const float bn = L.b[n];
  const float idn = L.idt ? L.idt[n] : 1.f;
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r = i * 16 + lg * 4 + e;
      if (r >= nrows) continue;
      const float z = acc[i][e] + bn;
      const float y = L.act ? tanhf(z) : z;
      const size_t o = (size_t)(row0 + r) * p.ldX + n;
      p.x1[o] = (_Float16)(y * idn);
      p.dv1[o] = (_Float16)(L.act ? idn * (1.f - y * y) : idn);
    }
}

// Layer 0 backward, dE/dD = dE/dz_0 W_0^T, tiled over rows and output
// columns: a workgroup of 8 waves takes one row tile and 16 column tiles of
// the descriptor (wave w: column tiles 16 cg + 2 w, + 1); dE/dz_0 rows stage
// through LDS.  Same product order as the fused kernel's descriptor gradient.
template <int RT>
__global__ void __launch_bounds__(512) k_fit_dgrad(FusedFitArgs p, int ntiles, int ngroups) {
  constexpr int kRows = 16 * RT;
  __shared__ _Float16 xs[kRows][kXld];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  int tix, cg;
  tile_group(blockIdx.x, p.nxcd, ngroups, &tix, &cg);
  if (tix >= ntiles) return;
  const int4 ti = p.tiles[tix];
  const int t = ti.x, row0 = ti.y, nrows = ti.z;
  if (nrows <= 0) return;
  const FusedLayer L = p.layers[(size_t)t * p.H];
  const int N8 = L.N / 8;
  for (int q = tid; q < kRows * N8; q += kWaves * kWave) {
    const int r = q / N8, c8 = q - r * N8;
    const half8 v =
        r < nrows ? *reinterpret_cast<const half8*>(p.gz0 + (size_t)(row0 + r) * p.ldX + c8 * 8) : zero8();
    *reinterpret_cast<half8*>(&xs[r][c8 * 8]) = v;
  }
  __syncthreads();
  const int KT = L.N / 32, OT = L.K / 16;
  const int ot0 = (cg * kWaves + wave) * 2;
  if (ot0 >= OT) return;
  const int lr = lane & 15, lg = lane >> 4;
  floatx4 acc[RT][2];
  zero_acc(acc);
  gemm_lds_a<RT, false>(acc, &xs[0][0], kXld, L.wb, L.wb, KT, OT, ot0, 2, lane);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    if (ot0 + j >= OT) continue;
    const int col = (ot0 + j) * 16 + lr;
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = i * 16 + lg * 4 + e;
        if (r < nrows) __builtin_nontemporal_store(acc[i][j][e], p.gD + (size_t)(row0 + r) * p.ldG + col);
      }
  }
}

}  // namespace

template <int RT, int NW = kWaves>
void launch_fit_fused_rt(const FusedFitArgs& a, int ntiles, hipStream_t st) {
  const dim3 block(NW * kWave);
  switch (a.H) {
    case 1: hipLaunchKernelGGL((k_fit_fused<1, RT, false, NW>), dim3(ntiles), block, 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_fit_fused<2, RT, false, NW>), dim3(ntiles), block, 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_fit_fused<3, RT, false, NW>), dim3(ntiles), block, 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_fit_fused<4, RT, false, NW>), dim3(ntiles), block, 0, st, a); break;
    default: throw std::invalid_argument("fused fitting supports 1 to 4 hidden layers");
  }
}

void launch_fit_fused(const FusedFitArgs& a, int ntiles, int rows_per_tile, hipStream_t st, bool wide) {
  if (ntiles <= 0) return;
  if (rows_per_tile == 64 && wide)
    launch_fit_fused_rt<4, 16>(a, ntiles, st);
  else if (rows_per_tile == 64)
    launch_fit_fused_rt<4>(a, ntiles, st);
  else if (rows_per_tile == 32)
    launch_fit_fused_rt<2>(a, ntiles, st);
  else
    throw std::invalid_argument("fused fitting tiles hold 32 or 64 rows");
  DPMI_LAUNCH_CHECK();
}

template <int RT>
void launch_fit_split_rt(const FusedFitArgs& a, int ntiles, const FusedLayer& l0, hipStream_t st) {
  const int nx = std::max(a.nxcd, 1);
  const int rounds = (ntiles + nx - 1) / nx;  // tiles per XCD
  const int g1 = (l0.N / 16 + 3) / 4, g3 = (l0.K / 16 + 2 * kWaves - 1) / (2 * kWaves);
  hipLaunchKernelGGL((k_fit_layer0<RT>), dim3(rounds * nx * g1), dim3(256), 0, st, a, ntiles, g1);
  switch (a.H) {
    case 2: hipLaunchKernelGGL((k_fit_fused<2, RT, true>), dim3(ntiles), dim3(512), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_fit_fused<3, RT, true>), dim3(ntiles), dim3(512), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_fit_fused<4, RT, true>), dim3(ntiles), dim3(512), 0, st, a); break;
    default: throw std::invalid_argument("split fitting supports 2 to 4 hidden layers");
  }
  hipLaunchKernelGGL((k_fit_dgrad<RT>), dim3(rounds * nx * g3), dim3(512), 0, st, a, ntiles, g3);
}

void launch_fit_split(const FusedFitArgs& a, int ntiles, int rows_per_tile, const FusedLayer& layer0_host,
                      hipStream_t st) {
  if (ntiles <= 0) return;
  if (rows_per_tile == 64)
    launch_fit_split_rt<4>(a, ntiles, layer0_host, st);
  else if (rows_per_tile == 32)
    launch_fit_split_rt<2>(a, ntiles, layer0_host, st);
  else
    throw std::invalid_argument("fused fitting tiles hold 32 or 64 rows");
  DPMI_LAUNCH_CHECK();
}

// Fragment order of B operands: Bt[n][k] (n output column, k reduction) of a
// [N, K] operand goes to ((k / 32) * (N / 16) + n / 16) * 512 + lane * 8 + e
// with lane = n % 16 + 16 * ((k % 32) / 8), e = k % 8.
void swizzle_bfrags(const _Float16* bt, int N, int K, _Float16* out) {
  const int NT = N / 16;
  for (int n = 0; n < N; ++n)
    for (int k = 0; k < K; ++k) {
      const int lane = n % 16 + 16 * ((k % 32) / 8);
      out[((size_t)(k / 32) * NT + n / 16) * 512 + lane * 8 + k % 8] = bt[(size_t)n * K + k];
    }
}

}  // namespace dpmi
