// Forces and virials from net_deriv = dE/d(em) (gfx950), and their
// adjoints w.r.t. net_deriv.
//
// prod_force_a (prod_force.cc:24-87): with f_s = sum_a nd[s,a] ed[s,a,:],
//   F_i -= sum_s f_s, F_j(s) += f_s.
// prod_virial_a (prod_virial.cc:23-69): with v_s[p][q] = sum_a nd[s,a]
//   ed[s,a,p] rij[s,q], virial += sum v_s and atom_virial[j(s)] += v_s.
// One wavefront per centre row, lanes over neighbour slots (4-, 12- and
// 3-wide rows read as contiguous spans); the scatter to neighbours uses FP64
// atomics, per-row sums use a wave butterfly, and totals are reduced per row
// into a partial buffer that a second kernel sums with compensation.
#include <algorithm>
#include <stdexcept>

#include "dpmi_common.h"
#include "env_core.h"
#include "kernels.h"

namespace dpmi {

namespace {

template <typename T>
__device__ __forceinline__ void slot_force(const T* nd, const T* ed, T* f) {
  f[0] = f[1] = f[2] = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) f[d] += nd[a] * ed[a * 3 + d];
}

template <typename T>
__device__ __forceinline__ void slot_virial(const T* nd, const T* ed, const T* r, T* v) {
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      T s = 0;
#pragma unroll
      for (int a = 0; a < 4; ++a) s += nd[a] * ed[a * 3 + p];
      v[p * 3 + q] = s * r[q];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_prod_force(T* __restrict__ force, const T* __restrict__ nd,
                                                    const T* __restrict__ ed, const int* __restrict__ nlist,
                                                    int nloc, int nall, int nnei, int nframes) {
  const int g = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = lane_id();
  if (g >= nframes * nloc) return;
  const int f = g / nloc, i = g % nloc;
  double fi[3] = {0., 0., 0.};
  for (int s = lane; s < nnei; s += kWave) {
    const size_t b = (size_t)g * nnei + s;
    T n4[4], e12[12], fs[3];
#pragma unroll
    for (int a = 0; a < 4; ++a) n4[a] = nd[b * 4 + a];
#pragma unroll
    for (int k = 0; k < 12; ++k) e12[k] = ed[b * 12 + k];
    slot_force(n4, e12, fs);
    fi[0] -= fs[0];
    fi[1] -= fs[1];
    fi[2] -= fs[2];
    const int j = nlist[b];
    if (j >= 0) {
      T* fj = force + ((size_t)f * nall + j) * 3;
      atomicAdd(fj + 0, fs[0]);
      atomicAdd(fj + 1, fs[1]);
      atomicAdd(fj + 2, fs[2]);
    }
  }
  for (int d = 0; d < 3; ++d) fi[d] = wave_sum(fi[d]);
  if (lane == 0) {
    T* fo = force + ((size_t)f * nall + i) * 3;
    atomicAdd(fo + 0, (T)fi[0]);
    atomicAdd(fo + 1, (T)fi[1]);
    atomicAdd(fo + 2, (T)fi[2]);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_prod_virial(double* __restrict__ vpart, T* __restrict__ atom_virial,
                                                     const T* __restrict__ nd, const T* __restrict__ ed,
                                                     const T* __restrict__ rij, const int* __restrict__ nlist,
                                                     int nloc, int nnei) {
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = lane_id();
  if (i >= nloc) return;
  double vs[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) vs[k] = 0.;
  for (int s = lane; s < nnei; s += kWave) {
    const size_t b = (size_t)i * nnei + s;
    const int j = nlist[b];
    if (j < 0) continue;
    T n4[4], e12[12], r3[3], v[9];
#pragma unroll
    for (int a = 0; a < 4; ++a) n4[a] = nd[b * 4 + a];
#pragma unroll
    for (int k = 0; k < 12; ++k) e12[k] = ed[b * 12 + k];
#pragma unroll
    for (int d = 0; d < 3; ++d) r3[d] = rij[b * 3 + d];
    slot_virial(n4, e12, r3, v);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      vs[k] += v[k];
      if (atom_virial) atomicAdd(atom_virial + (size_t)j * 9 + k, v[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) vs[k] = wave_sum(vs[k]);
  if (lane < 9) {
    double out = 0.;
#pragma unroll
    for (int k = 0; k < 9; ++k)
      if (lane == k) out = vs[k];
    vpart[(size_t)i * 9 + lane] = out;
  }
}

// Sum of rows [nrows, 9] -> out9 (one workgroup, compensated).
template <typename T>
__global__ void __launch_bounds__(256) k_reduce_rows9(T* __restrict__ out, const double* __restrict__ part,
                                                      int nrows) {
  __shared__ double sh[256][9];
  KahanD acc[9];
  for (int r = threadIdx.x; r < nrows; r += blockDim.x)
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k].add(part[(size_t)r * 9 + k]);
#pragma unroll
  for (int k = 0; k < 9; ++k) sh[threadIdx.x][k] = acc[k].value();
  __syncthreads();
  if (threadIdx.x < 9) {
    KahanD t;
    for (int r = 0; r < (int)blockDim.x; ++r) t.add(sh[r][threadIdx.x]);
    out[threadIdx.x] = (T)t.value();
  }
}

// prod_force_grad.cc:23-75 (ghost index folded with j % nloc)
template <typename T>
__global__ void k_prod_force_grad(T* __restrict__ gn, const T* __restrict__ grad,
                                  const T* __restrict__ ed, const int* __restrict__ nlist, int nloc, int nnei,
                                  int nframes) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long long)nframes * nloc * nnei) return;
  const long long g = tid / nnei;
  const int s = (int)(tid - g * nnei);
  const int f = (int)(g / nloc);
  int j = nlist[g * nnei + s];
  if (j >= nloc) j = j % nloc;
  const T* e = ed + ((size_t)g * nnei + s) * 12;
  const T gi[3] = {grad[g * 3 + 0], grad[g * 3 + 1], grad[g * 3 + 2]};
  T gj[3] = {0, 0, 0};
  if (j >= 0)
    for (int d = 0; d < 3; ++d) gj[d] = grad[((size_t)f * nloc + j) * 3 + d];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    T v = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) v -= gi[d] * e[a * 3 + d];
    if (j >= 0)
#pragma unroll
      for (int d = 0; d < 3; ++d) v += gj[d] * e[a * 3 + d];
    gn[((size_t)g * nnei + s) * 4 + a] = v;
  }
}

// prod_virial_grad.cc:22-66
template <typename T>
__global__ void k_prod_virial_grad(T* __restrict__ gn, const T* __restrict__ grad,
                                   const T* __restrict__ ed, const T* __restrict__ rij,
                                   const int* __restrict__ nlist, int nloc, int nnei) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long long)nloc * nnei) return;
  const int j = nlist[tid];
  const T* e = ed + (size_t)tid * 12;
  const T* r = rij + (size_t)tid * 3;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    T v = 0;
    if (j >= 0)
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) v += grad[p * 3 + q] * r[q] * e[a * 3 + p];
    gn[(size_t)tid * 4 + a] = v;
  }
}

// Engine: packed rows, extended force buffer per frame, virial partials.
// TI: storage and arithmetic type of the per-slot terms (double, or float in
// the mixed engine); forces and virials are accumulated in FP64.
template <typename TI>
__global__ void __launch_bounds__(256) k_force_virial_rows(const TI* __restrict__ nd, const TI* __restrict__ ed,
                                                           const TI* __restrict__ rij, const int* __restrict__ nlist,
                                                           const int* __restrict__ row_frame,
                                                           const int* __restrict__ row_center,
                                                           const int* __restrict__ frame_ext_off, long long nrows,
                                                           int nnei, double* __restrict__ force_ext,
                                                           double* __restrict__ vpart) {
  const long long g = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = lane_id();
  if (g >= nrows) return;
  const size_t eo = (size_t)frame_ext_off[row_frame[g]];
  const int i = row_center[g];
  double fi[3] = {0., 0., 0.};
  double vs[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) vs[k] = 0.;
  for (int s = lane; s < nnei; s += kWave) {
    const size_t b = (size_t)g * nnei + s;
    const int j = nlist[b];
    if (j < 0) continue;  // padding slots have zero em_deriv
    TI n4[4], e12[12], r3[3], fs[3], v[9];
#pragma unroll
    for (int a = 0; a < 4; ++a) n4[a] = nd[b * 4 + a];
#pragma unroll
    for (int k = 0; k < 12; ++k) e12[k] = ed[b * 12 + k];
#pragma unroll
    for (int d = 0; d < 3; ++d) r3[d] = rij[b * 3 + d];
    slot_force(n4, e12, fs);
    slot_virial(n4, e12, r3, v);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      fi[d] -= fs[d];
      atomicAdd(force_ext + (eo + j) * 3 + d, fs[d]);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) vs[k] += v[k];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) fi[d] = wave_sum(fi[d]);
#pragma unroll
  for (int k = 0; k < 9; ++k) vs[k] = wave_sum(vs[k]);
  if (lane < 3) {
    double o = lane == 0 ? fi[0] : (lane == 1 ? fi[1] : fi[2]);
    atomicAdd(force_ext + (eo + i) * 3 + lane, o);
  }
  if (lane < 9) {
    double out = 0.;
#pragma unroll
    for (int k = 0; k < 9; ++k)
      if (lane == k) out = vs[k];
    vpart[(size_t)g * 9 + lane] = out;
  }
}

// Same products with the neighbour scatter pre-summed in LDS.  Device-scope
// FP64 atomics on a multi-XCD chip are resolved beyond the per-XCD L2s, so
// their rate, not bandwidth, bounds the direct scatter.  Here one workgroup
// takes R centre rows in cell order (positions p of the cell-sorted atom
// list, so the rows' neighbour sets overlap), accumulates every force term
// into an LDS hash keyed by the owner atom (a ghost image counts for the
// local atom it copies: map_ext), and then issues one global atomic per
// distinct owner and component straight into the packed force rows, so no
// extended force buffer and no ghost fold are needed.  The table has 2^HB
// entries, the smallest power of two (>= 2^10) holding the R (nnei + 1)
// distinct keys a workgroup can insert, so every probe sequence ends, and
// smaller row groups leave LDS for more workgroups per compute unit.
constexpr int kHashBitsMax = 12, kHashSize = 1 << kHashBitsMax;

template <int HB>
__device__ __forceinline__ int hash_slot(int* keys, int key) {
  unsigned h = ((unsigned)key * 2654435761u) >> (32 - HB);
  while (true) {
    const int prev = atomicCAS(&keys[h], -1, key);
    if (prev == -1 || prev == key) return (int)h;
    h = (h + 1) & ((1u << HB) - 1);
  }
}

// Bounded probe for row groups that may hold more distinct keys than the
// table: -1 after kProbeMax occupied slots (the caller then adds straight
// into global memory).  The loop stays rolled: unrolled, every probe's
// divergent exit keeps an exec mask in scalar registers, and the kernel
// spilled them.
constexpr int kProbeMax = 32;
template <int HB>
__device__ __forceinline__ int hash_slot_bounded(int* keys, int key) {
  unsigned h = ((unsigned)key * 2654435761u) >> (32 - HB);
#pragma unroll 1
  for (int probe = 0; probe < kProbeMax; ++probe) {
    const int prev = atomicCAS(&keys[h], -1, key);
    if (prev == -1 || prev == key) return (int)h;
    h = (h + 1) & ((1u << HB) - 1);
  }
  return -1;
}

// Environment-matrix derivative and displacement of a slot of a centre at
// ci, rebuilt from the coordinates exactly as the environment kernel stores
// them (FP64 evaluation; TI float: rounded as the FP32-storing kernel does).
template <typename TI>
__device__ __forceinline__ void recompute_slot(const double (&ci)[3], const double* __restrict__ cj,
                                               const double* __restrict__ sd, double rmin, double rmax, TI* e12,
                                               TI* r3) {
  double rr[3], val[4], der[12];
#pragma unroll
  for (int d = 0; d < 3; ++d) rr[d] = cj[d] - ci[d];
  envc::slot_terms<sizeof(TI) == 4, true>(rr, rmin, rmax, val, der);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if constexpr (sizeof(TI) == 4) {
      const double isd = 1. / sd[a];
#pragma unroll
      for (int d = 0; d < 3; ++d) e12[a * 3 + d] = (TI)(der[a * 3 + d] * isd);
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d) e12[a * 3 + d] = der[a * 3 + d] / sd[a];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) r3[d] = (TI)rr[d];
}

// RECOMP: ed and rij are not stored; each slot's derivative and displacement
// are rebuilt from coord_ext and the centre type's dstd.
// AV: per-atom virials as well (prod_virial.cc:55-63: slot s of centre i adds
// its virial term to atom_virial[j(s)]), pre-summed per owner key in LDS like
// the forces and added into atom_virial [ntot, 9] (zeroed by the caller).
template <typename TI, bool RECOMP, int HB, bool AV = false>
__global__ void __launch_bounds__(1024) k_force_virial_hash(
    const TI* __restrict__ nd, const TI* __restrict__ ed, const TI* __restrict__ rij, const int* __restrict__ nlist,
    const int* __restrict__ row_frame, const int* __restrict__ row_center, const int* __restrict__ frame_ext_off,
    const FrameGeom* __restrict__ geom, const int* __restrict__ sorted_loc, long long nrows, int nnei, int R,
    const double* __restrict__ coord_ext, const int* __restrict__ type_ext, const double* __restrict__ dstd,
    double rmin, double rmax, const int* __restrict__ map_ext, double* __restrict__ force,
    double* __restrict__ vpart, double* __restrict__ atom_virial) {
  constexpr int NH = 1 << HB;
  constexpr int NV = AV ? NH : 1;
  __shared__ int hkey[NH];
  __shared__ double hval[NH][3];
  __shared__ double hvir[NV][9];
  for (int h = threadIdx.x; h < NH; h += blockDim.x) {
    hkey[h] = -1;
    hval[h][0] = hval[h][1] = hval[h][2] = 0.;
    if constexpr (AV) {
#pragma unroll
      for (int k = 0; k < 9; ++k) hvir[h][k] = 0.;
    }
  }
  __syncthreads();
  const int lane = lane_id(), wave = threadIdx.x / kWave, W = blockDim.x / kWave;
  // R rows per workgroup, W waves: wave w takes rows w, w + W, ... of the
  // group (consecutive in cell order across the waves); with R > the
  // table's guaranteed capacity the probes are bounded and a slot that finds
  // no room adds straight into global memory
  const bool bounded = (long long)R * (nnei + 1) > NH;
  for (int rr = wave; rr < R; rr += W) {
  const long long p = (long long)blockIdx.x * R + rr;
  if (p < nrows) {
    const int f = row_frame[p];
    const long long g = geom[f].loc_off + sorted_loc[p];  // packed row of the p-th atom in cell order
    const int eo = frame_ext_off[f];
    const int lo = geom[f].loc_off;
    const int i = row_center[g];
    double ci[3] = {0., 0., 0.};
    const double* sdrow = dstd;
    bool live = true;
    if constexpr (RECOMP) {
      const int ti = type_ext[eo + i];
      live = ti >= 0;  // rows of negative type carry no forces
      sdrow = dstd + (size_t)max(ti, 0) * nnei * 4;
#pragma unroll
      for (int d = 0; d < 3; ++d) ci[d] = coord_ext[(size_t)(eo + i) * 3 + d];
    }
    double fi[3] = {0., 0., 0.};
    double vs[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) vs[k] = 0.;
    // SU slots per lane are loaded before any is used (out-of-range slots
    // read slot 0 and are dropped), so their loads overlap
    constexpr int SU = sizeof(TI) == 8 ? 1 : (RECOMP ? 2 : 3);
    for (int s0 = 0; s0 < nnei; s0 += SU * kWave) {
      TI n4[SU][4], e12[SU][12], r3[SU][3];
      double cj[SU][3];
      int jj[SU], ss[SU], own[SU];
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const int s = s0 + u * kWave + lane;
        const bool ok = s < nnei && live;
        ss[u] = ok ? s : 0;
        const size_t b = (size_t)g * nnei + ss[u];
        jj[u] = ok ? nlist[b] : -1;
#pragma unroll
        for (int a = 0; a < 4; ++a) n4[u][a] = nd[b * 4 + a];
        const size_t e = (size_t)eo + (jj[u] >= 0 ? jj[u] : i);
        own[u] = map_ext ? map_ext[e] : (int)e;  // no map: keyed by the extended atom itself
        if constexpr (RECOMP) {
#pragma unroll
          for (int d = 0; d < 3; ++d) cj[u][d] = coord_ext[e * 3 + d];
        } else {
#pragma unroll
          for (int k = 0; k < 12; ++k) e12[u][k] = ed[b * 12 + k];
#pragma unroll
          for (int d = 0; d < 3; ++d) r3[u][d] = rij[b * 3 + d];
        }
      }
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        if (jj[u] < 0) continue;  // padding slots have zero em_deriv
        if constexpr (RECOMP) recompute_slot(ci, cj[u], sdrow + ss[u] * 4, rmin, rmax, e12[u], r3[u]);
        TI fs[3], v[9];
        slot_force(n4[u], e12[u], fs);
        slot_virial(n4[u], e12[u], r3[u], v);
        const int key = lo + own[u];
        const int h = bounded ? hash_slot_bounded<HB>(hkey, key) : hash_slot<HB>(hkey, key);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          fi[d] -= fs[d];
          if (h >= 0)
            atomicAdd(&hval[h][d], (double)fs[d]);
          else
            atomicAdd(force + (size_t)key * 3 + d, (double)fs[d]);
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) vs[k] += v[k];
        if constexpr (AV) {
#pragma unroll
          for (int k = 0; k < 9; ++k) atomicAdd(&hvir[h][k], (double)v[k]);
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) fi[d] = wave_sum(fi[d]);
#pragma unroll
    for (int k = 0; k < 9; ++k) vs[k] = wave_sum(vs[k]);
    if (lane == 0) {
      const int key = lo + (map_ext ? map_ext[eo + i] : eo + i);
      const int h = bounded ? hash_slot_bounded<HB>(hkey, key) : hash_slot<HB>(hkey, key);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (h >= 0)
          atomicAdd(&hval[h][d], fi[d]);
        else
          atomicAdd(force + (size_t)key * 3 + d, fi[d]);
      }
    }
    if (lane < 9) {
      double out = 0.;
#pragma unroll
      for (int k = 0; k < 9; ++k)
        if (lane == k) out = vs[k];
      vpart[(size_t)g * 9 + lane] = out;
    }
  }
  }
  __syncthreads();
  for (int h = threadIdx.x; h < NH; h += blockDim.x) {
    const int key = hkey[h];
    if (key >= 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) atomicAdd(force + (size_t)key * 3 + d, hval[h][d]);
      if constexpr (AV) {
#pragma unroll
        for (int k = 0; k < 9; ++k) atomicAdd(atom_virial + (size_t)key * 9 + k, hvir[h][k]);
      }
    }
  }
}

__global__ void k_fold_ghosts(const double* __restrict__ force_ext, const int* __restrict__ map_ext,
                              const int* __restrict__ xframe_of_ext, const FrameGeom* __restrict__ geom,
                              const int* __restrict__ nall, long long next_tot, double* __restrict__ force) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= next_tot) return;
  const int f = xframe_of_ext[e];
  const FrameGeom& g = geom[f];
  const int k = (int)(e - g.ext_off);
  if (k >= nall[f] || k >= g.ext_cap) return;
  const int owner = g.loc_off + map_ext[e];
#pragma unroll
  for (int d = 0; d < 3; ++d) atomicAdd(force + (size_t)owner * 3 + d, force_ext[e * 3 + d]);
}

// Two-level compensated per-frame sums of the atomic energies and of the
// per-row virial partials.  Level 1: one workgroup per chunk of at most
// kChunkRows rows of one frame (chunk list built on the host): each thread
// sums its rows with compensation, the block combines the thread sums by a
// pairwise tree into 10 values.  Level 2: the frame's chunk partials are
// summed with compensation in chunk order, either by the last chunk of the
// frame to finish (a per-frame ticket counted with a device-scope atomic,
// fences on both sides, reset to zero for the next call) or, without
// tickets, by one wave per frame in a second launch.  Both give the same
// bits.
constexpr int kChunkRows = 1024;

__device__ __forceinline__ void frame_total(const double* __restrict__ part, int c0, int c1, int q,
                                            double* __restrict__ energy, double* __restrict__ virial, int f) {
  // loads issued eight at a time ahead of the (sequential, chunk-ordered)
  // compensated additions, so the frame's partials cost one load latency per
  // eight chunks instead of one per chunk
  KahanD t;
  int c = c0;
  for (; c + 8 <= c1; c += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = part[(size_t)(c + u) * 10 + q];
#pragma unroll
    for (int u = 0; u < 8; ++u) t.add(v[u]);
  }
  for (; c < c1; ++c) t.add(part[(size_t)c * 10 + q]);
  if (q == 0)
    energy[f] = t.value();
  else
    virial[f * 9 + q - 1] = t.value();
}

__global__ void __launch_bounds__(256) k_chunk_reduce(const double* __restrict__ atom_e,
                                                      const double* __restrict__ vpart,
                                                      const int4* __restrict__ chunks, double* __restrict__ part,
                                                      int* __restrict__ ticket, const int* __restrict__ chunk_start,
                                                      double* __restrict__ energy, double* __restrict__ virial) {
  __shared__ double sh[10][256];
  const int4 ch = chunks[blockIdx.x];  // (first row, row count, frame, chunks of the frame)
  KahanD acc[10];
  // a thread's rows (at most kChunkRows / 256 = 4) loaded before they are added
  static_assert(kChunkRows <= 4 * 256, "a thread sums at most four rows");
  double v[4][10];
  if (ch.y > 0)
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int r = threadIdx.x + u * (int)blockDim.x;
    const size_t row = (size_t)ch.x + (r < ch.y ? r : 0);
    v[u][0] = atom_e[row];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[u][k + 1] = vpart[row * 9 + k];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if ((int)threadIdx.x + u * (int)blockDim.x < ch.y)
#pragma unroll
      for (int k = 0; k < 10; ++k) acc[k].add(v[u][k]);
#pragma unroll
  for (int k = 0; k < 10; ++k) sh[k][threadIdx.x] = acc[k].value();
  __syncthreads();
  // pairwise tree over the 256 thread sums (rounding error grows with
  // log2(256), not with the count; the order is fixed, so results repeat)
  // (levels 128 and 64 through LDS, the six within wave 0 by shuffles: the
  // same pairs (t, t + h) added in the same order)
  for (int h = 128; h >= kWave; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < 10; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x >= kWave) return;  // wave 0 writes the partials and takes the ticket
  const int q = threadIdx.x;
  double tv[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) tv[k] = sh[k][q];
#pragma unroll
  for (int h = kWave / 2; h > 0; h >>= 1)
#pragma unroll
    for (int k = 0; k < 10; ++k) tv[k] += __shfl_down(tv[k], h, kWave);
  // lane 0 holds the block's sums (every lane takes them: cross-lane reads
  // outside any condition); lane q < 10 writes value q
  double mine = 0.;
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    const double s0 = __shfl(tv[k], 0, kWave);
    mine = q == k ? s0 : mine;
  }
  if (q < 10) part[(size_t)blockIdx.x * 10 + q] = mine;
  if (!ticket) return;
  __threadfence();  // this chunk's partials before its ticket
  int last = 0;
  if (q == 0) last = atomicAdd(&ticket[ch.z], 1) == ch.w - 1;
  if (!__shfl(last, 0, kWave)) return;
  __threadfence();  // every ticket of the frame before its partials are read
  if (q < 10) frame_total(part, chunk_start[ch.z], chunk_start[ch.z + 1], q, energy, virial, ch.z);
  if (q == 0) ticket[ch.z] = 0;
}

__global__ void k_frame_combine(const double* __restrict__ part, const int* __restrict__ chunk_start, int nframes,
                                double* __restrict__ energy, double* __restrict__ virial) {
  const int f = blockIdx.x;
  const int q = threadIdx.x;
  if (f >= nframes || q >= 10) return;
  frame_total(part, chunk_start[f], chunk_start[f + 1], q, energy, virial, f);
}

}  // namespace

template <typename T>
void launch_prod_force_a(T* force, const T* net_deriv, const T* env_deriv, const int* nlist, int nloc, int nall,
                         int nnei, int nframes, hipStream_t st) {
  DPMI_CHECK(hipMemsetAsync(force, 0, sizeof(T) * (size_t)nframes * nall * 3, st));
  const long long rows = (long long)nframes * nloc;
  if (rows <= 0 || nnei <= 0) return;
  hipLaunchKernelGGL(k_prod_force<T>, dim3(ceil_div(rows * kWave, 256)), dim3(256), 0, st, force, net_deriv, env_deriv,
                     nlist, nloc, nall, nnei, nframes);
  DPMI_LAUNCH_CHECK();
}

template <typename T>
void launch_prod_virial_a(double* vpart, T* atom_virial, const T* net_deriv, const T* env_deriv, const T* rij,
                          const int* nlist, int nloc, int nall, int nnei, hipStream_t st) {
  if (atom_virial) DPMI_CHECK(hipMemsetAsync(atom_virial, 0, sizeof(T) * (size_t)nall * 9, st));
  if (nloc <= 0) return;
  hipLaunchKernelGGL(k_prod_virial<T>, dim3(ceil_div((long long)nloc * kWave, 256)), dim3(256), 0, st, vpart,
                     atom_virial, net_deriv, env_deriv, rij, nlist, nloc, nnei);
  DPMI_LAUNCH_CHECK();
}

template <typename T>
void launch_reduce_rows9(T* out9, const double* part, int nrows, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_rows9<T>, dim3(1), dim3(256), 0, st, out9, part, nrows);
  DPMI_LAUNCH_CHECK();
}

template <typename T>
void launch_prod_force_grad_a(T* grad_net, const T* grad, const T* env_deriv, const int* nlist, int nloc, int nnei,
                              int nframes, hipStream_t st) {
  const long long n = (long long)nframes * nloc * nnei;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_prod_force_grad<T>, dim3(ceil_div(n, 256)), dim3(256), 0, st, grad_net, grad, env_deriv,
                     nlist, nloc, nnei, nframes);
  DPMI_LAUNCH_CHECK();
}

template <typename T>
void launch_prod_virial_grad_a(T* grad_net, const T* grad, const T* env_deriv, const T* rij, const int* nlist,
                               int nloc, int nnei, hipStream_t st) {
  const long long n = (long long)nloc * nnei;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_prod_virial_grad<T>, dim3(ceil_div(n, 256)), dim3(256), 0, st, grad_net, grad, env_deriv,
                     rij, nlist, nloc, nnei);
  DPMI_LAUNCH_CHECK();
}

#define DPMI_FORCE_INST(T)                                                                                            \
  template void launch_prod_force_a<T>(T*, const T*, const T*, const int*, int, int, int, int, hipStream_t);          \
  template void launch_prod_virial_a<T>(double*, T*, const T*, const T*, const T*, const int*, int, int, int,         \
                                        hipStream_t);                                                                 \
  template void launch_reduce_rows9<T>(T*, const double*, int, hipStream_t);                                          \
  template void launch_prod_force_grad_a<T>(T*, const T*, const T*, const int*, int, int, int, hipStream_t);          \
  template void launch_prod_virial_grad_a<T>(T*, const T*, const T*, const T*, const int*, int, int, hipStream_t);
DPMI_FORCE_INST(double)
DPMI_FORCE_INST(float)

bool force_recompute_supported(int nnei) { return nnei + 1 <= kHashSize; }

bool force_rows_aggregated(const ForceRowsArgs& a) {
  return a.sorted_loc && a.geom && (a.map_ext || a.unfolded) && a.force && a.nnei + 1 <= kHashSize;
}

// Lower path set-up: clears the extended force [nall, 3] and per-atom virial
// [nall, 9] outputs and checks the caller's types: locals must equal the
// engine's types, every type must be below ntypes.
__global__ void k_lower_prep(double* __restrict__ force_ext, double* __restrict__ atom_virial_ext, int nall,
                             const int* __restrict__ atype_ext, const int* __restrict__ type_loc, int nloc,
                             int ntypes, int* __restrict__ flag) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nall) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) force_ext[(size_t)e * 3 + d] = 0.;
  if (atom_virial_ext)
#pragma unroll
    for (int k = 0; k < 9; ++k) atom_virial_ext[(size_t)e * 9 + k] = 0.;
  const int t = atype_ext[e];
  if (e < nloc && t != type_loc[e]) atomicOr(flag, kFlagTypeMismatch);
  if (t >= ntypes) atomicOr(flag, kFlagTypeRange);
}

void launch_lower_prep(double* force_ext, double* atom_virial_ext, int nall, const int* atype_ext,
                       const int* type_loc, int nloc, int ntypes, int* flag, hipStream_t st) {
  if (nall <= 0) return;
  hipLaunchKernelGGL(k_lower_prep, dim3(ceil_div(nall, 256)), dim3(256), 0, st, force_ext, atom_virial_ext, nall,
                     atype_ext, type_loc, nloc, ntypes, flag);
  DPMI_LAUNCH_CHECK();
}

void launch_force_virial_rows(const ForceRowsArgs& a, hipStream_t st) {
  if (a.nrows <= 0) return;
  if (!a.em_deriv && !force_rows_aggregated(a))
    throw std::invalid_argument("recomputed slot derivatives need the LDS-aggregated force kernel");
  if (force_rows_aggregated(a)) {
    if (a.atom_virial) {
      // per-atom virials: 9 more LDS sums per key, so the table stays at
      // 2^10 keys (about 100 KB of LDS) and the row group shrinks to fit
      if (a.nnei + 1 > (1 << 10)) throw Unsupported("per-atom virials need nnei < 1024");
      const int R = std::max(1, std::min(a.rows > 0 ? a.rows : 16, (1 << 10) / (a.nnei + 1)));
      const dim3 grid(ceil_div(a.nrows, R)), block(R * kWave);
      if (a.fp32 && !a.em_deriv)
        hipLaunchKernelGGL((k_force_virial_hash<float, true, 10, true>), grid, block, 0, st,
                           (const float*)a.net_deriv, nullptr, nullptr, a.nlist, a.row_frame, a.row_center,
                           a.frame_ext_off, a.geom, a.sorted_loc, a.nrows, a.nnei, R, a.coord_ext, a.type_ext, a.dstd,
                           a.rmin, a.rmax, a.map_ext, a.force, a.vpart, a.atom_virial);
      else if (!a.em_deriv)
        hipLaunchKernelGGL((k_force_virial_hash<double, true, 10, true>), grid, block, 0, st,
                           (const double*)a.net_deriv, nullptr, nullptr, a.nlist, a.row_frame, a.row_center,
                           a.frame_ext_off, a.geom, a.sorted_loc, a.nrows, a.nnei, R, a.coord_ext, a.type_ext, a.dstd,
                           a.rmin, a.rmax, a.map_ext, a.force, a.vpart, a.atom_virial);
      else if (a.fp32)
        hipLaunchKernelGGL((k_force_virial_hash<float, false, 10, true>), grid, block, 0, st,
                           (const float*)a.net_deriv, (const float*)a.em_deriv, (const float*)a.rij, a.nlist,
                           a.row_frame, a.row_center, a.frame_ext_off, a.geom, a.sorted_loc, a.nrows, a.nnei, R,
                           nullptr, nullptr, nullptr, 0., 0., a.map_ext, a.force, a.vpart, a.atom_virial);
      else
        hipLaunchKernelGGL((k_force_virial_hash<double, false, 10, true>), grid, block, 0, st,
                           (const double*)a.net_deriv, (const double*)a.em_deriv, (const double*)a.rij, a.nlist,
                           a.row_frame, a.row_center, a.frame_ext_off, a.geom, a.sorted_loc, a.nrows, a.nnei, R,
                           nullptr, nullptr, nullptr, 0., 0., a.map_ext, a.force, a.vpart, a.atom_virial);
      DPMI_LAUNCH_CHECK();
      return;
    }
    // up to 16 rows: one per wave and a table that holds every key; more
    // rows (a multiple of 16): several per wave, the 2^12-entry table with
    // bounded probes (the rows' neighbour sets overlap, so their distinct
    // owners stay far below R (nnei + 1))
    const int want = a.rows > 0 ? a.rows : 16;
    const int R = want > 16 ? (want + 15) / 16 * 16 : std::max(1, std::min(want, kHashSize / (a.nnei + 1)));
    const int W = std::min(R, 16);
    const dim3 grid(ceil_div(a.nrows, R)), block(W * kWave);
    const int keys = R * (a.nnei + 1);
    const int hb = keys <= (1 << 10) ? 10 : (keys <= (1 << 11) ? 11 : 12);
#define DPMI_HASH_LAUNCH(HB)                                                                                          \
  do {                                                                                                                \
    if (a.fp32 && !a.em_deriv)                                                                                        \
      hipLaunchKernelGGL((k_force_virial_hash<float, true, HB>), grid, block, 0, st, (const float*)a.net_deriv,       \
                         nullptr, nullptr, a.nlist, a.row_frame, a.row_center, a.frame_ext_off, a.geom, a.sorted_loc, \
                         a.nrows, a.nnei, R, a.coord_ext, a.type_ext, a.dstd, a.rmin, a.rmax, a.map_ext, a.force,     \
                         a.vpart, nullptr);                                                                           \
    else if (!a.em_deriv)                                                                                             \
      hipLaunchKernelGGL((k_force_virial_hash<double, true, HB>), grid, block, 0, st, (const double*)a.net_deriv,     \
                         nullptr, nullptr, a.nlist, a.row_frame, a.row_center, a.frame_ext_off, a.geom, a.sorted_loc, \
                         a.nrows, a.nnei, R, a.coord_ext, a.type_ext, a.dstd, a.rmin, a.rmax, a.map_ext, a.force,     \
                         a.vpart, nullptr);                                                                           \
    else if (a.fp32)                                                                                                  \
      hipLaunchKernelGGL((k_force_virial_hash<float, false, HB>), grid, block, 0, st, (const float*)a.net_deriv,      \
                         (const float*)a.em_deriv, (const float*)a.rij, a.nlist, a.row_frame, a.row_center,           \
                         a.frame_ext_off, a.geom, a.sorted_loc, a.nrows, a.nnei, R, nullptr, nullptr, nullptr, 0., 0., \
                         a.map_ext, a.force, a.vpart, nullptr);                                                       \
    else                                                                                                              \
      hipLaunchKernelGGL((k_force_virial_hash<double, false, HB>), grid, block, 0, st, (const double*)a.net_deriv,    \
                         (const double*)a.em_deriv, (const double*)a.rij, a.nlist, a.row_frame, a.row_center,         \
                         a.frame_ext_off, a.geom, a.sorted_loc, a.nrows, a.nnei, R, nullptr, nullptr, nullptr, 0., 0., \
                         a.map_ext, a.force, a.vpart, nullptr);                                                       \
  } while (0)
    if (hb == 10)
      DPMI_HASH_LAUNCH(10);
    else if (hb == 11)
      DPMI_HASH_LAUNCH(11);
    else
      DPMI_HASH_LAUNCH(12);
#undef DPMI_HASH_LAUNCH
    DPMI_LAUNCH_CHECK();
    return;
  }
  if (a.fp32)
    hipLaunchKernelGGL(k_force_virial_rows<float>, dim3(ceil_div(a.nrows * kWave, 256)), dim3(256), 0, st,
                       (const float*)a.net_deriv, (const float*)a.em_deriv, (const float*)a.rij, a.nlist,
                       a.row_frame, a.row_center, a.frame_ext_off, a.nrows, a.nnei, a.force_ext, a.vpart);
  else
    hipLaunchKernelGGL(k_force_virial_rows<double>, dim3(ceil_div(a.nrows * kWave, 256)), dim3(256), 0, st,
                       (const double*)a.net_deriv, (const double*)a.em_deriv, (const double*)a.rij, a.nlist,
                       a.row_frame, a.row_center, a.frame_ext_off, a.nrows, a.nnei, a.force_ext, a.vpart);
  DPMI_LAUNCH_CHECK();
}

void launch_fold_ghost_forces(const double* force_ext, const int* map_ext, const int* xframe_of_ext,
                              const FrameGeom* geom, const int* nall, long long next_tot, double* force,
                              hipStream_t st) {
  if (next_tot <= 0) return;
  hipLaunchKernelGGL(k_fold_ghosts, dim3(ceil_div(next_tot, 256)), dim3(256), 0, st, force_ext, map_ext,
                     xframe_of_ext, geom, nall, next_tot, force);
  DPMI_LAUNCH_CHECK();
}

void frame_chunks(const int* natoms, int nframes, std::vector<int>& chunks4, std::vector<int>& chunk_start) {
  chunks4.clear();
  chunk_start.assign(nframes + 1, 0);
  int off = 0;
  for (int f = 0; f < nframes; ++f) {
    chunk_start[f] = (int)chunks4.size() / 4;
    const int nc = (natoms[f] + kChunkRows - 1) / kChunkRows;
    for (int r = 0; r < natoms[f]; r += kChunkRows) {
      chunks4.push_back(off + r);
      chunks4.push_back(std::min(kChunkRows, natoms[f] - r));
      chunks4.push_back(f);
      chunks4.push_back(nc);
    }
    off += natoms[f];
  }
  chunk_start[nframes] = (int)chunks4.size() / 4;
}

void launch_frame_reduce(const double* atom_e, const double* vpart, const int* chunks4, int nchunks,
                         const int* chunk_start, int nframes, double* part, int* ticket, double* energy,
                         double* virial, hipStream_t st) {
  if (nchunks > 0) {
    hipLaunchKernelGGL(k_chunk_reduce, dim3(nchunks), dim3(256), 0, st, atom_e, vpart,
                       reinterpret_cast<const int4*>(chunks4), part, ticket, chunk_start, energy, virial);
    DPMI_LAUNCH_CHECK();
  }
  if (ticket) return;
  hipLaunchKernelGGL(k_frame_combine, dim3(nframes), dim3(kWave), 0, st, part, chunk_start, nframes, energy, virial);
  DPMI_LAUNCH_CHECK();
}

}  // namespace dpmi
