// Coordinate wrapping, cell binning and ghost-image emission for packed
// batches of periodic frames (gfx950).
//
// Behaviour follows the reference CPU path normalize_coord_cpu
// (coord.cc:14-28) -> copy_coord_cpu (coord.cc:31-66 -> neighbor_list.cc
// :43-148, 747-847): extended atoms are the locals followed, for every
// ghost cell of the extended grid in (x, y, z) lexicographic order, by the
// atoms of its periodic source cell in ascending index.  The layout is
// produced with counting sorts and prefix sums instead of the reference's
// per-cell vectors:
//   1. k_wrap_bin      one thread per local atom: wrap, bin, count per cell
//                      (the count's old value is the atom's slot in its
//                      cell; the atom's force row is cleared for the force
//                      kernel when asked)
//   2. k_scan_cells    one workgroup per frame: exclusive scan of the counts,
//                      then the prefix over extended cells (all atoms, and
//                      ghosts only) -> nall; for frames of at most 4096
//                      atoms the workgroup also does step 3
//   3. k_scatter_cells one thread per local atom: place into its cell slot
//   4. k_emit_ghosts   one wave per extended cell: sort the source cell's
//                      atoms by index in registers, write ghost images and
//                      the extended cells' atom records (coordinates,
//                      extended index, type) in cell order, inside a cell
//                      grouped by sub-cell along the third cell vector (with
//                      each sub-cell's first record, for the neighbour
//                      search's windows); the wave of a cell's own
//                      (non-ghost) image stores the cell's sorted atom list
//                      and clears its count for the next build
// Floating-point expressions that decide cell membership and image
// coordinates are evaluated without contraction, in the reference's order.
#include "dpmi_common.h"
#include "geometry.h"
#include "kernels.h"

namespace dpmi {

namespace {

__device__ __forceinline__ int image_shift(int idx, int n) {
  // neighbor_list.cc:730-745
  int s = 0;
  if (idx < 0) {
    s = 1;
    while (idx + s * n < 0) ++s;
  } else if (idx >= n) {
    s = -1;
    while (idx + s * n >= n) --s;
  }
  return s;
}

__global__ void k_wrap_bin(const double* __restrict__ coord_in, const int* __restrict__ type_in,
                           const int* __restrict__ frame_of, const FrameGeom* __restrict__ geom, int ntot,
                           double* __restrict__ coord_ext, int* __restrict__ type_ext, int* __restrict__ map_ext,
                           int* __restrict__ cell_of, int* __restrict__ cell_slot, int* __restrict__ cell_cnt,
                           int do_wrap, double* __restrict__ force_clear) {
#pragma clang fp contract(off)
  const int gidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (gidx >= ntot) return;
  if (force_clear) {  // the step's force accumulator, summed into from the force kernel on
#pragma unroll
    for (int d = 0; d < 3; ++d) force_clear[3 * (size_t)gidx + d] = 0.;
  }
  const FrameGeom& g = geom[frame_of[gidx]];
  const int i = gidx - g.loc_off;
  double p[3] = {coord_in[3 * gidx + 0], coord_in[3 * gidx + 1], coord_in[3 * gidx + 2]};
  if (do_wrap) {
    double f[3];
    for (int d = 0; d < 3; ++d)
      f[d] = p[0] * g.rec[d * 3 + 0] + p[1] * g.rec[d * 3 + 1] + p[2] * g.rec[d * 3 + 2];
    for (int d = 0; d < 3; ++d) {
      // fmod(f, 1.0) exactly: f - trunc(f) is representable, and fmod's
      // zero carries the sign of f (the library routine loops)
      f[d] = copysign(f[d] - trunc(f[d]), f[d]);
      if (f[d] < 0.) f[d] += 1.;
    }
    for (int d = 0; d < 3; ++d) p[d] = f[0] * g.t[0 * 3 + d] + f[1] * g.t[1 * 3 + d] + f[2] * g.t[2 * 3 + d];
  }
  const size_t e = (size_t)g.ext_off + i;
  coord_ext[3 * e + 0] = p[0];
  coord_ext[3 * e + 1] = p[1];
  coord_ext[3 * e + 2] = p[2];
  type_ext[e] = type_in[gidx];
  map_ext[e] = i;
  // bin on the wrapped coordinate (neighbor_list.cc:88-113)
  int id[3];
  for (int d = 0; d < 3; ++d) {
    const double fr = p[0] * g.rec[d * 3 + 0] + p[1] * g.rec[d * 3 + 1] + p[2] * g.rec[d * 3 + 2];
    const double cs = 1. / g.nc[d];
    int k = (int)((fr - 0.) / cs);
    if (fr - 0. < 0.) k--;
    if (k < 0) k = 0;
    if (k >= g.nc[d]) k = g.nc[d] - 1;
    id[d] = k;
  }
  const int c = (id[0] * g.nc[1] + id[1]) * g.nc[2] + id[2];
  cell_of[gidx] = c;
  cell_slot[gidx] = atomicAdd(&cell_cnt[g.cell_off + c], 1);
}

// Block-wide exclusive scan of n ints (one workgroup, blockDim = 1024): each
// thread sums a contiguous chunk, the chunk sums are scanned within each
// wave (shuffles) and across the 16 wave totals, then each thread writes
// its chunk.  Returns the total.
__device__ int block_exclusive_scan(const int* __restrict__ in, int* __restrict__ out, int n, int* sh) {
  const int t = threadIdx.x;
  const int nt = blockDim.x;
  const int lane = t & (kWave - 1), w = t / kWave, nw = nt / kWave;
  const int chunk = (n + nt - 1) / nt;
  const int b = t * chunk;
  const int e = min(n, b + chunk);
  int s = 0;
  for (int k = b; k < e; ++k) s += in[k];
  int inc = s;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int v = __shfl_up(inc, off, kWave);
    if (lane >= off) inc += v;
  }
  if (lane == kWave - 1) sh[w] = inc;
  __syncthreads();
  if (w == 0) {
    int tot = lane < nw ? sh[lane] : 0;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int v = __shfl_up(tot, off, kWave);
      if (lane >= off) tot += v;
    }
    if (lane < nw) sh[kWave + lane] = tot;  // inclusive prefix of the wave totals
  }
  __syncthreads();
  int run = inc - s + (w > 0 ? sh[kWave + w - 1] : 0);
  const int total = sh[kWave + nw - 1];
  for (int k = b; k < e; ++k) {
    const int v = in[k];
    out[k] = run;
    run += v;
  }
  __syncthreads();
  return total;
}

// Places every atom at its slot of its cell (slots from the counting
// atomics, so the order inside a cell is arbitrary until k_emit_ghosts).
__global__ void k_scatter_cells(const int* __restrict__ frame_of, const FrameGeom* __restrict__ geom, int ntot,
                                const int* __restrict__ cell_of, const int* __restrict__ cell_slot,
                                const int* __restrict__ cell_start, int* __restrict__ cell_list) {
  const int gidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (gidx >= ntot) return;
  const FrameGeom& g = geom[frame_of[gidx]];
  const int c = g.cell_off + cell_of[gidx];
  cell_list[g.loc_off + cell_start[c] + cell_slot[gidx]] = gidx - g.loc_off;
}

__device__ __forceinline__ int src_cell(const FrameGeom& g, int e, int* cx, bool* ghost) {
  int c[3];
  c[2] = e % g.nx[2];
  const int r = e / g.nx[2];
  c[1] = r % g.nx[1];
  c[0] = r / g.nx[1];
  bool gh = false;
  int s[3];
  for (int d = 0; d < 3; ++d) {
    c[d] -= g.ng[d];
    gh = gh || c[d] < 0 || c[d] >= g.nc[d];
    s[d] = c[d] + image_shift(c[d], g.nc[d]) * g.nc[d];
    if (cx) cx[d] = c[d];
  }
  *ghost = gh;
  return (s[0] * g.nc[1] + s[1]) * g.nc[2] + s[2];
}

// One workgroup per frame: exclusive scan of the cell counts (cell_start),
// then the per-extended-cell counts (all atoms, and ghosts only) and their
// scans -> nall.  With cell_list set (frames of at most kScatterInScan
// atoms) the workgroup also places the frame's atoms at their cell slots,
// instead of k_scatter_cells: each thread issues the loads of up to four
// atoms before their stores, so one round of dependent loads covers the
// frame (for larger frames the chain repeats and the separate kernel wins).
constexpr int kScatterPerThread = 4;
constexpr int kScatterInScan = kScatterPerThread * 1024;
__global__ void __launch_bounds__(1024) k_scan_cells(const FrameGeom* __restrict__ geom, const int* __restrict__ cell_cnt,
                                                     int* __restrict__ cell_start, int* __restrict__ xcnt, int* __restrict__ xstart,
                                                     int* __restrict__ xgcnt, int* __restrict__ xgstart,
                                                     int* __restrict__ nall_out, int* __restrict__ flag,
                                                     const int* __restrict__ cell_of, const int* __restrict__ cell_slot,
                                                     int* __restrict__ cell_list) {
  __shared__ int sh[2 * kWave];
  const int f = blockIdx.x;
  const FrameGeom& g = geom[f];
  const int nc = g.nc[0] * g.nc[1] * g.nc[2];
  block_exclusive_scan(cell_cnt + g.cell_off, cell_start + g.cell_off, nc, sh);  // ends with a barrier
  if (cell_list) {
    for (int i0 = threadIdx.x; i0 < g.nloc; i0 += kScatterPerThread * blockDim.x) {
      int c[kScatterPerThread], s[kScatterPerThread];
#pragma unroll
      for (int u = 0; u < kScatterPerThread; ++u) {
        const int i = i0 + u * blockDim.x;
        c[u] = i < g.nloc ? cell_of[g.loc_off + i] : 0;
        s[u] = i < g.nloc ? cell_slot[g.loc_off + i] : 0;
      }
#pragma unroll
      for (int u = 0; u < kScatterPerThread; ++u) c[u] = cell_start[g.cell_off + c[u]];
#pragma unroll
      for (int u = 0; u < kScatterPerThread; ++u) {
        const int i = i0 + u * blockDim.x;
        if (i < g.nloc) cell_list[g.loc_off + c[u] + s[u]] = i;
      }
    }
  }
  const int n = g.nx[0] * g.nx[1] * g.nx[2];
  const int off = g.xcell_off;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    bool gh;
    const int s = src_cell(g, e, nullptr, &gh);
    const int c = cell_cnt[g.cell_off + s];
    xcnt[off + e] = c;
    xgcnt[off + e] = gh ? c : 0;
  }
  __syncthreads();
  const int tot = block_exclusive_scan(xcnt + off, xstart + off, n, sh);
  const int gtot = block_exclusive_scan(xgcnt + off, xgstart + off, n, sh);
  if (threadIdx.x == 0) {
    xstart[off + n] = tot;
    const int nall = g.nloc + gtot;
    nall_out[f] = nall;
    if (nall > g.ext_cap) atomicOr(flag, kFlagGhostCapacity);
  }
}

// Ascending order of one int per lane across the wave (bitonic network).
__device__ __forceinline__ int wave_sort_asc(int v, int lane) {
  for (int k = 2; k <= kWave; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int p = __shfl_xor(v, j, kWave);
      const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
      v = keep_min ? min(v, p) : max(v, p);
    }
  }
  return v;
}

__global__ void k_emit_ghosts(const int* __restrict__ xcell_frame, const FrameGeom* __restrict__ geom, int nx_tot,
                              const int* __restrict__ cell_start, const int* __restrict__ cell_list,
                              int* __restrict__ sorted_loc, int* __restrict__ cell_cnt,
                              const int* __restrict__ xcnt, const int* __restrict__ xstart,
                              const int* __restrict__ xgstart, double* __restrict__ coord_ext,
                              int* __restrict__ type_ext, int* __restrict__ map_ext,
                              CellAtom* __restrict__ cell_atoms, int* __restrict__ cell_sub) {
#pragma clang fp contract(off)
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (wave >= nx_tot) return;
  const FrameGeom& g = geom[xcell_frame[wave]];
  const int e = wave - g.xcell_off;
  if (e >= g.nx[0] * g.nx[1] * g.nx[2]) return;  // the frame's scan sentinel
  int c[3];
  bool gh;
  const int s = src_cell(g, e, c, &gh);
  const int cnt = xcnt[wave];
  const int x0 = xstart[wave];
  const size_t base = (size_t)g.loc_off + cell_start[g.cell_off + s];
  const int* src = cell_list + base;
  double sv[3] = {0., 0., 0.};
  if (gh) {
    double sh[3];
    for (int d = 0; d < 3; ++d) sh[d] = (double)image_shift(c[d], g.nc[d]);
    for (int d = 0; d < 3; ++d) sv[d] = sh[0] * g.t[0 * 3 + d] + sh[1] * g.t[1 * 3 + d] + sh[2] * g.t[2 * 3 + d];
  }
  // record of atom p, the k-th of its cell in ascending index order (the
  // reference's ghost order); ghost images are written as a side effect
  auto record = [&](int p, int k, bool write) {
    const size_t a = (size_t)g.ext_off + p;
    CellAtom rec;
    rec.x = coord_ext[3 * a + 0];
    rec.y = coord_ext[3 * a + 1];
    rec.z = coord_ext[3 * a + 2];
    rec.type = write ? type_ext[a] : 0;
    rec.j = p;
    if (gh) {
      const int gi = g.nloc + xgstart[wave] + k;
      rec.x = rec.x - sv[0];
      rec.y = rec.y - sv[1];
      rec.z = rec.z - sv[2];
      rec.j = gi;
      if (write && gi < g.ext_cap) {
        const size_t b = (size_t)g.ext_off + gi;
        coord_ext[3 * b + 0] = rec.x;
        coord_ext[3 * b + 1] = rec.y;
        coord_ext[3 * b + 2] = rec.z;
        type_ext[b] = rec.type;
        map_ext[b] = p;
      }
    } else if (write) {
      sorted_loc[base + k] = p;
    }
    return rec;
  };
  // the records are stored grouped by sub-cell along the third cell vector
  // (kAxisBins per cell), in index order inside a sub-cell
  const double q = (double)(e % g.nx[2]);  // the cell spans positions [q, q + 1)
  auto sub_of = [&](const CellAtom& r) {   // sub-cell of a record (0 below the cell, last above)
    const double u = geom_axis2_pos(g, r.x, r.y, r.z);
    int b = 0;
#pragma unroll
    for (int k = 1; k < kAxisBins; ++k) b += u >= q + k * (1.0 / kAxisBins) ? 1 : 0;
    return u == u ? b : kAxisBins - 1;  // NaN: last
  };
  int first[kAxisBins];  // first position of each sub-cell's records in the cell
  if (cnt <= kWave) {
    const int v = wave_sort_asc(lane < cnt ? src[lane] : 0x7fffffff, lane);
    CellAtom rec;
    int sb = kAxisBins;
    if (lane < cnt) {
      rec = record(v, lane, true);
      sb = sub_of(rec);
    }
    int pos = 0, run = 0;
#pragma unroll
    for (int b = 0; b < kAxisBins; ++b) {
      const unsigned long long m = __ballot(sb == b);
      first[b] = run;
      if (sb == b) pos = run + __popcll(m & ((1ull << lane) - 1ull));
      run += __popcll(m);
    }
    if (lane < cnt && x0 + pos < g.ext_cap) cell_atoms[g.ext_off + x0 + pos] = rec;
  } else {
    // rank by counting (indices in a cell are distinct)
    int below[kAxisBins];  // this lane's records in each sub-cell
#pragma unroll
    for (int b = 0; b < kAxisBins; ++b) below[b] = 0;
    for (int k = lane; k < cnt; k += kWave) {
      const int p = src[k];
      int r = 0;
      for (int m = 0; m < cnt; ++m) r += src[m] < p;
      const CellAtom rec = record(p, r, true);
      const int sb = sub_of(rec);
      int pos = 0;
      for (int m = 0; m < cnt; ++m) {
        const int pm = src[m];
        const int bm = sub_of(record(pm, 0, false));
        pos += (bm < sb || (bm == sb && pm < p)) ? 1 : 0;
      }
      if (x0 + pos < g.ext_cap) cell_atoms[g.ext_off + x0 + pos] = rec;
#pragma unroll
      for (int b = 0; b < kAxisBins; ++b) below[b] += sb == b ? 1 : 0;
    }
    int run = 0;
#pragma unroll
    for (int b = 0; b < kAxisBins; ++b) {
      int c = below[b];
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) c += __shfl_xor(c, off, kWave);
      first[b] = run;
      run += c;
    }
  }
  if (cell_sub) {
    int mine = 0;
#pragma unroll
    for (int b = 0; b < kAxisBins; ++b)
      if (lane == b) mine = first[b];
    if (lane < kAxisBins) cell_sub[(size_t)wave * kAxisBins + lane] = x0 + mine;
  }
  if (!gh && lane == 0) cell_cnt[g.cell_off + s] = 0;  // counts start at zero in the next build
}

}  // namespace

void launch_build_extended(const CellWork& w, hipStream_t st) {
  // the emission leaves every count at zero; a first build clears them
  if (!w.counts_clear) launch_fill_int(w.cell_cnt, 0, w.ncells_tot, st);  // a kernel, not a memset node
  // one wave per workgroup: the per-atom kernels spread over more compute units
  const int tpb = 64;
  hipLaunchKernelGGL(k_wrap_bin, dim3(ceil_div(w.ntot, tpb)), dim3(tpb), 0, st, w.coord_in, w.type_in, w.frame_of,
                     w.geom, w.ntot, w.coord_ext, w.type_ext, w.map_ext, w.cell_of, w.cell_slot, w.cell_cnt,
                     w.do_wrap, w.force_clear);
  DPMI_LAUNCH_CHECK();
  const bool in_scan = w.max_nloc > 0 && w.max_nloc <= kScatterInScan;
  hipLaunchKernelGGL(k_scan_cells, dim3(w.nframes), dim3(1024), 0, st, w.geom, w.cell_cnt, w.cell_start, w.xcnt,
                     w.xstart, w.xgcnt, w.xgstart, w.nall, w.flag, w.cell_of, w.cell_slot,
                     in_scan ? w.cell_list : nullptr);
  DPMI_LAUNCH_CHECK();
  if (!in_scan) {
    hipLaunchKernelGGL(k_scatter_cells, dim3(ceil_div(w.ntot, tpb)), dim3(tpb), 0, st, w.frame_of, w.geom, w.ntot,
                       w.cell_of, w.cell_slot, w.cell_start, w.cell_list);
    DPMI_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_emit_ghosts, dim3(ceil_div((long long)w.nx_tot * kWave, tpb)), dim3(tpb), 0, st,
                     w.xcell_frame, w.geom, w.nx_tot, w.cell_start, w.cell_list, w.sorted_loc, w.cell_cnt, w.xcnt,
                     w.xstart, w.xgstart, w.coord_ext, w.type_ext, w.map_ext, w.cell_atoms, w.cell_sub);
  DPMI_LAUNCH_CHECK();
}

}  // namespace dpmi

namespace dpmi {
namespace {
__global__ void k_normalize(double* __restrict__ coord, int natom, FrameGeom g) {
#pragma clang fp contract(off)
  // coord.cc:14-28
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natom) return;
  double p[3] = {coord[3 * i], coord[3 * i + 1], coord[3 * i + 2]};
  double f[3];
  for (int d = 0; d < 3; ++d) f[d] = p[0] * g.rec[d * 3 + 0] + p[1] * g.rec[d * 3 + 1] + p[2] * g.rec[d * 3 + 2];
  for (int d = 0; d < 3; ++d) {
    f[d] = fmod(f[d], 1.0);
    if (f[d] < 0.) f[d] += 1.;
  }
  for (int d = 0; d < 3; ++d) coord[3 * i + d] = f[0] * g.t[0 * 3 + d] + f[1] * g.t[1 * 3 + d] + f[2] * g.t[2 * 3 + d];
}
struct Box9f {
  float t[9], rec[9];
};
__global__ void k_normalize_f32(float* __restrict__ coord, int natom, Box9f b) {
#pragma clang fp contract(off)
  // coord.cc:14-28 in FP32, as the float instantiation computes it
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natom) return;
  const float p[3] = {coord[3 * i], coord[3 * i + 1], coord[3 * i + 2]};
  float f[3];
  for (int d = 0; d < 3; ++d) f[d] = p[0] * b.rec[d * 3 + 0] + p[1] * b.rec[d * 3 + 1] + p[2] * b.rec[d * 3 + 2];
  for (int d = 0; d < 3; ++d) {
    f[d] = fmodf(f[d], 1.f);
    if (f[d] < 0.f) f[d] += 1.f;
  }
  for (int d = 0; d < 3; ++d) coord[3 * i + d] = f[0] * b.t[0 * 3 + d] + f[1] * b.t[1 * 3 + d] + f[2] * b.t[2 * 3 + d];
}
}  // namespace

void launch_normalize_f32(float* coord, int natom, const float* boxt9, const float* rec9, hipStream_t st) {
  if (natom <= 0) return;
  Box9f b;
  for (int k = 0; k < 9; ++k) {
    b.t[k] = boxt9[k];
    b.rec[k] = rec9[k];
  }
  hipLaunchKernelGGL(k_normalize_f32, dim3(ceil_div(natom, 256)), dim3(256), 0, st, coord, natom, b);
  DPMI_LAUNCH_CHECK();
}

namespace {
// Box of frame f from a device buffer (a captured step replayed for a new
// box): the frame's box and reciprocal box are replaced when the cell grid
// (cells and search layers per dimension) stays the same; otherwise the
// frame keeps its geometry and the grid change is flagged.
__global__ void k_geom_refresh(const double* __restrict__ box, FrameGeom* __restrict__ geom, int nframes, float rcut,
                               int* __restrict__ flag) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nframes) return;
  FrameGeom g = geom[f];
  geom_from_box(g, box + 9 * (size_t)f, rcut);
  const FrameGeom& c = geom[f];
  bool same = true;
  for (int d = 0; d < 3; ++d) same = same && g.nc[d] == c.nc[d] && g.ng[d] == c.ng[d] && g.nx[d] == c.nx[d];
  if (!same) {
    atomicOr(flag, kFlagGridChanged);
    return;
  }
  for (int k = 0; k < 9; ++k) {
    geom[f].t[k] = g.t[k];
    geom[f].rec[k] = g.rec[k];
  }
}
}  // namespace

void launch_geom_refresh(const double* box, FrameGeom* geom, int nframes, float rcut, int* flag, hipStream_t st) {
  if (nframes <= 0) return;
  hipLaunchKernelGGL(k_geom_refresh, dim3(ceil_div(nframes, 64)), dim3(64), 0, st, box, geom, nframes, rcut, flag);
  DPMI_LAUNCH_CHECK();
}

void launch_normalize(double* coord, int natom, const FrameGeom& g, hipStream_t st) {
  if (natom <= 0) return;
  hipLaunchKernelGGL(k_normalize, dim3(ceil_div(natom, 256)), dim3(256), 0, st, coord, natom, g);
  DPMI_LAUNCH_CHECK();
}
}  // namespace dpmi
