// Neighbour search and formatting (gfx950): one wavefront per centre atom.
//
// The formatted row of centre i holds, per neighbour type t, the sel[t]
// nearest candidates in the order of the reference CPU path
// (fmt_nlist.cc:99-144): candidates are the j != i with FP64
// |x_i - x_j|^2 < rcut^2 (neighbor_list.cc:905-913) and FP32
// |(float)x_j - (float)x_i|^2 <= rcut^2 computed from rounded coordinates
// (fmt_nlist.cc:114-128), ordered by (type, FP32 r^2, index).  Both
// distance expressions are evaluated without contraction in the reference's
// order, so rows are bit-identical to the reference's.
//
// Per wave: candidates are gathered (from the 27 -- or (2 ng + 1)^3 --
// surrounding cells of the extended cell grid, or from a caller-given
// candidate row) with ballot compaction into an LDS key array
//   key = type << 56 | float_bits(r^2) << 24 | index      (index < 2^24)
// which is sorted in LDS (by rank up to 128 keys, else a bitonic network);
// per-type first/end positions then map slot s of type t to
// key[first[t] + s - sec[t]].  With one ghost layer (cells >= rc) the 27
// cells are 9 runs of 3 consecutive cells whose records are grouped by
// sub-cell along the third cell vector, so each run is cut to the sub-cells
// that can hold pairs within rc (and dropped at a lateral distance >= rc):
// about 3x the cutoff sphere's volume is tested instead of 7x.
#include <type_traits>

#include "dpmi_common.h"
#include "env_core.h"
#include "geometry.h"
#include "kernels.h"

namespace dpmi {

namespace {

constexpr int kMaxTypes = 128;

__device__ __forceinline__ unsigned long long make_key(int t, float r2, int j) {
  return ((unsigned long long)(unsigned)t << 56) | ((unsigned long long)__float_as_uint(r2) << 24) |
         (unsigned long long)(unsigned)j;
}

// Appends the keys of the lanes with keep set; returns the new count.
// Lane l ^ j's key: DPP quad_perm moves for j = 1, 2 (vector ALU, no LDS
// permute), a shuffle otherwise (j is wave-uniform).
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)v, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(v >> 32), CTRL, 0xF, 0xF, false);
  return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
__device__ __forceinline__ unsigned long long xor_lane(unsigned long long v, int j) {
  if (j == 1) return dpp_u64<0xB1>(v);  // quad_perm [1, 0, 3, 2]
  if (j == 2) return dpp_u64<0x4E>(v);  // quad_perm [2, 3, 0, 1]
  return __shfl_xor(v, j, kWave);
}

template <int CAP>
__device__ __forceinline__ int wave_append(unsigned long long* keys, int n, bool keep, unsigned long long key,
                                           int* flag) {
  const unsigned long long mask = __ballot(keep);
  const int pos = n + __popcll(mask & ((1ull << lane_id()) - 1ull));
  if (keep) {
    if (pos < CAP)
      keys[pos] = key;
    else
      atomicOr(flag, kFlagNlistCapacity);
  }
  return n + __popcll(mask);
}

// FP64 acceptance then FP32 formatting test for candidate j of centre i.
__device__ __forceinline__ bool accept(const double* ci, const double* cj, double rc2, float rc2f, float* fr2) {
#pragma clang fp contract(off)
  const double d0 = ci[0] - cj[0], d1 = ci[1] - cj[1], d2 = ci[2] - cj[2];
  const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
  if (!(r2 < rc2)) return false;
  const float f0 = (float)cj[0] - (float)ci[0];
  const float f1 = (float)cj[1] - (float)ci[1];
  const float f2 = (float)cj[2] - (float)ci[2];
  const float q = f0 * f0 + f1 * f1 + f2 * f2;
  *fr2 = q;
  return q <= rc2f;
}

// Bitonic sort of the wave's P = 64 KR keys in registers: element
// t = lane + 64 r is kr[r].  Stages with partner distance j < 64 exchange
// across lanes (64-bit shuffles), the others swap registers within a lane.
template <int KR>
__device__ __forceinline__ void wave_bitonic(unsigned long long (&kr)[KR]) {
  constexpr int P = 64 * KR;
  const int lane = lane_id();
#pragma unroll
  for (int k = 2; k <= P; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= kWave) {
        const int rj = j / kWave;
#pragma unroll
        for (int r = 0; r < KR; ++r) {
          if (r & rj) continue;
          const int t = lane + kWave * r;
          const bool up = (t & k) == 0;
          const unsigned long long a = kr[r], b = kr[r | rj];
          const bool sw = (a > b) == up;
          kr[r] = sw ? b : a;
          kr[r | rj] = sw ? a : b;
        }
      } else {
        const bool lower = (lane & j) == 0;
#pragma unroll
        for (int r = 0; r < KR; ++r) {
          const int t = lane + kWave * r;
          const bool up = (t & k) == 0;
          const unsigned long long o = xor_lane(kr[r], j);
          // the lower element of an ascending pair keeps the minimum
          const bool keep_min = lower == up;
          kr[r] = keep_min ? (o < kr[r] ? o : kr[r]) : (o > kr[r] ? o : kr[r]);
        }
      }
    }
  }
}

// Sorts keys[0, n) (LDS) ascending and sets tfirst/tend[t] to the first and
// end positions of type t's keys.  UNIQUE: no two keys are equal (one record
// per extended atom); up to 128 such keys are placed by their rank (the
// number of smaller keys: independent compares against broadcast LDS reads
// instead of the network's chain of dependent shuffles; at 256 keys the
// quadratic compare count costs more than the chain), larger or possibly
// repeated sets go through a bitonic network.
template <int CAP, bool UNIQUE>
__device__ void sort_keys(unsigned long long* keys, int n, int* tfirst, int* tend, int ntypes) {
  const int lane = lane_id();
  if (n > CAP) n = CAP;
  int P = 64;
  while (P < n) P <<= 1;
  __syncthreads();  // the appended keys are visible to every lane
  for (int t = lane; t < ntypes; t += kWave) {
    tfirst[t] = 0;
    tend[t] = 0;
  }
  // registers: P / 64 keys per lane (padding keys are all-ones, last)
  auto sort_in_registers = [&](auto kr_tag) {
    constexpr int KR = decltype(kr_tag)::value;
    unsigned long long kr[KR];
#pragma unroll
    for (int r = 0; r < KR; ++r) {
      const int t = lane + kWave * r;
      kr[r] = t < n ? keys[t] : ~0ull;
    }
    wave_bitonic<KR>(kr);
#pragma unroll
    for (int r = 0; r < KR; ++r) keys[lane + kWave * r] = kr[r];
  };
  auto rank_in_registers = [&](auto kr_tag) {
    constexpr int KR = decltype(kr_tag)::value;
    unsigned long long kr[KR];
    int rk[KR];
#pragma unroll
    for (int r = 0; r < KR; ++r) {
      const int t = lane + kWave * r;
      kr[r] = t < n ? keys[t] : ~0ull;
      rk[r] = 0;
    }
#pragma unroll 4
    for (int m = 0; m < n; ++m) {
      const unsigned long long km = keys[m];
#pragma unroll
      for (int r = 0; r < KR; ++r) rk[r] += km < kr[r] ? 1 : 0;
    }
    __syncthreads();  // every read of the unsorted keys is done
#pragma unroll
    for (int r = 0; r < KR; ++r)
      if (lane + kWave * r < n) keys[rk[r]] = kr[r];
  };
  if (UNIQUE && P <= 64)
    rank_in_registers(std::integral_constant<int, 1>{});
  else if (UNIQUE && P <= 128)
    rank_in_registers(std::integral_constant<int, 2>{});
  else if (P <= 64)
    sort_in_registers(std::integral_constant<int, 1>{});
  else if (P <= 128)
    sort_in_registers(std::integral_constant<int, 2>{});
  else if (P <= 256)
    sort_in_registers(std::integral_constant<int, 4>{});
  else if (P <= 512)
    sort_in_registers(std::integral_constant<int, 8>{});
  else {
    // larger capacities: the same network through LDS
    for (int t = n + lane; t < P; t += kWave) keys[t] = ~0ull;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int t = lane; t < P; t += kWave) {
          const int u = t ^ j;
          if (u > t) {
            const unsigned long long a = keys[t], b = keys[u];
            const bool up = (t & k) == 0;
            if ((a > b) == up) {
              keys[t] = b;
              keys[u] = a;
            }
          }
        }
        __syncthreads();
      }
    }
  }
  __syncthreads();
  for (int t = lane; t < n; t += kWave) {
    const int ty = (int)(keys[t] >> 56);
    if (t == 0 || (int)(keys[t - 1] >> 56) != ty) tfirst[ty] = t;
    if (t == n - 1 || (int)(keys[t + 1] >> 56) != ty) tend[ty] = t + 1;
  }
  __syncthreads();
}

struct NoSlotWork {
  __device__ void operator()(int, int) const {}
};

// Sorts keys[0, n) (LDS) and writes the formatted row; slot_work(s, j) runs
// for every slot s as its entry j is written (same lane).
template <int CAP, bool UNIQUE, typename F = NoSlotWork>
__device__ void sort_and_slot(unsigned long long* keys, int n, int* tfirst, int* tend, const int* sec, int ntypes,
                              int nnei, int* __restrict__ row, F slot_work = F()) {
  const int lane = lane_id();
  sort_keys<CAP, UNIQUE>(keys, n, tfirst, tend, ntypes);
  for (int s = lane; s < nnei; s += kWave) {
    int ty = 0;
    while (ty + 1 < ntypes && s >= sec[ty + 1]) ++ty;
    const int r = s - sec[ty];
    const int cnt = tend[ty] - tfirst[ty];
    const int j = r < cnt ? (int)(keys[tfirst[ty] + r] & 0xFFFFFFull) : -1;
    row[s] = j;
    slot_work(s, j);
  }
}

// Records of the run of extended cells (cx + dx, cy + dy, cz - 1 .. cz + 1)
// (ghost layers of one cell) that can lie within sqrt(rc2) of the centre ci:
// [*lo, *lo + *cnt) of the frame's cell-ordered records.  Each cell's records
// are grouped by sub-cell of their position u along the third cell vector
// (cell q of the extended grid spans [q, q + 1), sub-cell b of it
// [q + b / kAxisBins, q + (b + 1) / kAxisBins)); sub[q * kAxisBins + b] is
// the first record of sub-cell b of cell q.  Records before the window's
// first sub-cell lie below u_i - h, records from the sub-cell after its last
// one above u_i + h.  The window is widened by margins far above the
// rounding of u and of the binning, so the records left out are ones that
// fail the distance test.  Orthorhombic
// boxes: h follows from the lateral gap to the run's cells (rc^2 - d0^2 -
// d1^2); other boxes: h = rc and only runs at lateral gap >= rc are dropped.
__device__ __forceinline__ void run_window(const FrameGeom& g, const double* ci, int cx, int cy, int cz, int dx,
                                           int dy, double rc2, const int* __restrict__ xs,
                                           const int* __restrict__ sub, int* lo_out, int* cnt_out) {
#pragma clang fp contract(off)
  const double* R = g.rec;
  const double s0 = ci[0] * R[0] + ci[1] * R[1] + ci[2] * R[2];
  const double s1 = ci[0] * R[3] + ci[1] * R[4] + ci[2] * R[5];
  const double ui = geom_axis2_pos(g, ci[0], ci[1], ci[2]);
  const double n0 = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
  const double n1 = sqrt(R[3] * R[3] + R[4] * R[4] + R[5] * R[5]);
  const double n2 = sqrt(R[6] * R[6] + R[7] * R[7] + R[8] * R[8]);
  // perpendicular gap (fractional, less a margin) between ci and the slab of
  // the neighbouring cell layer along one axis
  auto gap = [](int d, int c, int n, double s) {
    const double f = d > 0 ? (double)(c + 1) / n - s : d < 0 ? s - (double)c / n : 0.;
    return fmax(f - 1e-10, 0.);
  };
  const double d0 = gap(dx, cx, g.nc[0], s0) / n0;
  const double d1 = gap(dy, cy, g.nc[1], s1) / n1;
  const double* T = g.t;
  const bool ortho = T[1] == 0. && T[2] == 0. && T[3] == 0. && T[5] == 0. && T[6] == 0. && T[7] == 0.;
  const double rc2m = rc2 * (1. + 1e-9);
  double h;
  bool live;
  if (ortho) {
    const double rest = rc2m - d0 * d0 - d1 * d1;
    live = rest > 0.;
    h = sqrt(fmax(rest, 0.));
  } else {
    const double dm = fmax(d0, d1);
    live = dm * dm < rc2m;
    h = sqrt(rc2m);
  }
  const double hu = h * n2 * g.nc[2] + 1e-9;
  const double ulo = ui - hu, uhi = ui + hu;
  const double flo = floor(ulo), fhi = floor(uhi);
  const int q0 = cz + g.ng[2] - 1, q2 = q0 + 2;
  const int eb = ((cx + g.ng[0] + dx) * g.nx[1] + (cy + g.ng[1] + dy)) * g.nx[2];
  int lo = 0, cnt = 0;
  if (live && flo <= (double)q2 && fhi >= (double)q0) {
    if (flo < (double)q0) {
      lo = xs[eb + q0];
    } else {
      const int q = (int)flo;
      const int b = min(max((int)((ulo - flo) * kAxisBins), 0), kAxisBins - 1);
      lo = sub[(size_t)(eb + q) * kAxisBins + b];
    }
    int hi;
    if (fhi > (double)q2) {
      hi = xs[eb + q2 + 1];
    } else {
      const int q = (int)fhi;
      const int b = min(max((int)((uhi - fhi) * kAxisBins), 0), kAxisBins - 1) + 1;
      hi = b == kAxisBins ? xs[eb + q + 1] : sub[(size_t)(eb + q) * kAxisBins + b];
    }
    cnt = max(hi - lo, 0);
  }
  *lo_out = lo;
  *cnt_out = cnt;
}

// Engine path: candidates from the extended cell grid.  At most 64 VGPRs
// (eight waves per SIMD): the kernel waits on memory and LDS, not on issue.
template <int CAP>
__global__ void __launch_bounds__(64, 8) k_cell_nlist(const int* __restrict__ frame_of, const FrameGeom* __restrict__ geom,
                                                   int ntot, const double* __restrict__ coord_ext,
                                                   const int* __restrict__ type_ext, const int* __restrict__ cell_of,
                                                   const int* __restrict__ xstart, const CellAtom* __restrict__ cell_atoms,
                                                   const int* __restrict__ cell_sub, const int* __restrict__ sec_g, int ntypes, int nnei, double rc2,
                                                   float rc2f, int* __restrict__ nlist, int* __restrict__ ncand,
                                                   int* __restrict__ flag, const double* __restrict__ avg,
                                                   const double* __restrict__ stdv, double rmin, double rmax,
                                                   float* __restrict__ em_f, double* __restrict__ em_d) {
  __shared__ unsigned long long keys[CAP];
  __shared__ int tfirst[kMaxTypes], tend[kMaxTypes], sec[kMaxTypes + 1];
  __shared__ int cstart[kWave], cpre[kWave + 1];
  const int gi = blockIdx.x;
  const int lane = lane_id();
  if (gi >= ntot) return;
  for (int t = lane; t <= ntypes; t += kWave) sec[t] = sec_g[t];
  const FrameGeom& g = geom[frame_of[gi]];
  const int i = gi - g.loc_off;
  const size_t eo = (size_t)g.ext_off;
  const double ci[3] = {coord_ext[3 * (eo + i) + 0], coord_ext[3 * (eo + i) + 1], coord_ext[3 * (eo + i) + 2]};
  const int ti = type_ext[eo + i];
  const int c = cell_of[gi];
  const int cz = c % g.nc[2], cy = (c / g.nc[2]) % g.nc[1], cx = c / (g.nc[2] * g.nc[1]);
  const int r0 = g.ng[0], r1 = g.ng[1], r2 = g.ng[2];
  const int w1 = 2 * r1 + 1, w2 = 2 * r2 + 1;
  const int nsearch = (2 * r0 + 1) * w1 * w2;
  int n = 0;
  // candidate test shared by both walks: record `slot` of the frame's
  // cell-ordered extended atoms
  auto consider = [&](bool valid, int slot) {
    bool keep = false;
    unsigned long long key = 0;
    if (valid) {
      // slots and indices beyond the frame's capacity exist only after a
      // ghost-capacity overflow (already flagged): skip them
      CellAtom rec;
      rec.j = -1;
      if (slot < g.ext_cap) rec = cell_atoms[eo + slot];
      const int j = rec.j;
      const int tj = (j >= 0 && j < g.ext_cap) ? rec.type : -1;
      if (j != i && tj >= 0) {
        const double cj[3] = {rec.x, rec.y, rec.z};
        float fr2;
        if (accept(ci, cj, rc2, rc2f, &fr2)) {
          // out-of-range keys are flagged and dropped, so that the
          // per-type tables below are indexed by valid types only
          if (j >= (1 << 24)) atomicOr(flag, kFlagIndexRange);
          if (tj >= ntypes) atomicOr(flag, kFlagTypeRange);
          keep = j < (1 << 24) && tj < ntypes;
          key = make_key(tj, fr2, j);
        }
      }
    }
    n = wave_append<CAP>(keys, n, keep, key, flag);
  };
  const int nruns = (2 * r0 + 1) * w1;  // runs of w2 consecutive cells (z fastest)
  if (ti >= 0 && cell_sub && r0 == 1 && r1 == 1 && r2 == 1) {
    // 9 runs of 3 cells; lane r < 9 bounds run r to the records that can lie
    // within rc (a window along the third cell vector, the whole run dropped
    // when its lateral distance reaches rc), then the wave walks the
    // windows' records as one list of total candidates, 64 at a time
    int lo = 0, cnt = 0;
    if (lane < 9) run_window(g, ci, cx, cy, cz, lane / 3 - 1, lane % 3 - 1, rc2, xstart + g.xcell_off,
                             cell_sub + (size_t)g.xcell_off * kAxisBins, &lo, &cnt);
    int inc = cnt;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const int v = __shfl_up(inc, off, kWave);
      if (lane >= off) inc += v;
    }
    const int pre = inc - cnt;  // the run's first position in the list
    const int total = __builtin_amdgcn_readlane(inc, 8);
    const int dl = lo - pre;    // record of position p of the run: p + dl
    int rp[9], rd[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      rp[r] = __builtin_amdgcn_readlane(pre, r);
      rd[r] = __builtin_amdgcn_readlane(dl, r);
    }
    for (int p0 = 0; p0 < total; p0 += kWave) {
      const int p = p0 + lane;
      int slot = p + rd[0];
#pragma unroll
      for (int r = 1; r < 9; ++r) slot = p >= rp[r] ? p + rd[r] : slot;  // empty runs are overtaken
      consider(p < total, slot);
    }
  } else if (ti >= 0 && nruns <= 32) {
    // each run's records are contiguous: 64 / nruns lanes walk each run
    // (9 runs of 3 cells: 7 lanes per run), no search for a candidate's cell
    const int lpr = kWave / nruns, run = lane / lpr, sub = lane - run * lpr;
    int st = 0, cnt = 0;
    if (run < nruns) {
      const int dx = run / w1 - r0, dy = run % w1 - r1;
      const int e = ((cx + g.ng[0] + dx) * g.nx[1] + (cy + g.ng[1] + dy)) * g.nx[2] + (cz + g.ng[2] - r2);
      st = xstart[g.xcell_off + e];
      cnt = xstart[g.xcell_off + e + w2] - st;
    }
    int iters = (cnt + lpr - 1) / lpr;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) iters = max(iters, __shfl_xor(iters, off, kWave));
    for (int it = 0; it < iters; ++it) {
      const int p = it * lpr + sub;
      consider(run < nruns && p < cnt, st + p);
    }
  } else if (ti >= 0) {
    for (int cb = 0; cb < nsearch; cb += kWave) {
      const int k = cb + lane;
      int cnt = 0, st = 0;
      if (k < nsearch) {
        const int dz = k % w2 - r2, dy = (k / w2) % w1 - r1, dx = k / (w2 * w1) - r0;
        const int e = ((cx + g.ng[0] + dx) * g.nx[1] + (cy + g.ng[1] + dy)) * g.nx[2] + (cz + g.ng[2] + dz);
        st = xstart[g.xcell_off + e];
        cnt = xstart[g.xcell_off + e + 1] - st;
      }
      // inclusive scan of cnt over the wave
      int inc = cnt;
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const int v = __shfl_up(inc, off, kWave);
        if (lane >= off) inc += v;
      }
      cstart[lane] = st;
      cpre[lane + 1] = inc;
      if (lane == 0) cpre[0] = 0;
      const int total = __shfl(inc, kWave - 1, kWave);
      __syncthreads();
      for (int p0 = 0; p0 < total; p0 += kWave) {
        const int p = p0 + lane;
        int lo = 0;
        if (p < total) {
          int hi = kWave;  // find cell q with cpre[q] <= p < cpre[q+1]
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (cpre[mid] <= p)
              lo = mid;
            else
              hi = mid;
          }
        }
        // candidates of a cell are consecutive records (coalesced reads)
        consider(p < total, cstart[lo] + (p - cpre[lo]));
      }
      __syncthreads();
    }
  }
  if (ncand && lane == 0) ncand[gi] = n;
  int* row = nlist + (size_t)gi * nnei;
  // environment-matrix rows of the slots as the lane formats them
  if (em_f) {
    sort_and_slot<CAP, true>(keys, n, tfirst, tend, sec, ntypes, nnei, row, [&](int s, int j) {
      envc::slot_em<float>(ci, coord_ext + 3 * (eo + (j >= 0 ? j : i)), j, ti, s, nnei, avg, stdv, rmin, rmax,
                           em_f + ((size_t)gi * nnei + s) * 4);
    });
  } else if (em_d) {
    sort_and_slot<CAP, true>(keys, n, tfirst, tend, sec, ntypes, nnei, row, [&](int s, int j) {
      envc::slot_em<double>(ci, coord_ext + 3 * (eo + (j >= 0 ? j : i)), j, ti, s, nnei, avg, stdv, rmin, rmax,
                            em_d + ((size_t)gi * nnei + s) * 4);
    });
  } else {
    sort_and_slot<CAP, true>(keys, n, tfirst, tend, sec, ntypes, nnei, row);
  }
}

// Drop-in path: candidates from caller rows (InputNlist), one frame stride.
// Rows are jlist + r * jstride, or firstneigh[r] when firstneigh is given
// (InputNlist with device row pointers).
template <int CAP, typename TC>
__global__ void __launch_bounds__(64) k_row_format(int inum, const int* __restrict__ ilist,
                                                   const int* __restrict__ numneigh, const int* __restrict__ jlist,
                                                   const int* const* __restrict__ firstneigh, int jstride,
                                                   const TC* __restrict__ coord, const int* __restrict__ type,
                                                   int nloc, int nall, const int* __restrict__ sec_g, int ntypes,
                                                   int nnei, float rc2f, int* __restrict__ nlist,
                                                   int* __restrict__ flag) {
  __shared__ unsigned long long keys[CAP];
  __shared__ int tfirst[kMaxTypes], tend[kMaxTypes], sec[kMaxTypes + 1];
  const int r = blockIdx.x;
  const int lane = lane_id();
  if (r >= inum) return;
  for (int t = lane; t <= ntypes; t += kWave) sec[t] = sec_g[t];
  const int f = r / nloc;
  const int i = ilist[r];
  if (i < 0 || i >= nloc) {  // a centre outside the frame: flagged, its row is not written
    if (lane == 0) atomicOr(flag, kFlagIndexRange);
    return;
  }
  const TC* cf = coord + (size_t)f * nall * 3;
  const int* tf = type + (size_t)f * nall;
  const double ci[3] = {(double)cf[3 * i], (double)cf[3 * i + 1], (double)cf[3 * i + 2]};
  int cnt = numneigh[r];
  if (!firstneigh && cnt > jstride) {  // a row longer than its stride: flagged, read up to the stride
    if (lane == 0) atomicOr(flag, kFlagNlistCapacity);
    cnt = jstride;
  }
  const int* cand = firstneigh ? firstneigh[r] : jlist + (size_t)r * jstride;
  __syncthreads();
  int n = 0;
  for (int p0 = 0; p0 < cnt; p0 += kWave) {
    const int p = p0 + lane;
    bool keep = false;
    unsigned long long key = 0;
    if (p < cnt) {
      const int j = cand[p];
      // an index outside the frame is a caller error: flagged, never read
      if (j < 0 || j >= nall) atomicOr(flag, kFlagIndexRange);
      const int tj = (j >= 0 && j < nall) ? tf[j] : -1;
      if (tj >= 0) {
#pragma clang fp contract(off)
        const float f0 = (float)cf[3 * j + 0] - (float)ci[0];
        const float f1 = (float)cf[3 * j + 1] - (float)ci[1];
        const float f2 = (float)cf[3 * j + 2] - (float)ci[2];
        const float q = f0 * f0 + f1 * f1 + f2 * f2;
        if (q <= rc2f) {
          if (j >= (1 << 24)) atomicOr(flag, kFlagIndexRange);
          if (tj >= ntypes) atomicOr(flag, kFlagTypeRange);
          keep = j < (1 << 24) && tj < ntypes;
          key = make_key(tj, q, j);
        }
      }
    }
    n = wave_append<CAP>(keys, n, keep, key, flag);
  }
  __syncthreads();
  sort_and_slot<CAP, false>(keys, n, tfirst, tend, sec, ntypes, nnei, nlist + ((size_t)f * nloc + i) * nnei);
}

// Brute-force candidate rows in ascending j (the build_nlist contract):
// one wave per row, lanes over j, ballot compaction keeps the order.
// T: coordinate and distance type (the reference's FPTYPE).
template <typename T>
__global__ void __launch_bounds__(256) k_brute_nlist(const T* __restrict__ c, int nloc, int nall, int mem_size,
                                                     T rc2, int nframes, const int* __restrict__ type,
                                                     int* __restrict__ ilist, int* __restrict__ numneigh,
                                                     int* __restrict__ jlist, int* const* __restrict__ firstneigh) {
#pragma clang fp contract(off)
  const int r = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = lane_id();
  if (r >= nframes * nloc) return;
  const int i = r % nloc, f = r / nloc;
  const T* cf = c + (size_t)f * nall * 3;
  const int* tf = type ? type + (size_t)f * nall : nullptr;
  const T ci[3] = {cf[3 * i], cf[3 * i + 1], cf[3 * i + 2]};
  const bool ci_ok = !tf || tf[i] >= 0;
  int n = 0;
  int* out = firstneigh ? firstneigh[r] : jlist + (size_t)r * mem_size;
  for (int j0 = 0; j0 < nall; j0 += kWave) {
    const int j = j0 + lane;
    bool keep = false;
    if (j < nall && j != i && ci_ok && (!tf || tf[j] >= 0)) {
      const T d0 = ci[0] - cf[3 * j], d1 = ci[1] - cf[3 * j + 1], d2 = ci[2] - cf[3 * j + 2];
      keep = d0 * d0 + d1 * d1 + d2 * d2 < rc2;
    }
    const unsigned long long m = __ballot(keep);
    const int pos = n + __popcll(m & ((1ull << lane) - 1ull));
    if (keep && pos < mem_size) out[pos] = j;
    n += __popcll(m);
  }
  if (lane == 0) {
    ilist[r] = i;
    numneigh[r] = n;  // may exceed mem_size: the host reports the capacity error
  }
}

}  // namespace

#define DPMI_CAP_DISPATCH(CAPV, ...)                 \
  switch (CAPV) {                                    \
    case 128: { constexpr int CAP = 128; __VA_ARGS__; } break; \
    case 256: { constexpr int CAP = 256; __VA_ARGS__; } break; \
    case 512: { constexpr int CAP = 512; __VA_ARGS__; } break; \
    case 1024: { constexpr int CAP = 1024; __VA_ARGS__; } break; \
    case 2048: { constexpr int CAP = 2048; __VA_ARGS__; } break; \
    default: { constexpr int CAP = 4096; __VA_ARGS__; } break; \
  }

int nlist_row_capacity(int want) {
  int c = 128;
  while (c < want && c < 4096) c <<= 1;
  return c;
}

void launch_cell_nlist(const CellWork& w, const NlistArgs& a, hipStream_t st) {
  DPMI_CAP_DISPATCH(a.cap, hipLaunchKernelGGL(k_cell_nlist<CAP>, dim3(w.ntot), dim3(kWave), 0, st, w.frame_of,
                                              w.geom, w.ntot, w.coord_ext, w.type_ext, w.cell_of, w.xstart,
                                              w.cell_atoms, w.cell_sub, a.sec, a.ntypes, a.nnei, a.rc2, a.rc2f, a.nlist,
                                              a.ncand, w.flag, a.avg, a.stdv, a.rmin, a.rmax, a.em_f, a.em_d));
  DPMI_LAUNCH_CHECK();
}

template <typename TC>
void launch_row_format(int inum, const int* ilist, const int* numneigh, const int* jlist, int jstride,
                       const TC* coord, const int* type, int nloc, int nall, const NlistArgs& a, int* flag,
                       hipStream_t st, const int* const* firstneigh) {
  if (inum <= 0) return;
  DPMI_CAP_DISPATCH(a.cap, hipLaunchKernelGGL((k_row_format<CAP, TC>), dim3(inum), dim3(kWave), 0, st, inum, ilist,
                                              numneigh, jlist, firstneigh, jstride, coord, type, nloc, nall, a.sec,
                                              a.ntypes, a.nnei, a.rc2f, a.nlist, flag));
  DPMI_LAUNCH_CHECK();
}
template void launch_row_format<double>(int, const int*, const int*, const int*, int, const double*, const int*, int,
                                        int, const NlistArgs&, int*, hipStream_t, const int* const*);
template void launch_row_format<float>(int, const int*, const int*, const int*, int, const float*, const int*, int,
                                       int, const NlistArgs&, int*, hipStream_t, const int* const*);

template <typename T>
void launch_brute_nlist(const T* c, int nloc, int nall, int mem_size, T rc2, int nframes, const int* type,
                        int* ilist, int* numneigh, int* jlist, hipStream_t st, int* const* firstneigh) {
  const long long rows = (long long)nframes * nloc;
  if (rows <= 0) return;
  hipLaunchKernelGGL(k_brute_nlist<T>, dim3(ceil_div(rows * kWave, 256)), dim3(256), 0, st, c, nloc, nall, mem_size,
                     rc2, nframes, type, ilist, numneigh, jlist, firstneigh);
  DPMI_LAUNCH_CHECK();
}
template void launch_brute_nlist<double>(const double*, int, int, int, double, int, const int*, int*, int*, int*,
                                         hipStream_t, int* const*);
template void launch_brute_nlist<float>(const float*, int, int, int, float, int, const int*, int*, int*, int*,
                                        hipStream_t, int* const*);

}  // namespace dpmi
