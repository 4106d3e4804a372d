// Periodic box, cell grid and per-frame layout shared by host and device.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

namespace dpmi {

// One frame of a packed batch.  Local atoms of frame f are the packed rows
// [loc_off, loc_off + nloc); its ghost-extended atoms occupy
// [ext_off, ext_off + ext_cap) of the extended arrays (frame-relative
// extended index 0..nall-1, locals first).  The cell grid has nc local cells
// per dimension, ng ghost layers on each side and nx = nc + 2 ng extended
// cells per dimension.
struct FrameGeom {
  double t[9];    // box, rows are cell vectors
  double rec[9];  // reciprocal box (fractional = rec . phys)
  int nc[3], ng[3], nx[3];
  int nloc, loc_off;
  int ext_off, ext_cap;
  int cell_off;   // offset of the frame's local cells in packed cell arrays
  int xcell_off;  // offset of the frame's extended cells (+ frame index for the scan sentinel)
  int pad[2];
};

// Distance between opposite faces of the box along each cell vector
// (SimulationRegion_Impl.h:364-372 toFaceDistance: volume / |a x b|).
// Host and device share these functions, so a box refreshed on the device
// gives the same bits as on the host.
__host__ __device__ inline void geom_face_distances(const double* m, double* tf) {
#pragma clang fp contract(off)
  double vol = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[6] * m[5]) +
               m[2] * (m[3] * m[7] - m[6] * m[4]);
  vol = ::fabs(vol);
  const double* a[3] = {m, m + 3, m + 6};
  for (int d = 0; d < 3; ++d) {
    const double* u = a[(d + 1) % 3];
    const double* v = a[(d + 2) % 3];
    const double c0 = u[1] * v[2] - u[2] * v[1];
    const double c1 = u[2] * v[0] - u[0] * v[2];
    const double c2 = u[0] * v[1] - u[1] * v[0];
    tf[d] = vol * (1. / ::sqrt(c0 * c0 + c1 * c1 + c2 * c2));
  }
}

// Box inverse and cell grid with the same operation order as the reference
// (region.cc:40-85, SimulationRegion_Impl.h:364-372, coord.cc:69-108,
// neighbor_list.cc:758-766) so that fractional coordinates, cell indices and
// image shifts agree bit for bit.
__host__ __device__ inline void geom_from_box(FrameGeom& g, const double* boxt, float rcut) {
#pragma clang fp contract(off)
  for (int k = 0; k < 9; ++k) g.t[k] = boxt[k];
  const double* m = g.t;
  double vol = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[6] * m[5]) +
               m[2] * (m[3] * m[7] - m[6] * m[4]);
  vol = ::fabs(vol);
  const double vi = 1.0 / vol;
  double* r = g.rec;
  r[0] = (m[4] * m[8] - m[7] * m[5]) * vi;
  r[4] = (m[0] * m[8] - m[6] * m[2]) * vi;
  r[8] = (m[0] * m[4] - m[3] * m[1]) * vi;
  r[1] = (-m[3] * m[8] + m[6] * m[5]) * vi;
  r[2] = (m[3] * m[7] - m[6] * m[4]) * vi;
  r[3] = (-m[1] * m[8] + m[7] * m[2]) * vi;
  r[5] = (-m[0] * m[7] + m[6] * m[1]) * vi;
  r[6] = (m[1] * m[5] - m[4] * m[2]) * vi;
  r[7] = (-m[0] * m[5] + m[3] * m[2]) * vi;
  double tf[3];
  geom_face_distances(m, tf);
  const double rc = rcut;
  for (int d = 0; d < 3; ++d) {
    int n = (int)(tf[d] / rc);
    if (n == 0) n = 1;
    const double cs = tf[d] / n;
    g.nc[d] = n;
    g.ng[d] = int(rc / cs) + 1;
    g.nx[d] = n + 2 * g.ng[d];
  }
}

// Position of a point along the third cell vector in units of extended
// cells: the extended cells with third index q span [q, q + 1) (the
// fractional coordinate of the binning, scaled by nc and shifted by ng).
// Sub-cell b of cell q holds the records with q + b / kAxisBins <= u <
// q + (b + 1) / kAxisBins (the first and last also what lies below and
// above the cell by rounding); cell-ordered records are grouped by it.
__host__ __device__ inline double geom_axis2_pos(const FrameGeom& g, double x, double y, double z) {
#pragma clang fp contract(off)
  const double s = x * g.rec[6] + y * g.rec[7] + z * g.rec[8];
  return s * g.nc[2] + g.ng[2];
}

// Sub-cells per cell along the third axis in the records' position table.
constexpr int kAxisBins = 8;

inline int geom_local_cells(const FrameGeom& g) { return g.nc[0] * g.nc[1] * g.nc[2]; }
inline int geom_ext_cells(const FrameGeom& g) { return g.nx[0] * g.nx[1] * g.nx[2]; }

// Upper bound of nall for a frame: every extended cell holds at most the
// densest local cell; callers size with the average density and a margin.
inline int geom_nall_estimate(const FrameGeom& g, int nloc, double margin) {
  const double ratio = (double)geom_ext_cells(g) / (double)geom_local_cells(g);
  return (int)std::ceil(nloc * ratio * margin) + 64;
}

}  // namespace dpmi
