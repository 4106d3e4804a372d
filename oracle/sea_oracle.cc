// CPU restatement of the reference se_e2_a energy/force/virial path.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker the parity tests, smoke()
// and bench.py's cpu_baseline leg compare the HIP path against; the product
// library never links it and never falls back to it.
//
// Each function restates one reference routine (cited file:line, paths
// relative to the reference checkout) with the same floating-point operation
// order, so that integer outputs (ghost order, neighbour lists) are bit-exact
// and floating outputs agree to rounding.  Built with -ffp-contract=off.
//
// Pinning (tests/test_oracle_*.py):
//   * every function against the compiled reference library (oracle/_ref),
//     bit-exact for integers and to 1e-13 relative for floats;
//   * against the golden vectors of the reference's own tests
//     (source/lib/tests/test_*.cc, source/tests/pt/test_tabulate_fusion_se_a.py);
//   * against invariants: env-matrix derivative vs central finite
//     differences, brute-force O(N^2) minimum-image neighbour sets.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// ---------------------------------------------------------------- region --
// region.cc:40-85 (volume and reciprocal box by cofactors, volumei = 1/|V|)
struct Box {
  double t[9];
  double rec[9];
  double volume;
};

void box_init(Box& b, const double* boxt) {
  for (int i = 0; i < 9; ++i) b.t[i] = boxt[i];
  const double* m = b.t;
  double vol = m[0] * (m[4] * m[8] - m[7] * m[5]) -
               m[1] * (m[3] * m[8] - m[6] * m[5]) +
               m[2] * (m[3] * m[7] - m[6] * m[4]);
  vol = std::fabs(vol);
  b.volume = vol;
  const double vi = 1.0 / vol;
  double* r = b.rec;
  r[0] = (m[4] * m[8] - m[7] * m[5]) * vi;
  r[4] = (m[0] * m[8] - m[6] * m[2]) * vi;
  r[8] = (m[0] * m[4] - m[3] * m[1]) * vi;
  r[1] = (-m[3] * m[8] + m[6] * m[5]) * vi;
  r[2] = (m[3] * m[7] - m[6] * m[4]) * vi;
  r[3] = (-m[1] * m[8] + m[7] * m[2]) * vi;
  r[5] = (-m[0] * m[7] + m[6] * m[1]) * vi;
  r[6] = (m[1] * m[5] - m[4] * m[2]) * vi;
  r[7] = (-m[0] * m[5] + m[3] * m[2]) * vi;
}

// fractional = rec . phys  (region.cc:87-96, tensor_dot_vec)
inline void to_frac(const Box& b, const double* p, double* f) {
  for (int d = 0; d < 3; ++d)
    f[d] = p[0] * b.rec[d * 3 + 0] + p[1] * b.rec[d * 3 + 1] +
           p[2] * b.rec[d * 3 + 2];
}
// phys = box^T . fractional  (region.cc:98-108, tensor_t_dot_vec)
inline void to_phys(const Box& b, const double* f, double* p) {
  for (int d = 0; d < 3; ++d)
    p[d] = f[0] * b.t[0 * 3 + d] + f[1] * b.t[1 * 3 + d] +
           f[2] * b.t[2 * 3 + d];
}

// SimulationRegion_Impl.h:364-372 (volume * 1/sqrt(|a x b|^2))
void face_dist(const Box& b, double* dd) {
  const double* a0 = b.t;
  const double* a1 = b.t + 3;
  const double* a2 = b.t + 6;
  auto cross_norm_inv = [](const double* u, const double* v) {
    double c0 = u[1] * v[2] - u[2] * v[1];
    double c1 = u[2] * v[0] - u[0] * v[2];
    double c2 = u[0] * v[1] - u[1] * v[0];
    return 1. / std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  };
  dd[0] = b.volume * cross_norm_inv(a1, a2);
  dd[1] = b.volume * cross_norm_inv(a2, a0);
  dd[2] = b.volume * cross_norm_inv(a0, a1);
}

// neighbor_list.cc:730-745
int pbc_image_shift(int idx, int n) {
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

// ------------------------------------------------------------- tabulate --
// tabulate.cc:21-30 (index of the last knot interval below max)
template <typename T>
int high_tail_index(T lower, T upper, T mx, T s0, T s1) {
  const T edge = std::nextafter(mx, lower);
  return int((upper - lower) / s0) + int((edge - upper) / s1);
}

// tabulate.cc:46-73: segment index, local offset and linear-tail delta
template <typename T>
inline void table_locate(const T* info, T& x, int& idx, T& delta) {
  const T lower = info[0], upper = info[1], mx = info[2], s0 = info[3],
          s1 = info[4];
  const T x0 = x;
  delta = (T)0.;
  if (x < lower) {
    idx = 0;
    x = (T)0.;
    delta = x0 - lower;
  } else if (x < upper) {
    idx = (int)((x - lower) / s0);
    x -= (idx * s0 + lower);
  } else if (x < mx) {
    const int n0 = int((upper - lower) / s0);
    idx = n0 + (int)((x - upper) / s1);
    x -= ((idx - n0) * s1 + upper);
  } else {
    const int n0 = int((upper - lower) / s0);
    idx = high_tail_index(lower, upper, mx, s0, s1);
    x = mx - ((idx - n0) * s1 + upper);
    delta = x0 - mx;
  }
}

// tabulate.cc:122-158 (Horner forms of the quintic and its derivative)
template <typename T>
inline T quintic(const T* a, T x) {
  return a[0] + (a[1] + (a[2] + (a[3] + (a[4] + a[5] * x) * x) * x) * x) * x;
}
template <typename T>
inline T quintic_d(const T* a, T x) {
  return a[1] +
         ((T)2. * a[2] + ((T)3. * a[3] + ((T)4. * a[4] + (T)5. * a[5] * x) * x) * x) *
             x;
}

// True when neighbour jj is the first of the sorted zero-padding tail
// (tabulate.cc:188-201): same em_x as the last neighbour and a zero
// direction part.
template <typename T>
inline bool padding_tail(const T* ll, int nd, T x, T last) {
  (void)nd;
  return last == x && ll[1] == (T)0. && ll[2] == (T)0. && ll[3] == (T)0.;
}

template <typename T>
void tab_fwd(T* out, const T* table, const T* info, const T* em_x, const T* em,
             int nloc, int nnei, int M, bool sorted, int nd) {
  std::memset(out, 0, sizeof(T) * (size_t)nloc * nd * M);
  if (nnei <= 0) return;
#pragma omp parallel for
  for (int ii = 0; ii < nloc; ++ii) {
    const T last = em_x[(size_t)ii * nnei + nnei - 1];
    T* o = out + (size_t)ii * nd * M;
    for (int jj = 0; jj < nnei; ++jj) {
      const T* ll = em + ((size_t)ii * nnei + jj) * nd;
      T x = em_x[(size_t)ii * nnei + jj];
      const bool tail = sorted && padding_tail(ll, nd, x, last);
      int idx;
      T delta;
      table_locate(info, x, idx, delta);
      const T* row = table + (size_t)idx * M * 6;
      for (int kk = 0; kk < M; ++kk) {
        const T* a = row + 6 * kk;
        T g = quintic(a, x) + quintic_d(a, x) * delta;
        const T w = tail ? (T)(nnei - jj) * g : g;
        for (int mm = 0; mm < nd; ++mm) o[mm * M + kk] += w * ll[mm];
      }
      if (tail) break;
    }
  }
}

template <typename T>
void tab_grad(T* dx_out, T* dem_out, const T* table, const T* info,
              const T* em_x, const T* em, const T* dy, int nloc, int nnei,
              int M, bool sorted, int nd) {
  if (nnei <= 0) return;
  std::memset(dx_out, 0, sizeof(T) * (size_t)nloc * nnei);
  std::memset(dem_out, 0, sizeof(T) * (size_t)nloc * nnei * nd);
#pragma omp parallel for
  for (int ii = 0; ii < nloc; ++ii) {
    const T last = em_x[(size_t)ii * nnei + nnei - 1];
    const T* dyi = dy + (size_t)ii * nd * M;
    for (int jj = 0; jj < nnei; ++jj) {
      const size_t base = ((size_t)ii * nnei + jj) * nd;
      const T* ll = em + base;
      T x = em_x[(size_t)ii * nnei + jj];
      const bool tail = sorted && padding_tail(ll, nd, x, last);
      int idx;
      T delta;
      table_locate(info, x, idx, delta);
      const T* row = table + (size_t)idx * M * 6;
      T gx = (T)0.;
      for (int kk = 0; kk < M; ++kk) {
        const T* a = row + 6 * kk;
        const T gd = quintic_d(a, x);
        const T gv = quintic(a, x) + gd * delta;
        T proj = (T)0.;
        for (int mm = 0; mm < nd; ++mm) proj += ll[mm] * dyi[mm * M + kk];
        if (tail) {
          gx += gd * proj * (T)(nnei - jj);
          for (int mm = 0; mm < nd; ++mm)
            dem_out[base + mm] += gv * dyi[mm * M + kk] * (T)(nnei - jj);
        } else {
          gx += gd * proj;
          for (int mm = 0; mm < nd; ++mm)
            dem_out[base + mm] += gv * dyi[mm * M + kk];
        }
      }
      dx_out[(size_t)ii * nnei + jj] = gx;
      if (tail) break;
    }
  }
}

template <typename T>
void tab_grad_grad(T* dz_dy, const T* table, const T* info, const T* em_x,
                   const T* em, const T* dz_dx, const T* dz_dem, int nloc,
                   int nnei, int M, bool sorted, int nd) {
  std::memset(dz_dy, 0, sizeof(T) * (size_t)nloc * nd * M);
  if (nnei <= 0) return;
#pragma omp parallel for
  for (int ii = 0; ii < nloc; ++ii) {
    const T last = em_x[(size_t)ii * nnei + nnei - 1];
    T* o = dz_dy + (size_t)ii * nd * M;
    for (int jj = 0; jj < nnei; ++jj) {
      const size_t base = ((size_t)ii * nnei + jj) * nd;
      const T* ll = em + base;
      const T* hh = dz_dem + base;
      T x = em_x[(size_t)ii * nnei + jj];
      const T zx = dz_dx[(size_t)ii * nnei + jj];
      const bool tail = sorted && padding_tail(ll, nd, x, last);
      int idx;
      T delta;
      table_locate(info, x, idx, delta);
      const T* row = table + (size_t)idx * M * 6;
      const T rep = tail ? (T)(nnei - jj) : (T)1.;
      for (int kk = 0; kk < M; ++kk) {
        const T* a = row + 6 * kk;
        const T gd = quintic_d(a, x);
        const T gv = quintic(a, x) + gd * delta;
        for (int mm = 0; mm < nd; ++mm)
          o[mm * M + kk] += rep * (gv * hh[mm] + zx * gd * ll[mm]);
      }
      if (tail) break;
    }
  }
}

// ----------------------------------------------------------- env matrix --
// switcher.h:61-80 (quintic smooth switch on [rmin, rmax])
inline void smooth_switch(double& v, double& dv, double r, float rmin,
                          float rmax) {
  if (r < rmin) {
    v = 1.;
    dv = 0.;
  } else if (r < rmax) {
    const double u = (r - rmin) / (rmax - rmin);
    const double du = 1. / (rmax - rmin);
    v = u * u * u * (-6. * u * u + 15. * u - 10.) + 1.;
    dv = (3. * u * u * (-6. * u * u + 15. * u - 10.) +
          u * u * u * (-12. * u + 15.)) *
         du;
  } else {
    v = 0.;
    dv = 0.;
  }
}

// env_mat.cc:124-222 for one neighbour: 4 values and their 12 derivatives
// w.r.t. the neighbour displacement (row a holds d(value a)/d(x,y,z)).
inline void env_row(const double* rr, float rmin, float rmax, double* val,
                    double* der) {
  const double nr2 = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
  const double inr = 1. / std::sqrt(nr2);
  const double nr = nr2 * inr;
  const double inr2 = inr * inr;
  const double inr4 = inr2 * inr2;
  const double inr3 = inr4 * nr;
  double sw, dsw;
  smooth_switch(sw, dsw, nr, rmin, rmax);
  val[0] = 1. / nr;
  val[1] = rr[0] / nr2;
  val[2] = rr[1] / nr2;
  val[3] = rr[2] / nr2;
  for (int d = 0; d < 3; ++d)
    der[d] = rr[d] * inr3 * sw - val[0] * dsw * rr[d] * inr;
  for (int a = 1; a < 4; ++a) {
    for (int d = 0; d < 3; ++d) {
      double t = 2. * rr[a - 1] * rr[d] * inr4;
      if (d == a - 1) t -= inr2;
      der[a * 3 + d] = t * sw - val[a] * dsw * rr[d] * inr;
    }
  }
  for (int a = 0; a < 4; ++a) val[a] *= sw;
}

// One (type, dist, index) candidate of a neighbour row.
struct Cand {
  int type;
  float d2;
  int idx;
  bool operator<(const Cand& o) const {
    if (type != o.type) return type < o.type;
    if (d2 != o.d2) return d2 < o.d2;
    return idx < o.idx;
  }
};

// fmt_nlist.cc:99-144: keep typed candidates with float r^2 <= rc^2, sort by
// (type, float r^2, index), fill the first sel[t] slots of each type block.
void format_row(int* row, const double* coord, const int* type, int i_idx,
                const int* cand, int ncand, float rcut, const int* sec,
                int ntypes) {
  const int nnei = sec[ntypes];
  for (int k = 0; k < nnei; ++k) row[k] = -1;
  std::vector<Cand> c;
  c.reserve(ncand);
  const float rc2 = rcut * rcut;
  for (int k = 0; k < ncand; ++k) {
    const int j = cand[k];
    if (type[j] < 0) continue;
    float d[3];
    for (int dd = 0; dd < 3; ++dd)
      d[dd] = (float)coord[j * 3 + dd] - (float)coord[i_idx * 3 + dd];
    const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 <= rc2) c.push_back(Cand{type[j], r2, j});
  }
  std::sort(c.begin(), c.end());
  std::vector<int> fill(sec, sec + ntypes);
  for (const Cand& e : c) {
    if (fill[e.type] < sec[e.type + 1]) row[fill[e.type]++] = e.idx;
  }
}

}  // namespace

extern "C" {

// region.cc:104-108 + coord.cc:14-28: wrap into the primary cell.
void orc_normalize_coord(double* coord, int natom, const double* boxt) {
  Box b;
  box_init(b, boxt);
  for (int i = 0; i < natom; ++i) {
    double f[3];
    to_frac(b, coord + 3 * i, f);
    for (int d = 0; d < 3; ++d) {
      f[d] = std::fmod(f[d], 1.);
      if (f[d] < 0.) f[d] += 1.;
    }
    to_phys(b, f, coord + 3 * i);
  }
}

// coord.cc:69-108 (23 ints: nat_stt, ncell, ext_stt, ext_end, ngcell,
// cell_shift, cell_iter, loc_cellnum, total_cellnum)
void orc_compute_cell_info(int* ci, float rcut, const double* boxt) {
  Box b;
  box_init(b, boxt);
  double tf[3];
  face_dist(b, tf);
  for (int d = 0; d < 3; ++d) {
    ci[d] = 0;
    ci[3 + d] = (int)(tf[d] / rcut);
    if (ci[3 + d] == 0) ci[3 + d] = 1;
    const double cs = tf[d] / ci[3 + d];
    ci[12 + d] = int(rcut / cs) + 1;
    ci[6 + d] = -ci[12 + d];
    ci[9 + d] = ci[3 + d] + ci[12 + d];
    ci[15 + d] = ci[12 + d];
    ci[18 + d] = (int)(rcut / cs);
    if (ci[18 + d] * cs < rcut) ci[18 + d] += 1;
  }
  ci[21] = ci[3] * ci[4] * ci[5];
  ci[22] = (2 * ci[12] + ci[3]) * (2 * ci[13] + ci[4]) * (2 * ci[14] + ci[5]);
}

// coord.cc:31-66 -> neighbor_list.cc:43-148 (cell binning of the local
// atoms) and 747-847 (ghost emission): locals first, then for every ghost
// cell of the extended grid in (x, y, z) lexicographic order the atoms of
// its periodic source cell in ascending index order, shifted by the image
// vector.  Returns 1 when nall > mem_nall.
int orc_copy_coord(double* out_c, int* out_t, int* mapping, int* nall,
                   const double* in_c, const int* in_t, int nloc, int mem_nall,
                   float rcut, const double* boxt) {
  Box b;
  box_init(b, boxt);
  double tf[3];
  face_dist(b, tf);
  const double rc = rcut;
  int nc[3], ng[3];
  for (int d = 0; d < 3; ++d) {
    nc[d] = (int)(tf[d] / rc);
    if (nc[d] == 0) nc[d] = 1;
    const double cs = tf[d] / nc[d];
    ng[d] = int(rc / cs) + 1;
  }
  // bin local atoms (neighbor_list.cc:88-113 with nat = ext = [0, ncell))
  const int nloc_cells = nc[0] * nc[1] * nc[2];
  std::vector<std::vector<int>> cells(nloc_cells);
  for (int i = 0; i < nloc; ++i) {
    double f[3];
    to_frac(b, in_c + 3 * i, f);
    int id[3];
    for (int d = 0; d < 3; ++d) {
      const double csz = 1. / nc[d];
      id[d] = (int)((f[d] - 0.) / csz);
      if (f[d] - 0. < 0.) id[d]--;
      if (id[d] < 0) id[d] = 0;
      if (id[d] >= nc[d]) id[d] = nc[d] - 1;
    }
    cells[(id[0] * nc[1] + id[1]) * nc[2] + id[2]].push_back(i);
  }
  std::vector<double> oc(in_c, in_c + 3 * (size_t)nloc);
  std::vector<int> ot(in_t, in_t + nloc), om(nloc);
  for (int i = 0; i < nloc; ++i) om[i] = i;
  int c[3];
  for (c[0] = -ng[0]; c[0] < nc[0] + ng[0]; ++c[0]) {
    for (c[1] = -ng[1]; c[1] < nc[1] + ng[1]; ++c[1]) {
      for (c[2] = -ng[2]; c[2] < nc[2] + ng[2]; ++c[2]) {
        bool local = true;
        int src[3];
        double sh[3];
        for (int d = 0; d < 3; ++d) {
          local = local && c[d] >= 0 && c[d] < nc[d];
          const int s = pbc_image_shift(c[d], nc[d]);
          sh[d] = s;
          src[d] = c[d] + s * nc[d];
        }
        if (local) continue;
        double sv[3];
        to_phys(b, sh, sv);
        for (int p : cells[(src[0] * nc[1] + src[1]) * nc[2] + src[2]]) {
          for (int d = 0; d < 3; ++d) oc.push_back(in_c[3 * p + d] - sv[d]);
          ot.push_back(in_t[p]);
          om.push_back(p);
        }
      }
    }
  }
  *nall = (int)ot.size();
  if (*nall > mem_nall) return 1;
  std::copy(oc.begin(), oc.end(), out_c);
  std::copy(ot.begin(), ot.end(), out_t);
  std::copy(om.begin(), om.end(), mapping);
  return 0;
}

// neighbor_list.cc:876-928: all j != i of the frame with double
// |x_i - x_j|^2 < rcut^2 (rcut squared in float), ascending j.  Rows are
// jlist + row * mem_size.  Returns 1 if a row exceeds mem_size.
int orc_build_nlist(int* ilist, int* numneigh, int* jlist, int* max_list_size,
                    const double* c, int nloc, int nall, int mem_size,
                    float rcut, int nframes, const int* type) {
  *max_list_size = 0;
  const double rc2 = (double)(rcut * rcut);
  for (int row = 0; row < nframes * nloc; ++row) {
    const int i = row % nloc;
    const size_t off = (size_t)(row / nloc) * nall;
    ilist[row] = i;
    int n = 0;
    int* out = jlist + (size_t)row * mem_size;
    for (int j = 0; j < nall; ++j) {
      if (j == i) continue;
      if (type && (type[off + i] < 0 || type[off + j] < 0)) continue;
      double d[3];
      for (int dd = 0; dd < 3; ++dd)
        d[dd] = c[(off + i) * 3 + dd] - c[(off + j) * 3 + dd];
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (r2 < rc2) {
        if (n >= mem_size) {
          // count the rest to report the true row size
          int m = n;
          for (int j2 = j; j2 < nall; ++j2) {
            if (j2 == i) continue;
            if (type && (type[off + i] < 0 || type[off + j2] < 0)) continue;
            double e[3];
            for (int dd = 0; dd < 3; ++dd)
              e[dd] = c[(off + i) * 3 + dd] - c[(off + j2) * 3 + dd];
            if (e[0] * e[0] + e[1] * e[1] + e[2] * e[2] < rc2) ++m;
          }
          *max_list_size = m;
          return 1;
        }
        out[n++] = j;
      }
    }
    numneigh[row] = n;
    if (n > *max_list_size) *max_list_size = n;
  }
  return 0;
}

// fmt_nlist.cc:146-177 (one frame): formatted rows nlist[i_idx * nnei].
void orc_format_nlist(int* nlist, int inum, const int* ilist,
                      const int* numneigh, const int* jlist, int jstride,
                      const double* coord, const int* type, float rcut,
                      const int* sel, int ntypes) {
  std::vector<int> sec(ntypes + 1, 0);
  for (int t = 0; t < ntypes; ++t) sec[t + 1] = sec[t] + sel[t];
  const int nnei = sec[ntypes];
#pragma omp parallel for
  for (int r = 0; r < inum; ++r) {
    const int i = ilist[r];
    format_row(nlist + (size_t)i * nnei, coord, type, i,
               jlist + (size_t)r * jstride, numneigh[r], rcut, sec.data(),
               ntypes);
  }
}

// prod_env_mat.cc:15-132 with env_mat.cc:124-222: per row format the
// candidate list, compute the 4-vector environment row of each slot and its
// derivative, normalise by avg/std of the centre type.
void orc_prod_env_mat_a(double* em, double* em_deriv, double* rij, int* nlist,
                        const double* coord, const int* type, int inum,
                        const int* ilist, const int* numneigh,
                        const int* jlist, int jstride, const double* avg,
                        const double* std_, int nloc, int nall, int nframes,
                        float rcut, float rcut_smth, const int* sel,
                        int ntypes) {
  std::vector<int> sec(ntypes + 1, 0);
  for (int t = 0; t < ntypes; ++t) sec[t + 1] = sec[t] + sel[t];
  const int nnei = sec[ntypes];
  const int nem = nnei * 4;
  // rows of the input list, indexed by (frame, centre atom)
  std::vector<int> row_of((size_t)nframes * nloc, -1);
  for (int r = 0; r < inum; ++r) {
    const int f = r / nloc;
    row_of[(size_t)f * nloc + ilist[r]] = r;
  }
#pragma omp parallel for
  for (long long g = 0; g < (long long)nframes * nloc; ++g) {
    const int f = (int)(g / nloc), i = (int)(g % nloc);
    const double* c = coord + (size_t)f * nall * 3;
    const int* t = type + (size_t)f * nall;
    int* nl = nlist + (size_t)g * nnei;
    const int r = row_of[g];
    if (r >= 0)
      format_row(nl, c, t, i, jlist + (size_t)r * jstride, numneigh[r], rcut,
                 sec.data(), ntypes);
    else
      for (int k = 0; k < nnei; ++k) nl[k] = -1;
    double* e = em + (size_t)g * nem;
    double* ed = em_deriv + (size_t)g * nem * 3;
    double* rr = rij + (size_t)g * nnei * 3;
    std::vector<double> val(nem, 0.), der((size_t)nem * 3, 0.);
    for (int k = 0; k < nnei * 3; ++k) rr[k] = 0.;
    for (int s = 0; s < ntypes; ++s) {
      for (int k = sec[s]; k < sec[s + 1]; ++k) {
        const int j = nl[k];
        if (j < 0) break;
        for (int d = 0; d < 3; ++d) rr[k * 3 + d] = c[j * 3 + d] - c[i * 3 + d];
        env_row(rr + k * 3, rcut_smth, rcut, &val[k * 4], &der[k * 12]);
      }
    }
    const int ti = t[i];
    for (int k = 0; k < nem; ++k)
      e[k] = ti >= 0 ? (val[k] - avg[ti * nem + k]) / std_[ti * nem + k] : 0.;
    for (int k = 0; k < nem * 3; ++k)
      ed[k] = ti >= 0 ? der[k] / std_[ti * nem + k / 3] : 0.;
  }
}

// tabulate.cc:162-232, 234-341, 343-447 (FP64 and FP32)
void orc_tabulate_fusion_se_a_f64(double* out, const double* table,
                                  const double* info, const double* em_x,
                                  const double* em, int nloc, int nnei, int M,
                                  int sorted, int nd) {
  tab_fwd(out, table, info, em_x, em, nloc, nnei, M, sorted != 0, nd);
}
void orc_tabulate_fusion_se_a_f32(float* out, const float* table,
                                  const float* info, const float* em_x,
                                  const float* em, int nloc, int nnei, int M,
                                  int sorted, int nd) {
  tab_fwd(out, table, info, em_x, em, nloc, nnei, M, sorted != 0, nd);
}
void orc_tabulate_fusion_se_a_grad_f64(double* dx, double* dem,
                                       const double* table, const double* info,
                                       const double* em_x, const double* em,
                                       const double* dy, int nloc, int nnei,
                                       int M, int sorted, int nd) {
  tab_grad(dx, dem, table, info, em_x, em, dy, nloc, nnei, M, sorted != 0, nd);
}
void orc_tabulate_fusion_se_a_grad_f32(float* dx, float* dem,
                                       const float* table, const float* info,
                                       const float* em_x, const float* em,
                                       const float* dy, int nloc, int nnei,
                                       int M, int sorted, int nd) {
  tab_grad(dx, dem, table, info, em_x, em, dy, nloc, nnei, M, sorted != 0, nd);
}
void orc_tabulate_fusion_se_a_grad_grad_f64(
    double* dz_dy, const double* table, const double* info, const double* em_x,
    const double* em, const double* dz_dx, const double* dz_dem, int nloc,
    int nnei, int M, int sorted, int nd) {
  tab_grad_grad(dz_dy, table, info, em_x, em, dz_dx, dz_dem, nloc, nnei, M,
                sorted != 0, nd);
}
void orc_tabulate_fusion_se_a_grad_grad_f32(
    float* dz_dy, const float* table, const float* info, const float* em_x,
    const float* em, const float* dz_dx, const float* dz_dem, int nloc,
    int nnei, int M, int sorted, int nd) {
  tab_grad_grad(dz_dy, table, info, em_x, em, dz_dx, dz_dem, nloc, nnei, M,
                sorted != 0, nd);
}

// prod_force.cc:24-87: F_i -= sum_a nd_a ed_a ; F_j += sum_{a in j} nd_a ed_a
void orc_prod_force_a(double* force, const double* nd, const double* ed,
                      const int* nlist, int nloc, int nall, int nnei,
                      int nframes) {
  const int ndesc = 4 * nnei;
  std::memset(force, 0, sizeof(double) * (size_t)nframes * nall * 3);
  for (int g = 0; g < nframes * nloc; ++g) {
    const int f = g / nloc, i = g % nloc;
    double* fi = force + ((size_t)f * nall + i) * 3;
    const double* n = nd + (size_t)g * ndesc;
    const double* e = ed + (size_t)g * ndesc * 3;
    for (int a = 0; a < ndesc; ++a)
      for (int d = 0; d < 3; ++d) fi[d] -= n[a] * e[a * 3 + d];
    for (int k = 0; k < nnei; ++k) {
      const int j = nlist[(size_t)g * nnei + k];
      if (j < 0) continue;
      double* fj = force + ((size_t)f * nall + j) * 3;
      for (int a = 4 * k; a < 4 * k + 4; ++a)
        for (int d = 0; d < 3; ++d) fj[d] += n[a] * e[a * 3 + d];
    }
  }
}

// prod_virial.cc:23-69 (one frame):
//   virial[3p+q] += nd_a * rij_q * ed_{a,p}, summed per neighbour j also into
//   atom_virial[j].
void orc_prod_virial_a(double* virial, double* atom_virial, const double* nd,
                       const double* ed, const double* rij, const int* nlist,
                       int nloc, int nall, int nnei) {
  const int ndesc = 4 * nnei;
  for (int k = 0; k < 9; ++k) virial[k] = 0.;
  std::memset(atom_virial, 0, sizeof(double) * 9 * (size_t)nall);
  for (int i = 0; i < nloc; ++i) {
    for (int k = 0; k < nnei; ++k) {
      const int j = nlist[(size_t)i * nnei + k];
      if (j < 0) continue;
      for (int a = 4 * k; a < 4 * k + 4; ++a) {
        const double pref = -1.0 * nd[(size_t)i * ndesc + a];
        for (int p = 0; p < 3; ++p)
          for (int q = 0; q < 3; ++q) {
            const double v = pref * rij[((size_t)i * nnei + k) * 3 + q] *
                             ed[((size_t)i * ndesc + a) * 3 + p];
            virial[p * 3 + q] -= v;
            atom_virial[(size_t)j * 9 + p * 3 + q] -= v;
          }
      }
    }
  }
}

// prod_force_grad.cc:23-75 (note the j -> j % nloc fold of ghost indices)
void orc_prod_force_grad_a(double* grad_net, const double* grad,
                           const double* ed, const int* nlist, int nloc,
                           int nnei, int nframes) {
  const int ndesc = 4 * nnei;
  std::memset(grad_net, 0, sizeof(double) * (size_t)nframes * nloc * ndesc);
#pragma omp parallel for
  for (int g = 0; g < nframes * nloc; ++g) {
    const int f = g / nloc;
    double* gn = grad_net + (size_t)g * ndesc;
    const double* e = ed + (size_t)g * ndesc * 3;
    for (int a = 0; a < ndesc; ++a)
      for (int d = 0; d < 3; ++d) gn[a] -= grad[(size_t)g * 3 + d] * e[a * 3 + d];
    for (int k = 0; k < nnei; ++k) {
      int j = nlist[(size_t)g * nnei + k];
      if (j >= nloc) j = j % nloc;
      if (j < 0) continue;
      for (int a = 4 * k; a < 4 * k + 4; ++a)
        for (int d = 0; d < 3; ++d)
          gn[a] += grad[((size_t)f * nloc + j) * 3 + d] * e[a * 3 + d];
    }
  }
}

// prod_virial_grad.cc:22-66 (one frame)
void orc_prod_virial_grad_a(double* grad_net, const double* grad,
                            const double* ed, const double* rij,
                            const int* nlist, int nloc, int nnei) {
  const int ndesc = 4 * nnei;
  std::memset(grad_net, 0, sizeof(double) * (size_t)nloc * ndesc);
#pragma omp parallel for
  for (int i = 0; i < nloc; ++i) {
    for (int k = 0; k < nnei; ++k) {
      const int j = nlist[(size_t)i * nnei + k];
      if (j < 0) continue;
      for (int a = 4 * k; a < 4 * k + 4; ++a)
        for (int p = 0; p < 3; ++p)
          for (int q = 0; q < 3; ++q)
            grad_net[(size_t)i * ndesc + a] -=
                -1.0 * grad[p * 3 + q] * rij[((size_t)i * nnei + k) * 3 + q] *
                ed[((size_t)i * ndesc + a) * 3 + p];
    }
  }
}

}  // extern "C"
