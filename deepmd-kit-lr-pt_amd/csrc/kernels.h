// Internal launch interface between the host code and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "geometry.h"

namespace dpmi {

// ------------------------------------------------------------- coord.hip --
// Extended atom in the cell-sorted list: its coordinates, extended index and
// type, so that a neighbour search reads candidates as contiguous records.
struct __attribute__((aligned(16))) CellAtom {
  double x, y, z;
  int j, type;
};
struct CellWork {
  int nframes = 0;
  int ntot = 0;          // packed local atoms of all frames
  int max_nloc = 0;      // atoms of the largest frame (0: unknown; small frames are placed by the scan)
  int ncells_tot = 0;    // packed local cells of all frames
  int nx_tot = 0;        // packed extended cells (+1 sentinel per frame)
  int do_wrap = 1;
  const double* coord_in = nullptr;  // [ntot, 3]
  const int* type_in = nullptr;      // [ntot]
  const int* frame_of = nullptr;     // [ntot]
  const int* cell_frame = nullptr;   // [ncells_tot]
  const int* xcell_frame = nullptr;  // [nx_tot]
  const FrameGeom* geom = nullptr;   // [nframes]
  double* coord_ext = nullptr;       // [sum ext_cap, 3]
  int* type_ext = nullptr;
  int* map_ext = nullptr;
  CellAtom* cell_atoms = nullptr;    // [sum ext_cap], extended cells' atoms in cell order, each
                                     // cell's records grouped by sub-cell along the third cell
                                     // vector (index order inside a sub-cell)
  int* cell_sub = nullptr;           // [nx_tot, kAxisBins] optional: first record of each sub-cell
  int* cell_of = nullptr;            // [ntot]
  int* cell_slot = nullptr;          // [ntot] slot of the atom in its cell (counting order)
  int* cell_cnt = nullptr;           // [ncells_tot]; left zeroed by the build
  bool counts_clear = false;         // cell_cnt is zero on entry (a previous build left it so)
  int* cell_start = nullptr;
  int* cell_list = nullptr;          // [ntot] atoms by cell, counting order inside a cell
  int* sorted_loc = nullptr;         // [ntot] atoms by cell, ascending index inside a cell
  int* xcnt = nullptr;               // [nx_tot]
  int* xstart = nullptr;
  int* xgcnt = nullptr;
  int* xgstart = nullptr;
  int* nall = nullptr;               // [nframes]
  int* flag = nullptr;               // device status word
  double* force_clear = nullptr;     // [ntot, 3] zeroed by the wrap kernel when set
};
void launch_build_extended(const CellWork& w, hipStream_t st);
// one frame, caller buffers; returns a dpmi_status (capi.cpp)
int copy_coord_run(FrameGeom g, double* out_c, int* out_t, int* mapping, int* nall_host, const double* in_c,
                   const int* in_t, int nloc, int mem_nall, void* stream);
void launch_normalize(double* coord, int natom, const FrameGeom& g, hipStream_t st);
// geom[f].t / rec from box [nframes, 9] (device) when the cell grid stays the
// same; kFlagGridChanged otherwise
void launch_geom_refresh(const double* box, FrameGeom* geom, int nframes, float rcut, int* flag, hipStream_t st);
// normalize_coord_gpu<float>: FP32 arithmetic with the caller's box and
// reciprocal box (host copies of the Region's rows)
void launch_normalize_f32(float* coord, int natom, const float* boxt9, const float* rec9, hipStream_t st);

// ------------------------------------------------------------- nlist.hip --
struct NlistArgs {
  int cap = 256;        // LDS row capacity (power of two, 128..4096)
  const int* sec = nullptr;  // [ntypes + 1] device
  int ntypes = 0;
  int nnei = 0;
  double rc2 = 0;
  float rc2f = 0;
  int* nlist = nullptr;
  int* ncand = nullptr;
  // cell search only: when em_f or em_d is set, the kernel also stores the
  // environment-matrix rows of the formatted row (no derivatives), as the
  // environment kernel would, from avg/stdv [ntypes, nnei, 4], rmin, rmax
  const double* avg = nullptr;
  const double* stdv = nullptr;
  double rmin = 0, rmax = 0;
  float* em_f = nullptr;
  double* em_d = nullptr;
};
int nlist_row_capacity(int want);
void launch_cell_nlist(const CellWork& w, const NlistArgs& a, hipStream_t st);
// rows: jlist + r * jstride, or firstneigh[r] (device pointer array) if given;
// TC: coordinate type (double or float)
template <typename TC>
void launch_row_format(int inum, const int* ilist, const int* numneigh, const int* jlist, int jstride,
                       const TC* coord, const int* type, int nloc, int nall, const NlistArgs& a, int* flag,
                       hipStream_t st, const int* const* firstneigh = nullptr);
// T: coordinate / distance type of the pair test |x_i - x_j|^2 < rc2
template <typename T>
void launch_brute_nlist(const T* c, int nloc, int nall, int mem_size, T rc2, int nframes, const int* type,
                        int* ilist, int* numneigh, int* jlist, hipStream_t st, int* const* firstneigh = nullptr);

// --------------------------------------------------------------- env.hip --
struct EnvArgs {
  const double* coord = nullptr;
  const float* coord_f = nullptr;  // FP32 coordinates instead (geometry still evaluated in FP64)
  const int* type = nullptr;
  const int* nlist = nullptr;
  // packed rows: row g -> frame row_frame[g], extended offset
  // row_ext_off[frame], centre row_center[g]; when row_ext_off is null the
  // rows are strided frames of strided_nloc rows over nall_stride atoms.
  const int* row_frame = nullptr;
  const int* row_ext_off = nullptr;
  const int* row_center = nullptr;
  int strided_nloc = 0;
  long long nrows = 0;
  int nnei = 0;
  int nall_stride = 0;
  int ntypes = 1 << 30;  // rows of centres of type >= ntypes are zero
  const double* avg = nullptr;
  const double* std_ = nullptr;
  double rmin = 0, rmax = 0;
  double* em = nullptr;
  double* em_deriv = nullptr;
  double* rij = nullptr;
  int tpb = 0;  // threads per block (0: 256)
};
void launch_env_mat_a(const EnvArgs& a, hipStream_t st);
// same, storing em / em_deriv / rij in FP32 (em, em_deriv, rij of a are
// ignored); em_deriv null: em only
void launch_env_mat_a_f32(const EnvArgs& a, float* em, float* em_deriv, float* rij, hipStream_t st);

// ---------------------------------------------------------- tabulate.hip --
template <typename T>
int launch_tab_fwd(T* out, const T* table, const T* info_host, const T* em_x, const T* em, int nloc, int nnei,
                   int M, int sorted, int nd, hipStream_t st);
template <typename T>
int launch_tab_grad(T* dy_dem_x, T* dy_dem, const T* table, const T* info_host, const T* em_x, const T* em,
                    const T* dy, int nloc, int nnei, int M, int sorted, int nd, hipStream_t st);
template <typename T>
int launch_tab_grad_grad(T* dz_dy, const T* table, const T* info_host, const T* em_x, const T* em, const T* dz_dx,
                         const T* dz_dem, int nloc, int nnei, int M, int sorted, int nd, hipStream_t st);

// ---------------------------------------------------------- tablegen.hip --
// One embedding net (host pointers): width = {1, n0, ..., M}; per layer
// w [n_in, n_out] row-major, b [n_out], idt [n_out] or null.
struct TabNet {
  std::vector<int> width;
  std::vector<const double*> w, b, idt;
};
int tab_nspline(double lower, double upper, double extrapolate, double stride0, double stride1);
void tab_knots(std::vector<double>& xx, double lower, double upper, double extrapolate, double stride0,
               double stride1);
// table [nspline, 6 M] (host) of the net over [lower, extrapolate * upper]
void build_table_device(const TabNet& net, double lower, double upper, double extrapolate, double stride0,
                        double stride1, std::vector<double>& table, int& nspline, hipStream_t st);

// ------------------------------------------------------------- embed.hip --
// Engine descriptor: table embedding + contraction + descriptor (fwd) and the
// fused adjoint from dE/dD to net_deriv (bwd).  T: table/arithmetic type,
// TD: descriptor storage type (the fitting input), TG: dE/dD storage type.
struct EmbedArgs {
  void* xyz = nullptr;        // [nrows, 4, M] (T)
  void* D = nullptr;          // [nrows, M*A] (TD), rows permuted by pos
  const void* gD = nullptr;   // [nrows, M*A] (TG), rows permuted by pos
  void* net_deriv = nullptr;  // [nrows, nnei, 4] (T)
  const int* pos = nullptr;   // [nrows] row -> permuted descriptor row
  const void* tables = nullptr;
  const long long* tab_off = nullptr;  // [nnets] element offsets
  const void* infos = nullptr;         // device Info<T>[nnets]
  const int* net_of = nullptr;         // [ntypes * ntypes]: net of (ti, tj)
  const void* em = nullptr;            // [nrows, nnei, 4] (T); em_x is column 0
  const int* row_type = nullptr;       // [nrows]
  const int* sec = nullptr;            // [ntypes + 1]
  int ntypes = 0;
  long long nrows = 0;
  int nnei = 0;
  int M = 0;
  int A = 0;  // axis_neuron, at most 16
  int ldD = 0;  // row stride of D and gD (0: M*A)
  double inv_nnei = 0;
  int waves = 0;      // forward: waves (rows) per workgroup (0: 4)
  int waves_bwd = 0;  // backward: waves (rows) per workgroup (0: 1)
  // packed FP32 tables of the per-neighbour FP32 kernels (one 16-byte record
  // per column: a0, a1 FP32, a2..a5 FP16; rows of pk_stride(M) records) and
  // their per-net offsets in floats; null: the plain tables above
  const void* packed = nullptr;
  const long long* packed_off = nullptr;
  // matrix-core (table-space) kernels of the mixed engine
  const int* perm = nullptr;             // [nrows] descriptor position -> row
  const _Float16* tab16 = nullptr;       // split FP16 tables of every net
  const long long* tab16_off = nullptr;  // [4 nnets]: forward hi, forward lo, adjoint hi, adjoint lo
  const float* tab16_pscale = nullptr;   // [8 nnets]: 1 / sigma_p of the adjoint tables
  const int4* etiles = nullptr;          // [netiles] (type or -1, first position, count <= 4, 0)
  int netiles = 0;
  int* flag = nullptr;                   // device status word (kFlagHalfRange)
  // row-run adjoint of the mixed engine (matrix cores, one wavefront per row)
  const _Float16* rtab = nullptr;        // run tables of every net (build_run_tables)
  const long long* rtab_off = nullptr;   // [nnets] element offsets
  const float* rtab_pscale = nullptr;    // [8 nnets]: 1 / sigma_p
  // row-run forward of the mixed engine (matrix cores, one wavefront per row)
  const _Float16* ftab = nullptr;        // forward run tables of every net (build_fwd_run_tables)
  const long long* ftab_off = nullptr;   // [nnets] element offsets
  const float* ftab_scale = nullptr;     // [nnets]: 1 / sigma
};
// Uncompressed embedding (FP64 engine): slots <-> net rows, contraction and
// adjoint, one wavefront per descriptor position q in [q0, q1) (atom perm[q]);
// the slots of block tj of atom g are rows rowbase[g * ntypes + tj] + s.
struct NetArgs {
  const double* em = nullptr;   // [nrows, nnei, 4]
  const int* perm = nullptr;    // descriptor position -> row
  const int* pos = nullptr;     // row -> descriptor position
  const int* type = nullptr;    // [nrows]
  const int* rowbase = nullptr;
  const int* sec = nullptr;     // [ntypes + 1]
  const int* net_of = nullptr;  // [ntypes * ntypes]
  int ntypes = 0, nnei = 0, M = 0, A = 0, ldD = 0;
  double inv_nnei = 0;
  double* xyz = nullptr;        // [nrows, 4, M]
  double* D = nullptr;          // [nrows, M*A] by position
  const double* gD = nullptr;   // [nrows, M*A] by position
  double* net_deriv = nullptr;  // [nrows, nnei, 4]
};
// x0[row] = em[g, j, 0]
void launch_net_gather(const NetArgs& a, int q0, int q1, double* x0, hipStream_t st);
// xyz and D from the net outputs G [rows, M]
void launch_net_contract(const NetArgs& a, const double* G, int q0, int q1, hipStream_t st);
// dE/dG [rows, M] and dE/dz of the last layer (idt_top [nnets] device array of
// the nets' last-layer timesteps or nulls; ytop its act output), and
// net_deriv[g, j, c] = sum_k G dy (component 0 without the slope term)
void launch_net_adjoint(const NetArgs& a, const double* G, const double* ytop, const double* const* idt_top,
                        int act_top, int q0, int q1, double* gG, double* dz, hipStream_t st);
// net_deriv[g, j, 0] += ds[row]
void launch_net_slope(const NetArgs& a, const double* ds, int q0, int q1, hipStream_t st);

void* make_table_infos(const double* infos_host, int nnets, bool fp32);
size_t table_info_bytes(bool fp32);
void free_table_infos(void* p, bool fp32);
template <typename T, typename TD>
void launch_embed_fwd(const EmbedArgs& a, hipStream_t st);
template <typename T, typename TG>
void launch_embed_bwd(const EmbedArgs& a, hipStream_t st);
// Mixed engine, matrix cores: forward (xyz, D) and adjoint (net_deriv) of the
// table embedding computed in table space; M <= 128 and M % 4 == 0.
// Split tables: per net, rows [rows + 16] as FP16 hi + lo parts (p < 6;
// zero for p = 6, 7 and padded rows / columns): forward a_p s^p (s the
// stride of the row's segment) in layout [row][16 ceil(M/16)][8 p]; adjoint
// sigma_p a_p (sigma_p a power of two per net and p, 1 / sigma_p returned
// in pscale[8 net + p]) in layout [row][8 p][32 ceil(M/32)].
constexpr int kEmbedTileAtoms = 4;
bool embed_mfma_supported(int M);
// packed FP32 tables (records of 4 floats per column, rows padded with zero
// records to pk_stride(M) = a multiple of 16 records) and per-net offsets
void build_packed_tables(const double* table, const long long* tab_off, const int* nrows, int nnets, int M,
                         std::vector<float>& out, std::vector<long long>& off);
void build_split_tables(const double* table, const long long* tab_off, const int* nrows, const double* infos,
                        int nnets, int M, std::vector<_Float16>& out, std::vector<long long>& off,
                        std::vector<float>& pscale);
void launch_embed_fwd_mfma(const EmbedArgs& a, hipStream_t st);
void launch_embed_bwd_mfma(const EmbedArgs& a, hipStream_t st);
// Mixed engine, matrix cores: adjoint (net_deriv) from the table rows of each
// atom's neighbour runs, H[r, p, c] = sum_k a_p[r, k] dy[c, k] once per run.
// Run tables: per net [row][ceil(M/16)][4][8 A-rows][4 halves] (column
// 16 kc + 4 kq + i), A-rows sigma_0 a_0 hi, lo, sigma_1 a_1 hi, lo,
// sigma_p a_p (p = 2..5), sigma_p a power of two per net and p (1 / sigma_p
// in pscale[8 net + p]).
bool embed_runs_supported(int M, int A);
// throws std::invalid_argument when a net's table_info (FP64 or rounded to
// FP32) locates rows at or past its nrows
void check_table_rows(const double* infos, const int* nrows, int nnets);
void build_run_tables(const double* table, const long long* tab_off, const int* nrows, int nnets, int M,
                      std::vector<_Float16>& out, std::vector<long long>& off, std::vector<float>& pscale);
void launch_embed_bwd_runs(const EmbedArgs& a, hipStream_t st);
// Forward (xyz, D) from the table rows of each atom's neighbour runs,
// xyz = sum over runs of Q[run, p, c] a'_p[r, k] (a'_p = a_p s^p, s the
// stride of row r's segment).  Forward run tables: per net
// [row][ceil(M/16)][2][16][4 halves] (A-rows 4 quad + i of column
// 16 kb + j): sigma a'_0 hi, lo, sigma a'_1 hi, lo, sigma a'_p (p = 2..5),
// sigma one power of two per net (1 / sigma in scale[net]).
void build_fwd_run_tables(const double* table, const long long* tab_off, const int* nrows, const double* infos,
                          int nnets, int M, std::vector<_Float16>& out, std::vector<long long>& off,
                          std::vector<float>& scale);
void launch_embed_fwd_runs(const EmbedArgs& a, hipStream_t st);

// ------------------------------------------------------------- force.hip --
template <typename T>
void launch_prod_force_a(T* force, const T* net_deriv, const T* env_deriv, const int* nlist, int nloc, int nall,
                         int nnei, int nframes, hipStream_t st);
// per-row virial partials [nloc, 9] (FP64, caller reduces) and atom_virial
template <typename T>
void launch_prod_virial_a(double* vpart, T* atom_virial, const T* net_deriv, const T* env_deriv, const T* rij,
                          const int* nlist, int nloc, int nall, int nnei, hipStream_t st);
template <typename T>
void launch_reduce_rows9(T* out9, const double* part, int nrows, hipStream_t st);
// Rows of the fixed virial scratch (noalloc.hip) and its device address on
// the current device (looked up once per device, never allocated).
constexpr int kNoAllocVirialRows = 1 << 17;
double* noalloc_virial_scratch();
template <typename T>
void launch_prod_force_grad_a(T* grad_net, const T* grad, const T* env_deriv, const int* nlist, int nloc, int nnei,
                              int nframes, hipStream_t st);
template <typename T>
void launch_prod_virial_grad_a(T* grad_net, const T* grad, const T* env_deriv, const T* rij, const int* nlist,
                               int nloc, int nnei, hipStream_t st);

// engine: force/virial over packed rows (the LDS-aggregated kernel adds onto
// the owner rows of `force`; the row kernel into extended force buffers)
struct ForceRowsArgs {
  const void* net_deriv = nullptr;  // [nrows, nnei*4]  (double, or float when fp32)
  const void* em_deriv = nullptr;   // [nrows, nnei*12]; null: recomputed from coord_ext
  const void* rij = nullptr;        // [nrows, nnei*3]
  bool fp32 = false;
  // for recomputed slots: extended coordinates and types, dstd [ntypes, nnei, 4], rcut_smth, rcut
  const double* coord_ext = nullptr;
  const int* type_ext = nullptr;
  const double* dstd = nullptr;
  double rmin = 0, rmax = 0;
  // (the LDS-aggregated kernel's rows must then be given)
  const int* nlist = nullptr;
  const int* row_frame = nullptr;
  const int* row_center = nullptr;
  const int* frame_ext_off = nullptr;
  long long nrows = 0;
  int nnei = 0;
  double* force_ext = nullptr;        // [sum ext_cap, 3] (zeroed by caller; row kernel only)
  double* vpart = nullptr;            // [nrows, 9]
  // cell order of the rows (per-frame local indices, packed like the rows)
  // and the frame layout; when set, the LDS-aggregated kernel is used
  const int* sorted_loc = nullptr;
  const FrameGeom* geom = nullptr;
  int rows = 0;  // rows per workgroup of the LDS-aggregated kernel (0: 16)
  // the LDS-aggregated kernel adds every term straight onto the owner atom
  // (ghost images folded through map_ext) in force [ntot, 3] (zeroed by caller)
  const int* map_ext = nullptr;
  double* force = nullptr;
  // per-atom virials [ntot, 9] keyed like the forces (zeroed by caller), or null
  double* atom_virial = nullptr;
  // no ghost fold: terms go to the extended atom itself (force and
  // atom_virial are [nall, 3] / [nall, 9], map_ext unused)
  bool unfolded = false;
};
// clears force_ext / atom_virial_ext [nall] and flags local types that differ
// from type_loc (kFlagTypeMismatch) and types >= ntypes (kFlagTypeRange)
void launch_lower_prep(double* force_ext, double* atom_virial_ext, int nall, const int* atype_ext,
                       const int* type_loc, int nloc, int ntypes, int* flag, hipStream_t st);
// whether launch_force_virial_rows takes the LDS-aggregated kernel (which
// writes owner forces) for these arguments
bool force_rows_aggregated(const ForceRowsArgs& a);
// whether launch_force_virial_rows can rebuild the slot derivatives of rows
// with nnei slots (the LDS-aggregated FP32 kernel applies)
bool force_recompute_supported(int nnei);

// Launch-shape knobs of the engine stages, picked per size bucket by
// SeaEngine::autotune.
constexpr int kFitSplit = 256;
constexpr int kFitWide = 512;
struct Tune {
  int env_tpb = 256;
  int fwd_waves = 4;
  int bwd_waves = 1;
  int force_rows = 16;
  // mixed engine: bit 0 forward, bit 1 adjoint of the table embedding on the
  // matrix cores (table space, four atoms per wavefront); pays off when many
  // neighbours of an atom share table rows, so it is off until the autotuner
  // measures it faster; bit 2 the row-run adjoint on the matrix cores (one
  // atom per wavefront; takes precedence over bit 1); bit 3 the row-run
  // forward on the matrix cores (precedence over bit 0)
  int embed_mfma = 0;
  // fused fitting rows per workgroup tile (32 or 64; 0: chosen from the
  // size), plus kFitSplit: layer 0's products in their own kernels, or plus
  // kFitWide (64-row tiles): 16 waves per workgroup
  int fit_rows = 0;
};
void launch_force_virial_rows(const ForceRowsArgs& a, hipStream_t st);
// force[loc] += force_ext[e] for every extended atom e < nall of each frame
void launch_fold_ghost_forces(const double* force_ext, const int* map_ext, const int* xframe_of_ext,
                              const FrameGeom* geom, const int* nall, long long next_tot, double* force,
                              hipStream_t st);
// per-frame energy and virial (Neumaier-compensated, two levels): chunks of
// at most 1024 rows of one frame -> partials [nchunks, 10] -> frame totals.
// A chunk record is 4 ints: first row, rows, frame, chunks of that frame.
// With `ticket` ([nframes] ints, zero before the first call; the kernel
// leaves them zero) and every frame holding a chunk, the last chunk of a
// frame to finish combines it in the same launch.
void frame_chunks(const int* natoms, int nframes, std::vector<int>& chunks4, std::vector<int>& chunk_start);
void launch_frame_reduce(const double* atom_e, const double* vpart, const int* chunks4, int nchunks,
                         const int* chunk_start, int nframes, double* part, int* ticket, double* energy,
                         double* virial, hipStream_t st);

// ----------------------------------------------------------- fitting.hip --
void launch_fill_rows(double* g, const double* w, long long n, int width, hipStream_t st);

// ------------------------------------------------------------- fit64.hip --
// One FP64 fitting layer for every atom type (rows of type t are
// [p0[t], p0[t] + rows[t]) of the row-major activation buffers).
constexpr int kFit64MaxTypes = 8;
struct Fit64Layer {
  int ntypes = 0;
  long long p0[kFit64MaxTypes] = {}, rows[kFit64MaxTypes] = {};
  const double* w[kFit64MaxTypes] = {};         // fit64_swizzle of [K][N]: forward W, backward W^T
  const double* bias[kFit64MaxTypes] = {};      // forward [N]
  const double* idt[kFit64MaxTypes] = {};       // forward: this layer's timestep [N] (or null)
  const double* idt_next[kFit64MaxTypes] = {};  // backward: timestep [N] of the layer below (or null)
  const double* wout[kFit64MaxTypes] = {};      // top hidden layer: output weights [N] (forward) / [K] (backward)
  int K = 0, N = 0, act = 0, res_mode = 0;      // res_mode 1: + x, 2: + [x, x]
  int act_next = 0;                             // backward: activation of the layer below
  const double* a = nullptr;  // product operand rows: forward x, backward dE/dz of this layer
  int lda = 0;
  const double* g = nullptr;  // backward: dE/d(layer output) rows for the residual (unless wout)
  int ldg = 0;
  const double* y = nullptr;  // backward: act output of the layer below
  int ldy = 0;
  // forward: out0 = act(z + b) (or null), out1 = x_next, out2 = dE/dz when
  // wout is set.  backward: out1 null -> out0 = dE/d(input) (the descriptor
  // gradient); else out0 = dE/d(input) (or null), out1 = dE/dz of the layer below
  double* out0 = nullptr;
  double* out1 = nullptr;
  double* out2 = nullptr;
  int ldo = 0;
};
void launch_fit64(const Fit64Layer& p, bool backward, hipStream_t st);
// row-major [K][N] -> MFMA B fragments [ceil(K/16)][ceil(N/16)][4][64], zero padded
std::vector<double> fit64_swizzle(const double* w, int K, int N);
struct Fit64Out {
  int ntypes = 0;
  long long p0[kFit64MaxTypes] = {}, rows[kFit64MaxTypes] = {};
  const double* w[kFit64MaxTypes] = {};  // output weights [K]
  double b_out[kFit64MaxTypes] = {}, bias[kFit64MaxTypes] = {};
  const double* x = nullptr;  // last hidden activations [.., K]
  int K = 0;
  const int* perm = nullptr;  // row position -> atom
  double* atom_e = nullptr;
};
void launch_fit64_out(const Fit64Out& p, hipStream_t st);

// ------------------------------------------------------------- fit16.hip --
// One fitting layer of the mixed engine, widths padded to multiples of 64.
struct Fit16Layer {
  const _Float16* wt = nullptr;  // [n_pad][k_pad] forward operand (W^T), FP16 hi part
  const _Float16* w = nullptr;   // [k_pad][n_pad] backward operand (W), FP16 hi part
  const _Float16* wt_lo = nullptr;  // lo parts (W - hi in FP16), same layouts
  const _Float16* w_lo = nullptr;
  const float* b = nullptr;      // [n_pad]
  const float* idt = nullptr;    // [n_pad] or null
  const float* out_w = nullptr;  // [n_pad] output-layer weights (last hidden layer) or null
  int act = 1, res_mode = 0;
  int k_pad = 0, n_pad = 0;
};
// y = act(x W + b) (FP32), xo = y * idt + xres (FP16); with e_acc (last hidden
// layer) xo is not stored and e_acc[r] += xo . out_w instead
void launch_fit16_fwd(const Fit16Layer& L, const _Float16* x, int ldx, int rows, const _Float16* xres, int ldr,
                      float* y, int ldy, _Float16* xo, int ldxo, float* e_acc, hipStream_t st);
// g_in = (g * idt * (1 - y^2)) W^T + (residual ? g : 0); g_bcast: g is one row [n_pad]
void launch_fit16_bwd(const Fit16Layer& L, const float* g, int ldg, bool g_bcast, const float* y, int ldy, int rows,
                      bool residual, float* g_in, int ld_in, hipStream_t st);
// atom_e[g] = e_perm[pos[g]] + shift[type[g]]
void launch_energy_out(const float* e_perm, const int* pos, const int* row_type, const double* shift, long long n,
                       double* atom_e, hipStream_t st);

// --------------------------------------------------------- fitfused.hip --
// Whole fitting net (forward, energy, backward to dE/dD) of the mixed engine
// in one launch, one workgroup per tile of <= 64 rows of one type.
// Weights are split into FP16 hi + lo parts (W = hi + lo to about 22 bits):
// each product is x hi + x lo with FP32 accumulation, so rounding the
// weights to FP16 leaves no systematic error in the energy.
struct FusedLayer {
  const _Float16* wf = nullptr;     // forward B fragments: Bt = W^T [N][K], swizzle_bfrags order (hi)
  const _Float16* wb = nullptr;     // backward B fragments: Bt = W [K][N] (hi)
  const _Float16* wf_lo = nullptr;  // lo parts, same layouts
  const _Float16* wb_lo = nullptr;
  const float* b = nullptr;      // [N]
  const float* idt = nullptr;    // [N] or null
  const float* out_w = nullptr;  // [N] output weights (last hidden layer)
  int act = 1, res = 0;
  int K = 0, N = 0;              // padded input / output widths (multiples of 64, N <= 256)
};
struct FusedFitArgs {
  const _Float16* D = nullptr;  // [rows, ldD] descriptor (FP16), rows in type order
  int ldD = 0;
  float* gD = nullptr;          // [rows, ldG] dE/dD out
  int ldG = 0;
  const int4* tiles = nullptr;  // [ntiles] (type, first row, row count, 0)
  const FusedLayer* layers = nullptr;  // [ntypes * H] device
  const int* perm = nullptr;    // descriptor row -> atom
  double* atom_e = nullptr;     // [atoms] out
  const double* shift = nullptr;  // [ntypes] output bias + bias_atom_e
  int H = 0;                    // hidden layers, 1..4
  // layer 0 in its own kernels (launch_fit_split): x_1 and idt_0 act'(z_0)
  // from the forward product, dE/dz_0 into the backward product, [rows, ldX]
  // FP16 each; nxcd: XCDs the workgroups are dealt to round-robin
  _Float16* x1 = nullptr;
  _Float16* dv1 = nullptr;
  _Float16* gz0 = nullptr;
  int ldX = 0;
  int nxcd = 1;
};
// tiles of rows_per_tile (32 or 64) descriptor rows; wide (64-row tiles
// only): 16 waves of one column tile each instead of 8 waves of two (the
// same products; the energy's FP32 partial sums grouped differently)
void launch_fit_fused(const FusedFitArgs& a, int ntiles, int rows_per_tile, hipStream_t st, bool wide = false);
// the same results (bit for bit) from three launches: layer 0 forward as a
// product tiled over rows and output columns, the other layers fused per row
// tile, layer 0 backward (dE/dD) tiled over rows and output columns; needs
// H >= 2 and a.x1/dv1/gz0
// (layer0_host: the layer-0 widths K, N of the launch grid)
void launch_fit_split(const FusedFitArgs& a, int ntiles, int rows_per_tile, const FusedLayer& layer0_host,
                      hipStream_t st);
// host: Bt [N][K] (row-major) -> MFMA B-fragment order (N % 16 == 0, K % 32 == 0)
void swizzle_bfrags(const _Float16* bt, int N, int K, _Float16* out);

void launch_fill_int(int* p, int v, long long n, hipStream_t st);
void launch_fill_double(double* p, double v, long long n, hipStream_t st);
void launch_fill_float(float* p, float v, long long n, hipStream_t st);
void launch_copy_double(double* dst, const double* src, long long n, hipStream_t st);

}  // namespace dpmi
