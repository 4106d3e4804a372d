/*
 * dpmi -- MI355X (gfx950) se_e2_a energy/force/virial path, C ABI.
 *
 * Plain pointers and sizes only.  Every array argument is DEVICE memory
 * unless its name ends in _host.  All entry points are asynchronous on the
 * given HIP stream (NULL = the default stream) except where a host output
 * (suffix _host) forces a synchronisation, and return a dpmi_status; the
 * message of the last failure on the calling thread is dpmi_last_error().
 *
 * Two groups:
 *   1. kernel-level entry points, one per reference lib function on the path
 *      (reference: source/lib/include, the deepmd::..._gpu templates).
 *      Layouts, units and conventions are exactly the reference's; the
 *      deepmd:: C++ templates of include/deepmd_lib.h forward to these.
 *   2. the whole-path engine (dpmi_sea_*): cell list -> neighbour list ->
 *      environment matrix -> compressed embedding + G^T R contraction ->
 *      fitting net -> net_deriv -> force/virial, for one or many frames,
 *      in FP64 or in mixed precision.  This is what the PyTorch op layer,
 *      the C++ DeepPot-style caller and bench.py drive.
 */
#ifndef DPMI_H
#define DPMI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dpmi_status {
  DPMI_OK = 0,
  DPMI_ERR_ARGUMENT = 1,         /* invalid argument (std::invalid_argument in the reference) */
  DPMI_ERR_HIP = 2,              /* HIP runtime error */
  DPMI_ERR_NLIST_CAPACITY = 3,   /* a neighbour row exceeds its capacity (deepmd_exception_nlist_capacity) */
  DPMI_ERR_MEMORY = 4,           /* nall > mem_nall or allocation failure (reference returns 1 / deepmd_exception_oom) */
  DPMI_ERR_RANGE = 5,            /* index/type outside the sort-key range (reference traps), or a
                                    mixed-precision matrix-core operand outside the FP16 range */
  DPMI_ERR_UNSUPPORTED = 6       /* ndescrpt not in {4,9,16,25}, last_layer_size outside (0,1024] */
} dpmi_status;

/* Message of the last failure on this thread ("" if none). */
const char* dpmi_last_error(void);
/* Library build string (arch, precision flags). */
const char* dpmi_version(void);

/* ===================================================================== */
/* 1. kernel-level entry points (reference source/lib/include)          */
/* ===================================================================== */

/* coord.h:50-54 normalize_coord_gpu<double>: wrap coord[natom*3] into the
 * primary cell of the box given by boxt_host[9] (rows = cell vectors). */
int dpmi_normalize_coord_f64(double* coord, int natom, const double* boxt_host, void* stream);

/* coord.h:56-82 copy_coord_gpu<double>: ghost images of nloc normalised
 * atoms inside an rcut shell.  Output order equals the reference CPU
 * copy_coord (locals, then ghost cells in lexicographic order, each holding
 * its source cell's atoms in ascending index).  *nall_host receives nall;
 * returns DPMI_ERR_MEMORY (reference: return 1) when nall > mem_nall. */
int dpmi_copy_coord_f64(double* out_c, int* out_t, int* mapping, int* nall_host, const double* in_c,
                        const int* in_t, int nloc, int mem_nall, float rcut, const double* boxt_host,
                        void* stream);

/* neighbor_list.h:230-247 build_nlist_gpu<double>: for every row
 * r = frame*nloc + i, all j != i of that frame with |x_i - x_j|^2 < rcut^2
 * (and type >= 0 when type != NULL), ascending j, written to
 * jlist[r*mem_size ...]; ilist[r] = i, numneigh[r] = count.
 * *max_list_size_host receives the longest row; returns
 * DPMI_ERR_NLIST_CAPACITY when a row exceeds mem_size (reference: 1). */
int dpmi_build_nlist_f64(int* ilist, int* numneigh, int* jlist, int* max_list_size_host,
                         const double* c_cpy, int nloc, int nall, int mem_size, float rcut, int nframes,
                         const int* type, void* stream);

/* fmt_nlist.h:23-35 format_nbor_list_gpu + prod_env_mat.h:82-117
 * prod_env_mat_a_gpu<double>: formats the candidate rows (ilist/numneigh,
 * row r at jlist + r*jstride) into nlist[nf*nloc, nnei] sorted by
 * (type, float r^2, index) and truncated at sel, then writes
 * em[nf*nloc, nnei*4], em_deriv[nf*nloc, nnei*12], rij[nf*nloc, nnei*3]
 * normalised by avg/std[ntypes, nnei*4] of the centre type.
 * sel_host[ntypes]. */
int dpmi_prod_env_mat_a_f64(double* em, double* em_deriv, double* rij, int* nlist, const double* coord,
                            const int* type, int inum, const int* ilist, const int* numneigh,
                            const int* jlist, int jstride, const double* avg, const double* std_,
                            int nloc, int nall, int nframes, float rcut, float rcut_smth,
                            const int* sel_host, int ntypes, void* stream);

/* tabulate.h:175-218 tabulate_fusion_se_a{,_grad,_grad_grad}_gpu<T>
 * (two_embed == NULL: the se_a path).  table_info_host[6] is a host array
 * (the reference reads it on the host).  ndescrpt in {4,9,16,25}. */
int dpmi_tabulate_fusion_se_a_f64(double* out, const double* table, const double* table_info_host,
                                  const double* em_x, const double* em, int nloc, int nnei,
                                  int last_layer_size, int is_sorted, int ndescrpt, void* stream);
int dpmi_tabulate_fusion_se_a_f32(float* out, const float* table, const float* table_info_host,
                                  const float* em_x, const float* em, int nloc, int nnei,
                                  int last_layer_size, int is_sorted, int ndescrpt, void* stream);
int dpmi_tabulate_fusion_se_a_grad_f64(double* dy_dem_x, double* dy_dem, const double* table,
                                       const double* table_info_host, const double* em_x, const double* em,
                                       const double* dy, int nloc, int nnei, int last_layer_size,
                                       int is_sorted, int ndescrpt, void* stream);
int dpmi_tabulate_fusion_se_a_grad_f32(float* dy_dem_x, float* dy_dem, const float* table,
                                       const float* table_info_host, const float* em_x, const float* em,
                                       const float* dy, int nloc, int nnei, int last_layer_size,
                                       int is_sorted, int ndescrpt, void* stream);
int dpmi_tabulate_fusion_se_a_grad_grad_f64(double* dz_dy, const double* table, const double* table_info_host,
                                            const double* em_x, const double* em, const double* dz_dy_dem_x,
                                            const double* dz_dy_dem, int nloc, int nnei,
                                            int last_layer_size, int is_sorted, int ndescrpt, void* stream);
int dpmi_tabulate_fusion_se_a_grad_grad_f32(float* dz_dy, const float* table, const float* table_info_host,
                                            const float* em_x, const float* em, const float* dz_dy_dem_x,
                                            const float* dz_dy_dem, int nloc, int nnei,
                                            int last_layer_size, int is_sorted, int ndescrpt, void* stream);

/* prod_force.h:65-73 prod_force_a_gpu<double>: force[nf, nall, 3]. */
int dpmi_prod_force_a_f64(double* force, const double* net_deriv, const double* in_deriv, const int* nlist,
                          int nloc, int nall, int nnei, int nframes, void* stream);
/* prod_virial.h:30-39 prod_virial_a_gpu<double> (one frame):
 * virial[9], atom_virial[nall, 9]. */
int dpmi_prod_virial_a_f64(double* virial, double* atom_virial, const double* net_deriv,
                           const double* env_deriv, const double* rij, const int* nlist, int nloc, int nall,
                           int nnei, void* stream);
/* No counterpart in the reference: prod_virial_a_gpu<double> as
 * dpmi_prod_virial_a_f64, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-row partials go to a fixed device array of the
 * library (9 * 131072 doubles = 9 MiB per GPU, held until the process ends):
 * nloc above 131072 returns DPMI_ERR_ARGUMENT.  No hipMalloc/hipFree is
 * called.  Asynchronous on `stream`; every call on a GPU shares that array,
 * so calls that overlap on one GPU (different streams, or threads) race and
 * must be serialised by the caller. */
int dpmi_prod_virial_a_f64_noalloc(double* virial, double* atom_virial, const double* net_deriv,
                                   const double* env_deriv, const double* rij, const int* nlist, int nloc, int nall,
                                   int nnei, void* stream);
/* prod_force_grad.h:27-34 prod_force_grad_a_gpu<double>. */
int dpmi_prod_force_grad_a_f64(double* grad_net, const double* grad, const double* env_deriv,
                               const int* nlist, int nloc, int nnei, int nframes, void* stream);
/* prod_virial_grad.h:27-34 prod_virial_grad_a_gpu<double> (one frame). */
int dpmi_prod_virial_grad_a_f64(double* grad_net, const double* grad, const double* env_deriv,
                                const double* rij, const int* nlist, int nloc, int nnei, void* stream);
/* The same four for float (prod_force.h / prod_virial.h / _grad.h float
 * instantiations; per-row virial partials are summed in FP64). */
int dpmi_prod_force_a_f32(float* force, const float* net_deriv, const float* in_deriv, const int* nlist,
                          int nloc, int nall, int nnei, int nframes, void* stream);
int dpmi_prod_virial_a_f32(float* virial, float* atom_virial, const float* net_deriv, const float* env_deriv,
                           const float* rij, const int* nlist, int nloc, int nall, int nnei, void* stream);
int dpmi_prod_force_grad_a_f32(float* grad_net, const float* grad, const float* env_deriv, const int* nlist,
                               int nloc, int nnei, int nframes, void* stream);
int dpmi_prod_virial_grad_a_f32(float* grad_net, const float* grad, const float* env_deriv, const float* rij,
                                const int* nlist, int nloc, int nnei, void* stream);

/* deepmd/utils/tabulate.py:67-370 + tabulate_math.py:353-470 (the compressed
 * table of one embedding net, DPTabulate.build/_build_lower/_make_data),
 * built on the device: the net (nlayers layers, widths width_host[0..nlayers]
 * with width_host[0] == 1; w_host[l] [n_in, n_out] row-major, b_host[l],
 * idt_host[l] or NULL; idt_host itself may be NULL; all host arrays) is
 * evaluated with its first and second derivative at the knots
 * arange(lower, upper, stride0) ++ arange(upper, extrapolate*upper, stride1)
 * ++ [extrapolate*upper], and table_host [nspline, 6*M] (host) receives the
 * quintic Hermite coefficients per interval and column.  *nspline_host
 * receives nspline = int((upper-lower)/stride0 + (extrapolate*upper-upper)/stride1);
 * table_host NULL: only that count.  Synchronous. */
int dpmi_tabulate_build_se_a(double* table_host, int* nspline_host, int nlayers, const int* width_host,
                             const double* const* w_host, const double* const* b_host,
                             const double* const* idt_host, double lower, double upper, double extrapolate,
                             double stride0, double stride1, void* stream);

/* ===================================================================== */
/* 2. whole-path engine                                                  */
/* ===================================================================== */

typedef struct dpmi_sea_model dpmi_sea_model;
typedef struct dpmi_sea_engine dpmi_sea_engine;

/* Model description (all host pointers, copied at creation).
 * Embedding net e serves centre type ti / neighbour type tj with
 * e = tj*ntypes + ti, or e = tj when type_one_side (reference
 * deepmd/pt/model/descriptor/se_a.py:715-719). */
typedef struct dpmi_sea_model_desc {
  int ntypes;
  const int* sel;            /* [ntypes] */
  float rcut, rcut_smth;
  int last_layer_size;       /* M */
  int axis_neuron;           /* A */
  int type_one_side;
  const double* davg;        /* [ntypes, nnei*4] */
  const double* dstd;        /* [ntypes, nnei*4] */
  int nnets;                 /* ntypes or ntypes^2 */
  const int* nspline;        /* [nnets] table rows */
  const double* const* tables;     /* [nnets] -> [nspline, 6*M] */
  const double* table_info;  /* [nnets, 6] */
  int fit_nlayers;           /* hidden layers + output layer */
  const int* fit_width;      /* [fit_nlayers + 1]: M*A, h0, ..., 1 */
  const double* const* fit_w;   /* [ntypes*fit_nlayers] -> [n_in, n_out] row-major */
  const double* const* fit_b;   /* [ntypes*fit_nlayers] -> [n_out] */
  const double* const* fit_idt; /* [ntypes*fit_nlayers] -> [n_out] or NULL */
  const int* fit_act;        /* [fit_nlayers] 1 = tanh, 0 = linear */
  const int* fit_resnet;     /* [fit_nlayers] residual allowed (x or [x, x]) */
  const double* bias_atom_e; /* [ntypes] */
  /* Uncompressed embedding (deepmd/pt/model/descriptor/se_a.py:785-839
   * without compression), used when tables == NULL: embed_nlayers layers per
   * net, widths embed_width [embed_nlayers + 1] = 1, n0, ..., M; per (net,
   * layer) w [n_in, n_out] row-major, b [n_out], idt [n_out] or NULL; per
   * layer act (1 = tanh) and resnet (x or [x, x] residual allowed).  Such a
   * model runs in DPMI_FP64 engines only.  Zero embed_nlayers with tables
   * given: compressed only. */
  int embed_nlayers;
  const int* embed_width;
  const double* const* embed_w;   /* [nnets * embed_nlayers] */
  const double* const* embed_b;   /* [nnets * embed_nlayers] */
  const double* const* embed_idt; /* [nnets * embed_nlayers], entries may be NULL */
  const int* embed_act;           /* [embed_nlayers] */
  const int* embed_resnet;        /* [embed_nlayers] */
} dpmi_sea_model_desc;

int dpmi_sea_model_create(dpmi_sea_model** model, const dpmi_sea_model_desc* desc);
/* Loads a dpmi model file (an se_e2_a energy model as plain arrays: descriptor
 * settings, davg/dstd, embedding and fitting weights, bias_atom_e, compression
 * settings and type names; written by dpmi.model.save_model_file from a
 * DeePMD-kit serialized model) and builds its compressed tables on the
 * device (dpmi_tabulate_build_se_a).  No counterpart file format in the
 * reference: it stands in for the frozen model that DeepPot loads
 * (source/api_cc/src/DeepPotPT.cc:40-120). */
int dpmi_sea_model_load(dpmi_sea_model** model, const char* path, void* stream);
/* ntypes, rcut, nnei = sum(sel), and the type names separated by spaces
 * (any output may be NULL). */
int dpmi_sea_model_info(const dpmi_sea_model* model, int* ntypes_host, double* rcut_host, int* nnei_host,
                        char* type_map_host, int type_map_len);
void dpmi_sea_model_destroy(dpmi_sea_model* model);

/* precision */
#define DPMI_FP64 0   /* every stage FP64 (the plain path) */
#define DPMI_MIXED 1  /* FP64 image shift and s(r), FP32 env matrix / table embedding / contraction,
                         FP16 descriptor and fitting GEMMs on MFMA (FP32 accumulate), FP64 force,
                         virial and energy sums; returns DPMI_ERR_UNSUPPORTED for models with no
                         hidden fitting layer or width-doubling residual layers */

/* Engine for frames with fixed atom types.  nframes frames, frame f holding
 * natoms_host[f] atoms whose types are atype_host[offset_f ...] (frames
 * packed back to back).  Buffers are sized here for mem_nall_per_atom
 * ghost-extended atoms per local atom (0 = automatic). */
int dpmi_sea_engine_create(dpmi_sea_engine** engine, const dpmi_sea_model* model, int nframes,
                           const int* natoms_host, const int* atype_host, int precision,
                           double mem_nall_per_atom);
void dpmi_sea_engine_destroy(dpmi_sea_engine* engine);

/* One E+F+V evaluation of every frame.  coord [sum natoms, 3] device,
 * box_host [nframes, 9] host.  Outputs (device): energy [nframes],
 * force [sum natoms, 3], virial [nframes, 9]; atom_energy [sum natoms] and
 * atom_virial [sum natoms, 9] may be NULL.  The per-atom virial is that of
 * DeepPot's atomic output (deepmd/pt/model/model/transform_output.py:70-101
 * with the atomic correction, equal to source/lib/src/prod_virial.cc:55-63
 * on the extended atoms, ghost images summed onto their owners).
 * With capacity growth on (the default, dpmi_sea_engine_set_growth) the
 * call reads the step's status back: a ghost or neighbour-row overflow
 * grows that buffer and recomputes the step, as the reference retries
 * (source/op/tf/prod_env_mat_multi_device.cc:2015-2028), and the call
 * returns when the results are final.  With growth off the call is
 * asynchronous and device-side errors are reported by
 * dpmi_sea_engine_status(). */
int dpmi_sea_engine_compute(dpmi_sea_engine* engine, const double* coord, const double* box_host,
                            double* energy, double* force, double* virial, double* atom_energy, double* atom_virial,
                            void* stream);
/* forward_lower (deepmd/pt/model/model/make_model.py:372-453) with the lib's
 * external neighbour list (source/op/tf/prod_env_mat_multi_device.cc:2313-2396,
 * nei_mode 3/4 on the device, in the row layout of convert_nlist_gpu_device):
 * one frame of an engine created with nframes == 1, whose nloc atoms (types
 * fixed at creation) come first in coord_ext [nall, 3] / atype_ext [nall],
 * followed by nall - nloc ghosts the caller owns (no periodic images are
 * made).  Candidate row r of centre ilist[r] holds numneigh[r] extended
 * indices at jlist + r*jstride (jstride <= 4096), in any order and at any
 * radius >= rcut (an rcut + skin list is cut at rcut, sorted by
 * (type, r, index) and truncated at sel per type, as format_nlist does);
 * inum == nloc.  Outputs (device): energy [1], force_ext [nall, 3]
 * (extended_force: ghost terms are not folded), virial [9], atom_energy
 * [nloc] or NULL, atom_virial_ext [nall, 9] or NULL (extended_virial with the
 * atomic correction).  Asynchronous; dpmi_sea_engine_status reports local
 * types that differ from the engine's (DPMI_ERR_ARGUMENT), out-of-range
 * types or indices (DPMI_ERR_RANGE) and rows longer than jstride. */
int dpmi_sea_engine_compute_lower(dpmi_sea_engine* engine, const double* coord_ext, const int* atype_ext, int nall,
                                  int inum, const int* ilist, const int* numneigh, const int* jlist, int jstride,
                                  double* energy, double* force_ext, double* virial, double* atom_energy,
                                  double* atom_virial_ext, void* stream);
/* Capacity growth on (1, the default) or off (0): see dpmi_sea_engine_compute. */
int dpmi_sea_engine_set_growth(dpmi_sea_engine* engine, int enable);
/* Engines of uncompressed models evaluate the embedding net for every
 * neighbour slot, in chunks of atoms whose row buffers stay under this many
 * bytes (default 16 GiB; more than one chunk recomputes the net in the
 * backward).  Process-wide, for engines created afterwards.  No counterpart
 * in the reference (which holds every slot's net output at once). */
int dpmi_sea_set_net_chunk_bytes(long long bytes);

/* hipGraph of one step (no counterpart in the reference; the batched-launch
 * pattern): runs one eager step, then captures the same step on `stream`
 * (must not be the NULL stream).  dpmi_sea_engine_graph_launch replays it:
 * the whole path (wrap, ghosts, cell list, neighbour list, ... , force,
 * virial, energies) on the coordinates currently stored at `coord`, with the
 * boxes given at capture, into the same output buffers.  A new box needs a
 * new capture.  Both precisions. */
int dpmi_sea_engine_graph_capture(dpmi_sea_engine* engine, const double* coord, const double* box_host,
                                  double* energy, double* force, double* virial, double* atom_energy,
                                  void* stream);
int dpmi_sea_engine_graph_launch(dpmi_sea_engine* engine, void* stream);
/* The same capture with the boxes in device memory, box_dev [nframes, 9]:
 * every replay re-reads them (NPT: the caller rescales coordinates and boxes
 * in place between replays).  A box whose cell grid (cells and search
 * layers per dimension) differs from the captured one is not applied: the
 * step runs with the previous box and dpmi_sea_engine_status reports
 * DPMI_ERR_ARGUMENT (capture again).  No counterpart in the reference. */
int dpmi_sea_engine_graph_capture_device_box(dpmi_sea_engine* engine, const double* coord, const double* box_dev,
                                             double* energy, double* force, double* virial, double* atom_energy,
                                             void* stream);

/* Autotuner (no counterpart in the reference): times `reps` steps for each
 * candidate launch shape of the env-mat, embedding and force stages and
 * keeps the fastest per stage; the choice is cached per (precision, size
 * bucket = floor(log2 of the packed atom count)) and applied to engines
 * created later for the same bucket.  dpmi_sea_engine_tuning returns the
 * current knobs (6 ints): {env threads per block, forward embedding waves
 * per workgroup, adjoint waves per workgroup, force rows per workgroup
 * (1 to 16: one row per wavefront; 32, 64, 128 or 256: several rows per
 * wavefront of a 16-wavefront workgroup),
 * matrix-core embedding mask (DPMI_MIXED: bit 0 forward, bit 1 adjoint in
 * table space on the matrix cores, bit 2 adjoint and bit 3 forward from
 * each atom's table-row runs on the matrix cores (precedence over bits 1
 * and 0); else the per-neighbour kernels), fused
 * fitting rows per tile (32 or 64; 0 = chosen from the atom count; plus 256:
 * layer 0's products in their own kernels tiled over rows and columns, the
 * same bits, faster when the row tiles leave compute units idle; or 64 plus
 * 512: 16 wavefronts of one column tile each per 64-row tile)}.
 * dpmi_sea_engine_set_tuning sets them (a knob <= 0, or < 0 for the mask,
 * keeps its value; out-of-range values give DPMI_ERR_ARGUMENT). */
int dpmi_sea_engine_autotune(dpmi_sea_engine* engine, const double* coord, const double* box_host, int reps,
                             void* stream);
int dpmi_sea_engine_tuning(dpmi_sea_engine* engine, int* knobs_host);
int dpmi_sea_engine_set_tuning(dpmi_sea_engine* engine, const int* knobs_host);

/* Synchronises the stream of the last compute and returns the device-side
 * status of the computes (and graph replays) since the previous status call
 * (DPMI_OK, DPMI_ERR_NLIST_CAPACITY, DPMI_ERR_MEMORY, DPMI_ERR_RANGE); the
 * bits are cleared when read, so steps between two calls are not reset one
 * by one. */
int dpmi_sea_engine_status(dpmi_sea_engine* engine);

/* Copies an intermediate of the last compute of frame `frame` to host memory
 * (tests and diagnostics).  what: 0 nall (int, 1), 1 extended coord
 * (double, nall*3), 2 extended type (int, nall), 3 mapping (int, nall),
 * 4 nlist (int, nloc*nnei), 5 em (double, nloc*nnei*4), 6 descriptor
 * (double, nloc*M*A), 7 net_deriv (double, nloc*nnei*4), 8 atom energy
 * (double, nloc), 9 extended capacity of the frame (int, 1), 10 capacity
 * increases made by growth so far (int, 1), 11 neighbour-row candidate
 * capacity (int, 1).  *count_host receives the element count; dst_host may
 * be NULL to query it. */
int dpmi_sea_engine_fetch(dpmi_sea_engine* engine, int frame, int what, void* dst_host, long long* count_host);

/* Device-time breakdown of the last compute in milliseconds per stage,
 * recorded when timing is enabled: stage names are written to names_host as
 * a ';'-separated list.  Returns the number of stages. */
int dpmi_sea_engine_set_timing(dpmi_sea_engine* engine, int enable);
int dpmi_sea_engine_timing(dpmi_sea_engine* engine, double* ms_host, int max_stages, char* names_host,
                           int names_len);

#ifdef __cplusplus
}
#endif

#endif /* DPMI_H */
